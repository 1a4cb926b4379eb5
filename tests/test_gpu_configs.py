"""The precision-form workload (M = 64, d = 8, chi = 2000) and a fixed subset of the batched
workload (random_batch(256, seed=1000), each instance at its own M, d and chi) on the GPU path:
sampled site-tensor entries against the C oracle (bit for bit, the largest hafnians of each site
included), the canonical identities of every site tensor, and sampling properties."""
import numpy as np
import pytest

from conftest import same_bits
from test_gpu_identities import _canonical_report
from test_gpu_mps import _oracle_entries

pytestmark = pytest.mark.gpu


def _check_sampled_entries(inst, kres, knum, kS, sites, n_random=200, seed=0):
    from bosonsupremacy_amd import gaussian, mps, runtime
    rng = np.random.default_rng(seed)
    S_full = gaussian.williamson(inst.sq_cov)[1]
    for k in sites:
        G, L = mps.site_tensor(k, inst.M, inst.sq_cov, kres, knum, kS, inst.d, inst.chi, S_full)
        Gh = runtime.to_host(G)
        nL = int((kres[k] > 1e-10).sum())
        nR = len(knum[k - 1])
        la = knum[k].reshape(len(kres[k]), -1)[:nL].sum(1)
        lb = knum[k - 1].reshape(nR, -1).sum(1)
        picks = []
        while len(picks) < n_random:
            a, b, j = int(rng.integers(nL)), int(rng.integers(nR)), int(rng.integers(inst.d))
            if (la[a] + lb[b] + j) % 2 == 0:
                picks.append((a, b, j))
        big = np.argwhere((la[:, None] + lb[None, :]) == (la.max() + lb.max()))[:2]
        for a, b in big:
            j = inst.d - 1 if (la[a] + lb[b] + inst.d - 1) % 2 == 0 else inst.d - 2
            picks.append((int(a), int(b), j))
        print(f"{inst.name} site {k}: max hafnian size {la.max() + lb.max() + inst.d - 1}", flush=True)
        want = _oracle_entries(k, inst.M, inst.sq_cov, kres, knum, kS, inst.d, inst.chi, picks)
        got = np.array([Gh[b, a, j] for a, b, j in picks])
        assert same_bits(got, want), (inst.name, k)
        del G, Gh


def test_m64_sampled_entries_match_oracle():
    from bosonsupremacy_amd import instances, kron
    inst = instances.workload("m64_d8_chi2000")
    kres, knum, kS = kron.kron_all(inst.sq_cov, inst.d, inst.chi)
    _check_sampled_entries(inst, kres, knum, kS, (1, 9, 32, 57, 62), n_random=150)


def test_m64_canonical_and_sampling_property():
    """Whole-tensor identities of all 64 site tensors, then 10^5 FP64 samples: Rao-Blackwellised
    first-order marginals against the exact lossy-state values (chi = 2000 keeps 0.958 of the
    mid-chain Schmidt weight, so the MPS is a truncation), and the FP16x3 form along the FP64
    sample paths within the 1e-6 bar."""
    import torch
    from bosonsupremacy_amd import batched, instances, metrics, mps, runtime, sampling
    inst = instances.workload("m64_d8_chi2000")
    G, L, kron_out = mps.build_mps(inst.sq_cov, inst.d, inst.chi, pruned=True)
    for k, over, deficit, bound in _canonical_report(G, L, kron_out[0], inst.sq_cov, inst.d):
        assert over <= 1e-12 and -1e-12 <= deficit <= bound * (1 + 1e-9) + 1e-13, k
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    normals, uniforms = sampling.random_inputs(inst.M, inst.n_samples, seed=5)
    smp, probs = sampling.Sampler(G, L, sqrtW, inst.dd).run(normals, uniforms, return_probs=True)
    rb = metrics.rb_first_order(probs)
    mean, _ = metrics.exact_photon_moments(inst.cov)
    rel = np.abs(rb - mean) / mean
    print(f"M=64 RB first order vs exact: max rel {rel.max():.3e}, median {np.median(rel):.3e}", flush=True)
    assert np.all(np.abs(probs.sum(dim=2).cpu().numpy() - 1) < 1e-12)
    assert np.median(rel) < 0.05 and rel.max() < 0.15
    e2 = metrics.rb_second_order(smp, probs)
    mid = batched.mid_uniforms(smp, probs)
    del probs
    s16, p16 = sampling.Sampler(G, L, sqrtW, inst.dd, precision="fp16x3").run(normals, mid, return_probs=True)
    assert int((s16 != smp).any(dim=1).sum()) <= 5
    e1h = metrics.rb_first_order(p16)
    e2h = metrics.rb_second_order(s16, p16)
    big = np.abs(e2) > 1e-3 * np.abs(e2).max()
    r1 = np.max(np.abs(e1h - rb) / rb)
    r2 = np.max(np.abs(e2h[big] - e2[big]) / np.abs(e2[big]))
    print(f"FP16x3 vs FP64 along the FP64 paths: first order {r1:.2e}, second order {r2:.2e}", flush=True)
    assert r1 < 1e-6 and r2 < 1e-6
    torch.cuda.synchronize()


def _batch_subset():
    from bosonsupremacy_amd import instances
    batch = instances.random_batch(256, seed=1000)
    # the largest chi, the largest M, the largest d with a large chi, and a small instance
    pick = [max(range(256), key=lambda i: batch[i].chi), max(range(256), key=lambda i: batch[i].M),
            max(range(256), key=lambda i: (batch[i].d, batch[i].chi)), min(range(256), key=lambda i: batch[i].M)]
    pick = list(dict.fromkeys(pick + list(range(256))))[:4]
    return [batch[i] for i in pick]


@pytest.mark.parametrize("idx", range(4))
def test_batched_subset_entries_and_identities(idx):
    from bosonsupremacy_amd import kron, mps, sampling, runtime
    inst = _batch_subset()[idx]
    print(f"{inst.name}: M={inst.M} d={inst.d} chi={inst.chi}", flush=True)
    kres, knum, kS = kron.kron_all(inst.sq_cov, inst.d, inst.chi)
    _check_sampled_entries(inst, kres, knum, kS, (1, inst.M // 2, inst.M - 2), n_random=100, seed=idx)
    G, L, kron_out = mps.build_mps(inst.sq_cov, inst.d, inst.chi, kron=(kres, knum, kS), pruned=True)
    for k, over, deficit, bound in _canonical_report(G, L, kres, inst.sq_cov, inst.d):
        assert over <= 1e-12 and -1e-12 <= deficit <= bound * (1 + 1e-9) + 1e-13, k
    # a short FP64 pass: probabilities normalised, draws in range
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    normals, uniforms = sampling.random_inputs(inst.M, 2000, seed=inst.seed)
    smp, probs = sampling.Sampler(G, L, sqrtW, inst.dd).run(normals, uniforms, return_probs=True)
    assert float((probs.sum(dim=2) - 1).abs().max()) < 1e-12
    s = runtime.to_host(smp)
    assert s.min() >= 0 and s.max() < inst.dd
