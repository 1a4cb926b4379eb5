"""GPU Schmidt spectra and site tensors vs the oracle (bit-exact)."""
import numpy as np
import pytest
from scipy.special import factorial

from conftest import same_bits
from oracle import c_oracle, gbs_oracle as O

pytestmark = pytest.mark.gpu


def test_kron_m8_matches_reference(golden_m8):
    from bosonsupremacy_amd import kron
    g = golden_m8
    res, num, S = kron.kron_all(g["sq_cov"], int(g["d"]), int(g["chi"]))
    for k in range(7):
        assert same_bits(num[k], g[f"kron_num_{k}"]), k
        assert same_bits(res[k].astype(np.float16), g[f"kron_res16_{k}"]), k
        tp, _ = kron.cut_spectrum_inputs(g["sq_cov"], k + 1, int(g["d"]))
        ores, onum = O.kron_from_thermal(list(tp), int(g["chi"]))
        assert same_bits(res[k], ores) and same_bits(num[k], onum), k


@pytest.mark.parametrize("max_dim", [10 ** 5, 300])
def test_kron_m32_matches_oracle(max_dim):
    from bosonsupremacy_amd import instances, kron
    inst = instances.workload("m32_d6_chi512")
    tps = [kron.cut_spectrum_inputs(inst.sq_cov, k + 1, inst.d)[0] for k in range(inst.M - 1)]
    got = kron.kron_cuts(tps, inst.chi, max_dim=max_dim)
    for k, tp in enumerate(tps):
        ores, onum = O.kron_from_thermal(list(tp), inst.chi, max_dim=max_dim)
        print(f"cut {k}: {len(ores)} weights", flush=True)
        assert same_bits(got[k][0], ores), k
        assert same_bits(got[k][1], onum), k


def test_kron_edge_cases():
    from bosonsupremacy_amd import kron
    # one-mode cut (no loop), a zero-temperature mode (exact zeros), d = 1
    tp1 = np.array([[0.7, 0.21, 0.063, 0.0189]])
    tp0 = np.stack([O.thermal_photons(0.0, 3), O.thermal_photons(0.2, 3), O.thermal_photons(0.1, 3)])
    tpd1 = np.stack([O.thermal_photons(0.3, 1)] * 4)
    for tp, chi in ((tp1, 3), (tp1, 10), (tp0, 5), (tpd1, 4)):
        (r, n), = kron.kron_cuts([tp], chi)
        orr, onn = O.kron_from_thermal(list(tp), chi)
        assert same_bits(r, orr) and same_bits(n, onn)


def test_sites_m8_match_oracle():
    from bosonsupremacy_amd import instances, mps, runtime
    inst = instances.workload("m8_d4_chi64")
    G, L, kron_out = mps.build_mps(inst.sq_cov, inst.d, inst.chi)
    oG, oL = O.build_mps(inst.sq_cov, inst.d, inst.chi, kron=(kron_out[0], kron_out[1], kron_out[2]))
    for k in range(inst.M):
        print(f"site {k}", flush=True)
        assert same_bits(runtime.to_host(G[k]), oG[k]), k
        assert same_bits(L[k], oL[k]), k


def _oracle_entries(k, M, sq_cov, kres, knum, kS, d, chi, picks, err_tol=1e-10):
    """Gamma[b, a, j] for picked (a, b, j) via the C hafnian and NumPy's division order
    (MPS_cpu.py:478-519)."""
    Sigma, Z = O.site_sigma(k, M, sq_cov, kS)
    Mp = Sigma.shape[0] // 2
    res = kres[k]
    num = knum[k].reshape(len(res), -1)[res > err_tol]
    lam = np.sqrt(res[res > err_tol])
    num_pre = knum[k - 1].reshape(len(knum[k - 1]), -1)
    lt, rt = O.get_target(num), O.get_target(num_pre)
    lden = np.sqrt(np.prod(np.array(factorial(num)), axis=1))
    rden = np.sqrt(np.prod(np.array(factorial(num_pre)), axis=1))
    out = []
    for a, b, j in picks:
        ta = lt[a][lt[a] != 0]
        tb = rt[b][rt[b] != 0] + (Mp - 1)
        tg = np.concatenate([np.zeros(j, np.int32), ta, tb]).astype(np.int32)
        if len(tg) == 0:
            out.append(np.ones(1) / Z / lam[a])
            continue
        h = c_oracle.hafnian_targets(Sigma, tg[None])
        den = np.sqrt(factorial(j)) * lden[a:a + 1] * rden[b:b + 1]
        out.append((h / den) / Z / lam[a:a + 1])
    return np.concatenate(out)


def test_sites_m32_sampled_entries_match_oracle():
    from bosonsupremacy_amd import instances, kron, mps, runtime, gaussian
    inst = instances.workload("m32_d6_chi512")
    kres, knum, kS = kron.kron_all(inst.sq_cov, inst.d, inst.chi)
    rng = np.random.default_rng(0)
    S_full = gaussian.williamson(inst.sq_cov)[1]
    for k in (1, 8, 16, 29):
        G, L = mps.site_tensor(k, inst.M, inst.sq_cov, kres, knum, kS, inst.d, inst.chi, S_full)
        Gh = runtime.to_host(G)
        nL = int((kres[k] > 1e-10).sum())
        nR = len(knum[k - 1])
        la = knum[k].reshape(len(kres[k]), -1)[:nL].sum(1)
        lb = knum[k - 1].reshape(nR, -1).sum(1)
        picks = []
        while len(picks) < 300:
            a, b, j = int(rng.integers(nL)), int(rng.integers(nR)), int(rng.integers(inst.d))
            if (la[a] + lb[b] + j) % 2 == 0:
                picks.append((a, b, j))
        # always include the largest hafnians of the site
        big = np.argwhere((la[:, None] + lb[None, :]) == (la.max() + lb.max()))[:3]
        for a, b in big:
            j = inst.d - 1 if (la[a] + lb[b] + inst.d - 1) % 2 == 0 else inst.d - 2
            picks.append((int(a), int(b), j))
        print(f"site {k}: max hafnian size {la.max() + lb.max() + inst.d - 1}", flush=True)
        want = _oracle_entries(k, inst.M, inst.sq_cov, kres, knum, kS, inst.d, inst.chi, picks)
        got = np.array([Gh[b, a, j] for a, b, j in picks])
        assert same_bits(got, want), k
        # odd-size entries stay exactly zero
        odd = (la[:, None] + lb[None, :]) % 2 == 1
        assert np.all(Gh[:nR, :nL, 0].T[odd] == 0)


def test_site_in_right_label_runs_is_bit_identical():
    """A site computed in several calls over runs of its right labels equals the one-call site."""
    from bosonsupremacy_amd import instances, kron, mps, runtime, gaussian
    inst = instances.workload("m32_d6_chi512")
    kres, knum, kS = kron.kron_all(inst.sq_cov, inst.d, inst.chi)
    S_full = gaussian.williamson(inst.sq_cov)[1]
    for k in (0, 9, 31):
        G1, L1 = mps.site_tensor(k, inst.M, inst.sq_cov, kres, knum, kS, inst.d, inst.chi, S_full)
        seen = []
        G5, L5 = mps.site_tensor(k, inst.M, inst.sq_cov, kres, knum, kS, inst.d, inst.chi, S_full,
                                 right_runs=5, on_run=lambda i, n: seen.append((i, n)))
        print(f"site {k}: {len(seen)} runs", flush=True)
        assert same_bits(runtime.to_host(G5), runtime.to_host(G1)) and same_bits(L5, L1), k
        # at most the runs asked for (labels of equal work are not split), one call per run
        assert 1 <= len(seen) <= 5 and [i for i, _ in seen] == list(range(seen[-1][1])), k
        assert k != 9 or len(seen) == 5
