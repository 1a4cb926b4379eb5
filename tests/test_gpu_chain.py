"""Site-major sampling, the construct-and-sample chain, and the drop-in CLIs with their file
contract (kron_cpu.py:132-153, MPS_cpu.py:377-525, sampling_cpu.py:163-183)."""
import os

import numpy as np
import pytest
import torch

from conftest import same_bits
from oracle import gbs_oracle as O

pytestmark = pytest.mark.gpu


def _m8():
    from bosonsupremacy_amd import instances, mps
    inst = instances.workload("m8_d4_chi64")
    G, L, kron_out = mps.build_mps(inst.sq_cov, inst.d, inst.chi)
    return inst, G, L, kron_out


def test_host_tensors_are_refused_by_the_abi_wrapper():
    from bosonsupremacy_amd import runtime
    runtime.context()
    with pytest.raises(ValueError):
        runtime.ptr(torch.zeros(4, dtype=torch.float64))
    with pytest.raises(ValueError):
        runtime.ptr(torch.zeros(4, dtype=torch.float64).pin_memory())


@pytest.mark.parametrize("precision", ["fp64", "fp16x3"])
def test_site_sampler_matches_sampler(precision):
    """Site-major walks in chunks, with the site tensors handed over as host arrays, padded or
    pruned, give the same draws and distributions as the all-resident sampler."""
    from bosonsupremacy_amd import mps, runtime, sampling
    inst, G, L, kron_out = _m8()
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    n = 1000
    normals, uniforms = sampling.random_inputs(inst.M, n, seed=3)
    s, p = sampling.Sampler(G, L, sqrtW, inst.dd, precision=precision).run(normals, uniforms, return_probs=True)
    Gp, Lp, _ = mps.build_mps(inst.sq_cov, inst.d, inst.chi, kron=kron_out, pruned=True)
    for gammas, lambdas, chunk in ((G, L, 300), ([runtime.to_host(g) for g in Gp], Lp, 1000)):
        w = sampling.SiteSampler(inst.M, inst.dd, sqrtW, normals, uniforms, inst.chi, inst.d, precision=precision,
                                 chunk=chunk)
        for k in range(inst.M):
            w.push(k, gammas[k], lambdas[k] if k < inst.M - 1 else None)
        assert torch.equal(w.samples, s)
        assert torch.equal(w.probs, p)


def test_sampler_accepts_host_site_tensors():
    """Pinned host tensors from build_mps(to_host=True) are copied to the device by the sampler."""
    from bosonsupremacy_amd import mps, sampling
    inst, G, L, kron_out = _m8()
    Gh, Lh, _ = mps.build_mps(inst.sq_cov, inst.d, inst.chi, kron=kron_out, to_host=True)
    assert not Gh[3].is_cuda
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    normals, uniforms = sampling.random_inputs(inst.M, 500, seed=4)
    a = sampling.Sampler(G, L, sqrtW, inst.dd).run(normals, uniforms)
    b = sampling.Sampler(Gh, Lh, sqrtW, inst.dd).run(normals, uniforms)
    assert torch.equal(a, b)


@pytest.mark.parametrize("precision", ["fp64", "fp16x3"])
def test_site_major_walks_with_parity_split(precision, monkeypatch):
    """Every site past the first split by photon-number parity (the size threshold forced to 0):
    the site-major walk in chunks, with host tensors, equals the all-resident sampler bit for bit,
    the construct-and-sample chain draws the same samples, and the split sampler's FP64
    probabilities equal the unsplit sampler's to rounding."""
    from bosonsupremacy_amd import chain, runtime, sampling
    inst, G, L, kron_out = _m8()
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    n = 1000
    normals, uniforms = sampling.random_inputs(inst.M, n, seed=5)
    s0, p0 = sampling.Sampler(G, L, sqrtW, inst.dd, precision=precision).run(normals, uniforms, return_probs=True)
    monkeypatch.setattr(sampling, "PARITY_MIN_ENTRIES", 0)
    real, split = sampling.parity_blocks, []

    def counted(g):
        b = real(g)
        split.append(b is not None)
        return b
    monkeypatch.setattr(sampling, "parity_blocks", counted)
    print(f"parity split walks {precision} M={inst.M} n={n}", flush=True)
    s, p = sampling.Sampler(G, L, sqrtW, inst.dd, precision=precision).run(normals, uniforms, return_probs=True)
    assert len(split) == inst.M - 1 and all(split)
    if precision == "fp64":
        assert torch.equal(s, s0)
        assert (p - p0).abs().max().item() < 1e-13
    w = sampling.SiteSampler(inst.M, inst.dd, sqrtW, normals, uniforms, inst.chi, inst.d, precision=precision,
                             chunk=300)
    for k in range(inst.M):
        w.push(k, runtime.to_host(G[k]), L[k] if k < inst.M - 1 else None)
    assert torch.equal(w.samples, s) and torch.equal(w.probs, p)
    assert len(split) == 2 * (inst.M - 1) and all(split)
    if precision == "fp64":
        c = chain.Chain(inst.sq_cov, inst.cov, inst.d, inst.chi, inst.dd, n, seed=5, chunk=400, kron=kron_out)
        c.run()
        assert same_bits(c.samples(0), runtime.to_host(s).astype(np.int64))
        assert all(split)


def test_chain_matches_sampler_and_resumes(tmp_path):
    """Construct-and-sample over the M = 32 workload: the FP64 walk equals the all-resident
    sampler's, the FP16x3 replay stays on its paths within 1e-6, and a chain stopped after site
    13, checkpointed and restored in a new driver ends with the same results."""
    from bosonsupremacy_amd import chain, instances, mps, runtime, sampling
    inst = instances.workload("m32_d6_chi512")
    n = 2000
    forms = [chain.Form("fp64"), chain.Form("fp16x3", "fp16x3")]
    c = chain.Chain(inst.sq_cov, inst.cov, inst.d, inst.chi, inst.dd, n, seed=9, forms=forms, chunk=700)
    c.run(on_site=lambda k, t: print(f"site {k}: {t:.3f} s", flush=True))
    G, L, _ = mps.build_mps(inst.sq_cov, inst.d, inst.chi, kron=c.kron)
    normals, uniforms = sampling.random_inputs(inst.M, n, seed=9)
    s, p = sampling.Sampler(G, L, c.sqrtW, inst.dd).run(normals, uniforms, return_probs=True)
    assert same_bits(c.samples(0), runtime.to_host(s).astype(np.int64))
    assert np.allclose(c.first_order(0), (runtime.to_host(p).mean(axis=1) * np.arange(inst.dd)).sum(axis=1),
                       rtol=1e-13, atol=0)
    err = c.errors()["fp16x3"]
    print(err, flush=True)
    assert err["first_order_max_rel_err"] < 1e-6 and err["second_order_max_rel_err"] < 1e-6
    # two sessions: sites 0..13, checkpoint, then 14..31 in a fresh driver
    c1 = chain.Chain(inst.sq_cov, inst.cov, inst.d, inst.chi, inst.dd, n, seed=9, forms=forms, chunk=700,
                     kron=c.kron)
    c1.run(sites=range(14))
    c1.save(str(tmp_path / "ck.npz"))
    del c1
    c2 = chain.Chain(inst.sq_cov, inst.cov, inst.d, inst.chi, inst.dd, n, seed=9, forms=forms, chunk=700,
                     kron=c.kron)
    c2.restore(str(tmp_path / "ck.npz"))
    c2.run()
    for fi in range(2):
        assert same_bits(c2.samples(fi), c.samples(fi))
        assert same_bits(c2.mean_probs[fi], c.mean_probs[fi]) and same_bits(c2.cond_mean[fi], c.cond_mean[fi])


def _write_instance(root, inst):
    os.makedirs(root, exist_ok=True)
    np.save(os.path.join(root, "cov.npy"), inst.cov)
    np.save(os.path.join(root, "sq_cov.npy"), inst.sq_cov)


@pytest.mark.parametrize("reference_dtypes", [False, True], ids=["fp64_files", "reference_files"])
def test_cli_file_contract_and_resume(tmp_path, capsys, reference_dtypes):
    """The three CLIs through {dir}: spectra, site tensors and samples against the oracle on the
    same inputs; with --reference-dtypes the files are float16 res, int8 num, complex64 Gamma and
    float32 Lambda (kron_cpu.py:127, MPS_cpu.py:397-398) and the later stages read them; a site
    run killed before its files were written resumes at the first missing site."""
    from bosonsupremacy_amd import instances, kron, mps, runtime, sampling
    inst = instances.workload("m8_d4_chi64")
    root = str(tmp_path) + "/"
    _write_instance(root, inst)
    flags = ["--reference-dtypes"] if reference_dtypes else []
    args = ["--d", str(inst.d), "--chi", str(inst.chi), "--dir", root]
    kron.main(args + flags)
    path = root + f"d_{inst.d}_chi_{inst.chi}/"
    M = inst.M
    kres = [np.load(path + f"res_{k}.npy") for k in range(M - 1)]
    knum = [np.load(path + f"num_{k}.npy") for k in range(M - 1)]
    kS = [np.load(path + f"S_{k}.npy") for k in range(M - 1)]
    assert all(r.dtype == (np.float16 if reference_dtypes else np.float64) for r in kres)
    assert all(x.dtype == np.int8 for x in knum)
    okron = O.run_kron(inst.sq_cov, inst.d, inst.chi)
    for k in range(M - 1):
        assert same_bits(kres[k], okron[0][k].astype(kres[k].dtype)) and same_bits(knum[k], okron[1][k]), k
        assert same_bits(kS[k], okron[2][k]), k
    mps.main(args + flags)
    # oracle sites from the same (possibly float16) spectra
    ofeed = ([r.astype(np.float64) for r in kres], knum, kS)
    oG, oL = O.build_mps(inst.sq_cov, inst.d, inst.chi, kron=ofeed)
    gdt, ldt = (np.complex64, np.float32) if reference_dtypes else (np.complex128, np.float64)
    for k in range(M):
        g = np.load(path + f"Gamma_{k}.npy")
        assert g.dtype == gdt and same_bits(g, oG[k].astype(gdt)), k
        if k < M - 1:
            lam = np.load(path + f"Lambda_{k}.npy")
            assert lam.dtype == ldt and same_bits(lam, oL[k].astype(ldt)), k
    # resume: a run that stopped after site 4 (site 5's Gamma never written)
    for k in range(5, M):
        os.remove(path + f"Gamma_{k}.npy")
    os.remove(path + "Lambda_6.npy")
    stamp = os.path.getmtime(path + "Gamma_2.npy")
    capsys.readouterr()
    mps.main(args + flags)
    assert "resuming at site 5" in capsys.readouterr().out
    assert os.path.getmtime(path + "Gamma_2.npy") == stamp     # finished sites are not recomputed
    for k in range(M):
        assert same_bits(np.load(path + f"Gamma_{k}.npy"), oG[k].astype(gdt)), k
    # sampling: batches of n, one Gamma_i at a time from disk
    N, nb = 900, 400
    sampling.main(["--N", str(N), "--n", str(nb), "--iter", "2", "--d", str(inst.d), "--dd", str(inst.dd),
                   "--chi", str(inst.chi), "--dir", root, "--seed", "7"])
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    Gf = [np.load(path + f"Gamma_{k}.npy").astype(np.complex128) for k in range(M)]
    Lam = np.stack([np.load(path + f"Lambda_{k}.npy").astype(np.float64) for k in range(M - 1)], axis=1)
    for it in range(2):
        got = np.load(root + f"samples_{it}.npy")
        assert got.dtype == np.int64 and got.shape == (N, M)
        want = []
        for b0 in range(0, N, nb):
            nrm, uni = sampling.random_inputs(M, min(N, b0 + nb) - b0, 7 + it, b0)
            want.append(O.sampling(Gf, Lam, sqrtW, inst.dd, runtime.to_host(nrm), runtime.to_host(uni))[0])
        want = np.concatenate(want)
        # equal draws except where rounding puts a uniform on a cdf boundary
        assert np.mean(np.all(got == want, axis=1)) > 0.998, it
