"""GPU sampling path vs the oracle: displacements, contraction, probabilities, draws, RNG."""
import numpy as np
import pytest
import torch
from scipy.linalg import expm

from conftest import same_bits
from oracle import gbs_oracle as O, philox

pytestmark = pytest.mark.gpu


def test_batch_displaces_matches_expm():
    from bosonsupremacy_amd import sampling
    rng = np.random.default_rng(0)
    for dd in (4, 10, 20):
        al = (rng.normal(size=(40, 3)) + 1j * rng.normal(size=(40, 3))) * 0.8
        got = sampling.batch_displaces(dd, al)
        want = O.batch_displaces(dd, al)
        assert np.abs(got - want).max() < 1e-12, dd


@pytest.fixture(params=[0, 1], ids=["blocks", "gauss"])
def contract_form(request):
    """Both site-contraction forms (real 2x2 blocks, Gauss three products) for FP64 and FP32."""
    from bosonsupremacy_amd import runtime
    ctx = runtime.context()
    for prec in (0, 1):
        ctx.call("bsmps_set_contract_form", prec, request.param)
    yield request.param
    for prec in (0, 1):
        ctx.call("bsmps_set_contract_form", prec, 0)


def test_site_contract_matches_matmul(contract_form):
    from bosonsupremacy_amd import runtime
    ctx = runtime.context()
    rng = np.random.default_rng(1)
    for n, chi_p, chi, d in ((70, 33, 17, 3), (256, 64, 64, 4), (5, 100, 7, 6), (300, 130, 70, 5)):
        pre = rng.normal(size=(n, chi_p)) + 1j * rng.normal(size=(n, chi_p))
        lam = rng.uniform(0.1, 1, size=chi_p)
        G = rng.normal(size=(chi_p, chi, d)) + 1j * rng.normal(size=(chi_p, chi, d))
        want = ((pre * lam) @ G.reshape(chi_p, chi * d)).reshape(n, chi, d)
        pre_d, lam_d, G_d = runtime.to_device(pre), runtime.to_device(lam), runtime.to_device(G)
        T = torch.empty((n, chi, d), dtype=torch.complex128, device=ctx.device)
        ctx.call("bsmps_site_contract", n, chi_p, chi, d, runtime.ptr(pre_d), runtime.ptr(lam_d),
                 runtime.ptr(G_d), runtime.ptr(T))
        got = runtime.to_host(T)
        assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max(), (n, chi_p, chi, d)


def test_philox_matches_host():
    from bosonsupremacy_amd import sampling, runtime
    for seed, stream0, first, n in ((0, 0, 0, 1000), (12345678901, 7, 10 ** 9, 257)):
        u = runtime.to_host(sampling.philox(3, n, seed, stream0, first))
        z = runtime.to_host(sampling.philox(2, n, seed, stream0, first, normal=True))
        for r in range(3):
            assert same_bits(u[r], philox.stream(seed, stream0 + r, first, n))
        for r in range(2):
            assert np.abs(z[r] - philox.stream(seed, stream0 + r, first, n, normal=True)).max() < 1e-13


def _mps_m8():
    from bosonsupremacy_amd import instances, mps
    inst = instances.workload("m8_d4_chi64")
    G, L, kron_out = mps.build_mps(inst.sq_cov, inst.d, inst.chi)
    return inst, G, L


@pytest.mark.parametrize("parity", ["full", "split"])
def test_sampling_m8_matches_oracle(contract_form, parity, monkeypatch):
    """GPU sampler vs the oracle on the M = 8 workload, with every site contracted in one product
    and with every site that splits by photon-number parity contracted as its two blocks."""
    from bosonsupremacy_amd import runtime, sampling
    monkeypatch.setattr(sampling, "PARITY_MIN_ENTRIES", 0 if parity == "split" else 1 << 62)
    inst, G, L = _mps_m8()
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    n = 3000
    rng = np.random.default_rng(3)
    normals = rng.normal(size=(2 * inst.M, n))
    uniforms = rng.uniform(size=(inst.M, n))
    s = sampling.Sampler(G, L, sqrtW, inst.dd)
    assert any(f.blocks is not None for f in s.forms) == (parity == "split")
    got, probs = s.run(runtime.to_device(normals), runtime.to_device(uniforms), return_probs=True)
    got, probs = runtime.to_host(got), runtime.to_host(probs)
    oG = [runtime.to_host(g) for g in G]
    want, oprobs = O.sampling(oG, np.stack(L[:-1], axis=1), sqrtW, inst.dd, normals, uniforms)
    oprobs = np.transpose(oprobs, (1, 0, 2))
    # conditional distributions agree to FP64 rounding while the draws agree
    first_diff = np.argmax(np.any(got != want, axis=0)) if np.any(got != want) else inst.M
    agree = np.all(got == want, axis=1)
    assert agree.mean() > 0.999
    for i in range(inst.M):
        ok = np.all(got[:, :i] == want[:, :i], axis=1)
        assert np.abs(probs[i][ok] - oprobs[i][ok]).max() < 1e-11, i
    # a disagreeing draw must sit on a cumulative-probability boundary
    for s_ in np.where(~agree)[0]:
        i = int(np.argmax(got[s_] != want[s_]))
        cdf = np.cumsum(oprobs[i][s_])
        assert np.min(np.abs(cdf - uniforms[i, s_])) < 1e-9
    del first_diff


def test_sampling_m32_first_order_marginals():
    """Full-size property: 10^5 samples of the M = 32 workload; the Rao-Blackwellised
    first-order marginals match the exact lossy-state values (truncation + noise)."""
    from bosonsupremacy_amd import instances, mps, runtime, sampling
    inst = instances.workload("m32_d6_chi512")
    G, L, _ = mps.build_mps(inst.sq_cov, inst.d, inst.chi)
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    s = sampling.Sampler(G, L, sqrtW, inst.dd)
    normals, uniforms = sampling.random_inputs(inst.M, inst.n_samples, seed=5)
    samples, probs = s.run(normals, uniforms, return_probs=True)
    probs = runtime.to_host(probs)
    rb = (probs * np.arange(inst.dd)[None, None, :]).sum(axis=2).mean(axis=1)
    mean, _ = O.photon_number_moments(inst.cov)
    print("max rel err of E[n_i]:", np.max(np.abs(rb - mean) / mean), flush=True)
    assert np.all(np.abs(rb - mean) <= 0.05 * mean + 0.01)
    assert np.all(np.abs(probs.sum(axis=2) - 1) < 1e-12)
    smp = runtime.to_host(samples)
    assert smp.min() >= 0 and smp.max() < inst.dd


def test_site_contract_c64_matches_matmul(contract_form):
    from bosonsupremacy_amd import runtime
    ctx = runtime.context()
    rng = np.random.default_rng(7)
    for n, chi_p, chi, d in ((70, 33, 17, 3), (256, 64, 64, 4), (300, 130, 70, 5)):
        pre = (rng.normal(size=(n, chi_p)) + 1j * rng.normal(size=(n, chi_p))).astype(np.complex64)
        lam = rng.uniform(0.1, 1, size=chi_p)
        G = (rng.normal(size=(chi_p, chi, d)) + 1j * rng.normal(size=(chi_p, chi, d))).astype(np.complex64)
        want = ((pre.astype(np.complex128) * lam) @ G.astype(np.complex128).reshape(chi_p, chi * d)).reshape(n, chi, d)
        pre_d, lam_d, G_d = runtime.to_device(pre), runtime.to_device(lam), runtime.to_device(G)
        T = torch.empty((n, chi, d), dtype=torch.complex64, device=ctx.device)
        ctx.call("bsmps_site_contract_c64", n, chi_p, chi, d, runtime.ptr(pre_d), runtime.ptr(lam_d),
                 runtime.ptr(G_d), runtime.ptr(T))
        got = runtime.to_host(T).astype(np.complex128)
        assert np.abs(got - want).max() <= 2e-5 * np.abs(want).max(), (n, chi_p, chi, d)


def test_sampling_fp32_form_m8(contract_form):
    """FP32 contraction form: conditional distributions within 1e-5 of the FP64 form on the
    same random numbers, and the same draws away from cdf boundaries."""
    from bosonsupremacy_amd import runtime, sampling
    inst, G, L = _mps_m8()
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    n = 3000
    rng = np.random.default_rng(4)
    normals = runtime.to_device(rng.normal(size=(2 * inst.M, n)))
    uniforms = runtime.to_device(rng.uniform(size=(inst.M, n)))
    s64 = sampling.Sampler(G, L, sqrtW, inst.dd)
    s32 = sampling.Sampler(G, L, sqrtW, inst.dd, precision="fp32")
    a, pa = s64.run(normals, uniforms, return_probs=True)
    b, pb = s32.run(normals, uniforms, return_probs=True)
    a, b, pa, pb = (runtime.to_host(x) for x in (a, b, pa, pb))
    same = np.all(a == b, axis=1)
    assert same.mean() > 0.99
    assert np.abs(pa[:, same] - pb[:, same]).max() < 1e-5


def test_site_contract_bf16x3_matches_matmul():
    from bosonsupremacy_amd import runtime
    ctx = runtime.context()
    rng = np.random.default_rng(8)
    for n, chi_p, chi, d in ((70, 33, 17, 3), (300, 64, 64, 4)):
        pre = (rng.normal(size=(n, chi_p)) + 1j * rng.normal(size=(n, chi_p))).astype(np.complex64)
        lam = rng.uniform(0.1, 1, size=chi_p)
        G = (rng.normal(size=(chi_p, chi, d)) + 1j * rng.normal(size=(chi_p, chi, d))).astype(np.complex64)
        want = ((pre.astype(np.complex128) * lam) @ G.astype(np.complex128).reshape(chi_p, chi * d)).reshape(n, chi, d)
        pre_d, lam_d, G_d = runtime.to_device(pre), runtime.to_device(lam), runtime.to_device(G)
        hi = torch.empty((2 * G.size,), dtype=torch.bfloat16, device=ctx.device)
        lo = torch.empty((2 * G.size,), dtype=torch.bfloat16, device=ctx.device)
        ctx.call("bsmps_split_bf16", runtime.ptr(G_d), 2 * G.size, runtime.ptr(hi), runtime.ptr(lo))
        T = torch.empty((n, chi, d), dtype=torch.complex64, device=ctx.device)
        ctx.call("bsmps_site_contract_bf16x3", n, chi_p, chi, d, runtime.ptr(pre_d), runtime.ptr(lam_d),
                 runtime.ptr(hi), runtime.ptr(lo), runtime.ptr(T))
        got = runtime.to_host(T).astype(np.complex128)
        err = np.abs(got - want).max() / np.abs(want).max()
        print("bf16x3 rel err", err, flush=True)
        assert err <= 2e-4, (n, chi_p, chi, d)


def test_sampling_bf16x3_form_m8():
    from bosonsupremacy_amd import runtime, sampling
    inst, G, L = _mps_m8()
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    n = 2000
    rng = np.random.default_rng(5)
    normals = runtime.to_device(rng.normal(size=(2 * inst.M, n)))
    uniforms = runtime.to_device(rng.uniform(size=(inst.M, n)))
    a, pa = sampling.Sampler(G, L, sqrtW, inst.dd).run(normals, uniforms, return_probs=True)
    b, pb = sampling.Sampler(G, L, sqrtW, inst.dd, precision="bf16x3").run(normals, uniforms, return_probs=True)
    a, b, pa, pb = (runtime.to_host(x) for x in (a, b, pa, pb))
    same = np.all(a == b, axis=1)
    assert same.mean() > 0.98
    assert np.abs(pa[:, same] - pb[:, same]).max() < 1e-3


def _bf16x6_contract(ctx, pre, lam, G):
    from bosonsupremacy_amd import runtime
    n, chi_p = pre.shape
    _, chi, d = G.shape
    pre_d, lam_d, G_d = runtime.to_device(pre), runtime.to_device(lam), runtime.to_device(G)
    kpad, cpad = -(-chi_p // 32) * 32, -(-(chi * d) // 64) * 64
    nbytes = ctx.lib.bsmps_bf16x6_plane_bytes(chi_p, chi, d)
    assert nbytes == 6 * 2 * kpad * cpad
    rbytes = ctx.lib.bsmps_bf16x6_row_plane_bytes(n, chi_p)
    assert rbytes == 6 * 2 * n * kpad
    planes = torch.empty((nbytes // 2,), dtype=torch.bfloat16, device=ctx.device)
    rows = torch.empty((rbytes // 2,), dtype=torch.bfloat16, device=ctx.device)
    ctx.call("bsmps_split_bf16x6", chi_p, chi, d, runtime.ptr(G_d), runtime.ptr(lam_d), runtime.ptr(planes))
    ctx.call("bsmps_split_rows_bf16x6", n, chi_p, runtime.ptr(pre_d), runtime.ptr(rows))
    # the three parts of every value add up to it exactly
    parts = rows.view(n, kpad // 32, 6, 32).permute(2, 0, 1, 3).reshape(6, n, kpad).float()
    re = (parts[0].double() + parts[1].double() + parts[2].double())[:, :chi_p]
    im = (parts[3].double() + parts[4].double() + parts[5].double())[:, :chi_p]
    assert torch.equal(re.float(), pre_d.real) and torch.equal(im.float(), pre_d.imag)
    T = torch.empty((n, chi, d), dtype=torch.complex64, device=ctx.device)
    ctx.call("bsmps_site_contract_bf16x6", n, chi_p, chi, d, runtime.ptr(rows), runtime.ptr(planes), runtime.ptr(T))
    return runtime.to_host(T).astype(np.complex128)


def test_site_contract_bf16x6_matches_matmul():
    """FP32-accurate BF16x6 form: every entry within the FP32 GEMM error bound
    K * 2^-24 * sum |a| |b| of the exact product, on ragged shapes (rows, k and columns not
    multiples of the tile) and on operands spanning twelve decades."""
    from bosonsupremacy_amd import runtime
    ctx = runtime.context()
    rng = np.random.default_rng(9)
    # (200, 2100, ...) has k >= 2048 (128-column tiles) and a last tile past the padded columns;
    # (96, 9000, ...) has k > 4096, so it runs in k segments added in FP64; the two-stage k loop
    # sees one step (k = 20), an odd count (k = 90: 3, k = 130: 5) and even counts
    for n, chi_p, chi, d, spread in ((70, 33, 17, 3, 0), (300, 64, 64, 4, 0), (257, 130, 70, 5, 0),
                                     (129, 100, 30, 6, 6), (200, 2100, 50, 3, 2), (96, 9000, 20, 3, 0),
                                     (50, 20, 10, 3, 0), (64, 90, 16, 2, 3)):
        scale_a = 10.0 ** rng.uniform(-spread, spread, size=(n, chi_p))
        scale_g = 10.0 ** rng.uniform(-spread, spread, size=(chi_p, chi, d))
        pre = ((rng.normal(size=(n, chi_p)) + 1j * rng.normal(size=(n, chi_p))) * scale_a).astype(np.complex64)
        lam = rng.uniform(0.1, 1, size=chi_p)
        G = ((rng.normal(size=(chi_p, chi, d)) + 1j * rng.normal(size=(chi_p, chi, d))) * scale_g).astype(np.complex64)
        A = pre.astype(np.complex128) * lam
        B = G.astype(np.complex128).reshape(chi_p, chi * d)
        want = (A @ B).reshape(n, chi, d)
        bound = ((np.abs(A.real) + np.abs(A.imag)) @ (np.abs(B.real) + np.abs(B.imag))).reshape(n, chi, d)
        got = _bf16x6_contract(ctx, pre, lam, G)
        ratio = (np.abs(got - want) / bound).max()
        print(f"bf16x6 {n}x{chi_p}x{chi * d} spread {spread}: max err / bound = {ratio:.2e}", flush=True)
        assert ratio <= 2 * chi_p * 2.0 ** -24, (n, chi_p, chi, d)


def test_sampling_bf16x6_form_m8():
    """BF16x6 form: conditional distributions within 1e-5 of the FP64 form on the same random
    numbers (the FP32 form's bar), the same draws away from cdf boundaries."""
    from bosonsupremacy_amd import runtime, sampling
    inst, G, L = _mps_m8()
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    n = 3000
    rng = np.random.default_rng(6)
    normals = runtime.to_device(rng.normal(size=(2 * inst.M, n)))
    uniforms = runtime.to_device(rng.uniform(size=(inst.M, n)))
    a, pa = sampling.Sampler(G, L, sqrtW, inst.dd).run(normals, uniforms, return_probs=True)
    b, pb = sampling.Sampler(G, L, sqrtW, inst.dd, precision="bf16x6").run(normals, uniforms, return_probs=True)
    a, b, pa, pb = (runtime.to_host(x) for x in (a, b, pa, pb))
    same = np.all(a == b, axis=1)
    assert same.mean() > 0.99
    assert np.abs(pa[:, same] - pb[:, same]).max() < 1e-5


def _fp16x3_contract(ctx, pre, lam, G):
    from bosonsupremacy_amd import runtime
    n, chi_p = pre.shape
    _, chi, d = G.shape
    pre_d, lam_d, G_d = runtime.to_device(pre), runtime.to_device(lam), runtime.to_device(G)
    kpad, cpad = -(-chi_p // 32) * 32, -(-(chi * d) // 64) * 64
    assert ctx.lib.bsmps_fp16x3_plane_bytes(chi_p, chi, d) == 4 * 2 * kpad * cpad
    assert ctx.lib.bsmps_fp16x3_row_plane_bytes(n, chi_p) == 4 * 2 * n * kpad
    planes = torch.empty((4 * kpad * cpad,), dtype=torch.float16, device=ctx.device)
    colscale = torch.empty((cpad,), dtype=torch.float32, device=ctx.device)
    rows = torch.empty((4 * n * kpad,), dtype=torch.float16, device=ctx.device)
    rowscale = torch.empty((n,), dtype=torch.float32, device=ctx.device)
    ctx.call("bsmps_split_fp16x3", chi_p, chi, d, runtime.ptr(G_d), runtime.ptr(lam_d), runtime.ptr(planes),
             runtime.ptr(colscale))
    ctx.call("bsmps_split_rows_fp16x3", n, chi_p, runtime.ptr(pre_d), runtime.ptr(rows), runtime.ptr(rowscale))
    # the scales are the largest components, and the two parts of every scaled value add up to it
    # within fp16's 22-bit two-part precision
    want_rs = np.maximum(np.abs(pre.real), np.abs(pre.imag)).max(axis=1)
    assert np.allclose(runtime.to_host(rowscale), want_rs, rtol=0, atol=0)
    parts = rows.view(n, kpad // 32, 4, 32).permute(2, 0, 1, 3).reshape(4, n, kpad).double()
    re = (parts[0] + parts[1])[:, :chi_p] * rowscale.double()[:, None]
    assert torch.abs(re - pre_d.real.double()).max() <= 2.0 ** -21 * float(want_rs.max())
    T = torch.empty((n, chi, d), dtype=torch.complex64, device=ctx.device)
    ctx.call("bsmps_site_contract_fp16x3", n, chi_p, chi, d, runtime.ptr(rows), runtime.ptr(rowscale),
             runtime.ptr(planes), runtime.ptr(colscale), runtime.ptr(T))
    return runtime.to_host(T).astype(np.complex128)


def _parity_gamma(rng, rows, cols, d, parity_a, parity_b):
    """Random complex Gamma [rows][cols][d] that is zero wherever p(a) != p(b) xor (j mod 2)."""
    g = rng.normal(size=(rows, cols, d)) + 1j * rng.normal(size=(rows, cols, d))
    j = np.arange(d)
    keep = (parity_a[:, None, None] ^ parity_b[None, :, None] ^ (j[None, None, :] % 2)) == 0
    return np.where(keep, g, 0)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x6", "fp16x3"])
def test_parity_split_contraction_matches_full(prec, monkeypatch):
    """Site tensors whose entries vanish by photon-number parity are contracted as two half-size
    products (gathered labels, columns stored through the block's column map); the result equals
    the full product (FP64 within 1e-13 of its scale, the others within their FP32-level bound)
    and, bit for bit, the same blocks written to a buffer and scattered; tensors without the
    structure are not split, and a tensor of two disconnected label groups still is."""
    from bosonsupremacy_amd import runtime, sampling
    monkeypatch.setattr(sampling, "PARITY_MIN_ENTRIES", 0)   # a small tensor, split regardless
    ctx = runtime.context()
    rng = np.random.default_rng(31)
    rows, cols, d, n = 150, 90, 5, 300
    pa, pb = rng.integers(0, 2, size=rows), rng.integers(0, 2, size=cols)
    G = _parity_gamma(rng, rows, cols, d, pa, pb)
    G[7] = 0                                               # a label with no non-zero entry
    lam = rng.uniform(0.1, 1, size=rows)
    pre = rng.normal(size=(n, rows)) + 1j * rng.normal(size=(n, rows))
    want = ((pre * lam) @ G.reshape(rows, cols * d))
    ctype = torch.complex128 if prec == "fp64" else torch.complex64
    print(f"parity split {prec} rows={rows} cols={cols} d={d} n={n}", flush=True)
    lam_d = runtime.to_device(lam)
    f = sampling._prepare_site(ctx, prec, 1, torch.from_numpy(G), rows, cols, lam_d)
    assert f.blocks is not None and f.gamma is None
    (ia0, ic0), (ia1, ic1) = [(b.ia.cpu().numpy(), b.ic.cpu().numpy()) for b in f.blocks]
    assert sorted(np.concatenate([ia0, ia1]).tolist()) == list(range(rows))
    assert sorted(np.concatenate([ic0, ic1]).tolist()) == list(range(cols * d))
    work = sampling._Work(ctx, prec, n, rows, cols * d, ctype)
    pre_d = torch.from_numpy(pre).to(device=ctx.device, dtype=ctype).reshape(-1).contiguous()
    T = torch.empty((n * cols * d,), dtype=ctype, device=ctx.device)
    sampling.contract_site(ctx, prec, f, n, pre_d, lam_d, T, work)
    T2 = torch.full((n * cols * d,), 7.0, dtype=ctype, device=ctx.device)
    sampling.contract_site(ctx, prec, f, n, pre_d, lam_d, T2, work, scatter=True)
    assert torch.equal(T, T2)
    got = runtime.to_host(T).reshape(n, cols * d).astype(np.complex128)
    if prec == "fp64":
        assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()
    else:
        bound = (np.abs(pre * lam) @ np.abs(G.reshape(rows, cols * d)))
        assert (np.abs(got - want) / bound).max() < 1e-5
    # one entry against the structure: no split
    G2 = G.copy()
    a_bad = int(np.flatnonzero(pa == 0)[0]); b_bad = int(np.flatnonzero(pb == 0)[0])
    G2[a_bad, b_bad, 1] = 1.0
    assert sampling.parity_blocks(torch.from_numpy(G2).to(ctx.device)) is None
    # two disconnected label groups (block-diagonal in the labels): still split, checked entry-wise
    G3 = np.zeros_like(G)
    G3[:75, :45] = G[:75, :45]
    G3[75:, 45:] = G[75:, 45:]
    blocks = sampling.parity_blocks(torch.from_numpy(G3).to(ctx.device))
    assert blocks is not None
    flat = G3.reshape(rows, cols * d)
    (ia0, ic0), (ia1, ic1) = [(a.cpu().numpy(), c.cpu().numpy()) for a, c in blocks]
    assert not flat[np.ix_(ia0, ic1)].any() and not flat[np.ix_(ia1, ic0)].any()


@pytest.mark.parametrize("prec,chi_p", [("fp64", 300), ("fp32", 300), ("bf16x6", 300), ("fp16x3", 300),
                                         ("bf16x6", 4500), ("fp16x3", 4500)])
def test_column_mapped_contractions_match_plain(prec, chi_p, contract_form):
    """The _cols entry points store output column c at T[s * ldT + cmap[c]]: every mapped column
    equals the plain call's column bit for bit, entries of cmap outside [0, ldT) are not written,
    and T's unmapped columns keep what they held -- in both FP64 / FP32 forms and, at chi_prev =
    4500, through the reduced forms' k segments and their combine."""
    from bosonsupremacy_amd import runtime, sampling
    ctx = runtime.context()
    rng = np.random.default_rng(chi_p + len(prec))
    n, chi, d, ld = 200, 37, 3, 150
    ctype = torch.complex128 if prec == "fp64" else torch.complex64
    print(f"column map {prec} chi_prev={chi_p} n={n} cols={chi * d} ld={ld} form={contract_form}", flush=True)
    G = rng.normal(size=(chi_p, chi, d)) + 1j * rng.normal(size=(chi_p, chi, d))
    lam = rng.uniform(0.1, 1, size=chi_p)
    pre = rng.normal(size=(n, chi_p)) + 1j * rng.normal(size=(n, chi_p))
    G_d = torch.from_numpy(G).to(device=ctx.device, dtype=ctype).contiguous()
    pre_d = torch.from_numpy(pre).to(device=ctx.device, dtype=ctype).reshape(-1).contiguous()
    lam_d = runtime.to_device(lam)
    split = sampling._split_operand(ctx, prec, G_d, chi_p, chi, d, lam_d)
    work = sampling._Work(ctx, prec, n, chi_p, chi * d, ctype)
    plain = torch.empty((n * chi * d,), dtype=ctype, device=ctx.device)
    sampling._contract(ctx, prec, n, chi_p, chi, d, runtime.ptr(pre_d), lam_d, G_d, split, plain, work)
    cmap = rng.permutation(ld)[:chi * d].astype(np.int32)
    cmap[5], cmap[40] = -1, ld                      # not written
    Tm = torch.full((n * ld,), 3.0 - 1.0j, dtype=ctype, device=ctx.device)
    sampling._contract(ctx, prec, n, chi_p, chi, d, runtime.ptr(pre_d), lam_d, G_d, split, Tm, work,
                       out_cols=(ld, torch.from_numpy(cmap).to(ctx.device)))
    got, want = runtime.to_host(Tm).reshape(n, ld), runtime.to_host(plain).reshape(n, chi * d)
    ok = (cmap >= 0) & (cmap < ld)
    assert np.array_equal(np.ascontiguousarray(got[:, cmap[ok]]).view(np.uint8),
                          np.ascontiguousarray(want[:, ok]).view(np.uint8))
    untouched = np.setdiff1d(np.arange(ld), cmap[ok])
    assert len(untouched) > 0 and (got[:, untouched] == 3.0 - 1.0j).all()


def test_column_mapped_contractions_refuse_bad_maps():
    """A non-positive row length ldT or a missing column map is refused before any launch, with
    the entry point named; T is left untouched."""
    from bosonsupremacy_amd import _lib, runtime
    ctx = runtime.context()
    T = torch.full((8,), 5.0, dtype=torch.float64, device=ctx.device)
    cmap = torch.zeros((4,), dtype=torch.int32, device=ctx.device)
    x = runtime.ptr(T)
    for ld, cm in ((0, runtime.ptr(cmap)), (-3, runtime.ptr(cmap)), (8, None)):
        st = ctx.lib.bsmps_site_contract_cols(ctx.handle, 2, 2, 2, 2, x, x, x, x, ld, cm)
        assert st == _lib.BSMPS_ERR_INVALID and b"bsmps_site_contract_cols" in ctx.lib.bsmps_last_error()
        st = ctx.lib.bsmps_site_contract_fp16x3_cols(ctx.handle, 2, 2, 2, 2, x, x, x, x, x, ld, cm)
        assert st == _lib.BSMPS_ERR_INVALID and b"fp16x3_cols" in ctx.lib.bsmps_last_error()
    torch.cuda.synchronize()
    assert (T == 5.0).all()


def test_split_gamma_planes_and_column_scales_exact():
    """The Gamma planes of both split forms, entry by entry, against a NumPy restatement of the
    split, on a ragged shape (k and columns padded) with all-zero columns: BF16x6 parts add up to
    float32(lambda gamma) exactly; FP16x3's column scale is the exact float32 maximum of
    max(|Re|, |Im|) over k (1 for an all-zero column and for the padded columns) and its two parts
    are the float16 roundings h = fp16(x / scale), l = fp16(x / scale - h)."""
    from bosonsupremacy_amd import runtime
    ctx = runtime.context()
    rng = np.random.default_rng(21)
    chi_p, chi, d = 45, 14, 5                        # kpad 64, ncols 70 -> ncols_pad 128
    ncols, kpad, cpad = chi * d, 64, 128
    G = ((rng.normal(size=(chi_p, chi, d)) + 1j * rng.normal(size=(chi_p, chi, d)))
         * 10.0 ** rng.uniform(-3, 3, size=(chi_p, chi, d))).astype(np.complex64)
    G[:, 3, 2] = 0                                   # column 17 all zero
    G[:, 9, 0] = 0                                   # column 45 all zero
    lam = rng.uniform(0.01, 1, size=chi_p)
    # the kernels' scaling: float32(float64(g) * lambda)
    x = G.reshape(chi_p, ncols)
    xr = (x.real.astype(np.float64) * lam[:, None]).astype(np.float32)
    xi = (x.imag.astype(np.float64) * lam[:, None]).astype(np.float32)
    G_d, lam_d = runtime.to_device(G), runtime.to_device(lam)
    print(f"split planes chi_p={chi_p} ncols={ncols}", flush=True)

    planes = torch.empty((ctx.lib.bsmps_bf16x6_plane_bytes(chi_p, chi, d) // 2,), dtype=torch.bfloat16,
                         device=ctx.device)
    ctx.call("bsmps_split_bf16x6", chi_p, chi, d, runtime.ptr(G_d), runtime.ptr(lam_d), runtime.ptr(planes))
    P = planes.view(cpad, kpad // 32, 6, 32).permute(2, 0, 1, 3).reshape(6, cpad, kpad).double().cpu().numpy()
    re, im = P[0] + P[1] + P[2], P[3] + P[4] + P[5]   # [column][k]
    assert np.array_equal(re[:ncols, :chi_p], xr.T.astype(np.float64))
    assert np.array_equal(im[:ncols, :chi_p], xi.T.astype(np.float64))
    assert not P[:, ncols:, :].any() and not P[:, :, chi_p:].any()

    planes = torch.empty((ctx.lib.bsmps_fp16x3_plane_bytes(chi_p, chi, d) // 2,), dtype=torch.float16,
                         device=ctx.device)
    colscale = torch.full((cpad,), -7.0, dtype=torch.float32, device=ctx.device)  # overwritten entirely
    ctx.call("bsmps_split_fp16x3", chi_p, chi, d, runtime.ptr(G_d), runtime.ptr(lam_d), runtime.ptr(planes),
             runtime.ptr(colscale))
    m = np.maximum(np.abs(xr), np.abs(xi)).max(axis=0)
    want_sc = np.ones(cpad, dtype=np.float32)
    want_sc[:ncols] = np.where(m > 0, m, np.float32(1))
    got_sc = runtime.to_host(colscale)
    assert np.array_equal(got_sc, want_sc)
    assert got_sc[17] == 1 and got_sc[45] == 1
    P = planes.view(cpad, kpad // 32, 4, 32).permute(2, 0, 1, 3).reshape(4, cpad, kpad).cpu().numpy()
    for part, v in ((0, xr), (2, xi)):
        y = (v / want_sc[None, :ncols]).astype(np.float32).T          # [column][k], float32 division
        h = y.astype(np.float16)
        lo = (y - h.astype(np.float32)).astype(np.float16)
        assert np.array_equal(P[part][:ncols, :chi_p], h)
        assert np.array_equal(P[part + 1][:ncols, :chi_p], lo)
    assert not P[:, ncols:, :].any() and not P[:, :, chi_p:].any()


def test_site_contract_fp16x3_matches_matmul():
    """FP16x3 form: every entry within K * 2^-19 * (row scale) * (column scale) of the exact
    product (per product the error is ~2^-21 of the scaled operands), on ragged shapes and on
    operands spanning twelve decades."""
    from bosonsupremacy_amd import runtime
    ctx = runtime.context()
    rng = np.random.default_rng(10)
    # (200, 2100, ...) has k >= 2048 (128-column tiles) and a last tile past the padded columns;
    # (96, 9000, ...) has k > 4096, so it runs in k segments added in FP64
    for n, chi_p, chi, d, spread in ((70, 33, 17, 3, 0), (300, 64, 64, 4, 0), (257, 130, 70, 5, 0),
                                     (129, 100, 30, 6, 6), (200, 2100, 50, 3, 2), (96, 9000, 20, 3, 0)):
        scale_a = 10.0 ** rng.uniform(-spread, spread, size=(n, chi_p))
        scale_g = 10.0 ** rng.uniform(-spread, spread, size=(chi_p, chi, d))
        pre = ((rng.normal(size=(n, chi_p)) + 1j * rng.normal(size=(n, chi_p))) * scale_a).astype(np.complex64)
        lam = rng.uniform(0.1, 1, size=chi_p)
        G = ((rng.normal(size=(chi_p, chi, d)) + 1j * rng.normal(size=(chi_p, chi, d))) * scale_g).astype(np.complex64)
        A = pre.astype(np.complex128)
        B = G.astype(np.complex128).reshape(chi_p, chi * d) * lam[:, None]
        want = (A @ B).reshape(n, chi, d)
        rs = np.maximum(np.abs(A.real), np.abs(A.imag)).max(axis=1)
        cs = np.maximum(np.abs(B.real), np.abs(B.imag)).max(axis=0)
        bound = (chi_p * np.outer(rs, cs)).reshape(n, chi, d)
        got = _fp16x3_contract(ctx, pre, lam, G)
        ratio = (np.abs(got - want) / bound).max()
        typical = np.median(np.abs(got - want) / np.sqrt(chi_p * np.outer(rs, cs)).reshape(n, chi, d))
        print(f"fp16x3 {n}x{chi_p}x{chi * d} spread {spread}: max err / (K rowscale colscale) = {ratio:.2e}, "
              f"median err / (sqrt(K) rowscale colscale) = {typical:.2e}", flush=True)
        assert ratio <= 2.0 ** -19, (n, chi_p, chi, d)


def test_sampling_fp16x3_form_m8():
    """FP16x3 form: conditional distributions within 1e-5 of the FP64 form on the same random
    numbers, the same draws away from cdf boundaries."""
    from bosonsupremacy_amd import runtime, sampling
    inst, G, L = _mps_m8()
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    n = 3000
    rng = np.random.default_rng(11)
    normals = runtime.to_device(rng.normal(size=(2 * inst.M, n)))
    uniforms = runtime.to_device(rng.uniform(size=(inst.M, n)))
    a, pa = sampling.Sampler(G, L, sqrtW, inst.dd).run(normals, uniforms, return_probs=True)
    b, pb = sampling.Sampler(G, L, sqrtW, inst.dd, precision="fp16x3").run(normals, uniforms, return_probs=True)
    a, b, pa, pb = (runtime.to_host(x) for x in (a, b, pa, pb))
    same = np.all(a == b, axis=1)
    print("fp16x3 max prob diff", np.abs(pa[:, same] - pb[:, same]).max(), "same", same.mean(), flush=True)
    assert same.mean() > 0.99
    assert np.abs(pa[:, same] - pb[:, same]).max() < 1e-5
