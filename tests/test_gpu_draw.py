"""Kernel-level checks of the draw (displacement, conditional probabilities, inverse CDF, carried
vector; sampling_cpu.py:81-113) on synthetic T, Lambda and alpha, in both storage forms, on the
shapes each kernel variant serves:

* d = 10, dd = 20 and d = 16, dd = 40: the one-wave-per-sample kernel (d > 8 or dd > 16; the
  M = 144 workload uses d = 10, dd = 20);
* d = 8, dd = 16 with chi = 1200 and d = 6, dd = 12 with chi = 600: the Gram kernel with the
  sample's rows staged through LDS in several chunks (more rows than one stage holds);
* d = 4, dd = 8 with chi = 100: the Gram kernel with all rows resident;
* the last site (no Lambda: row 0 only).
"""
import numpy as np
import pytest
import torch

from oracle import gbs_oracle as O

pytestmark = pytest.mark.gpu


def _reference(T, alpha, lam, u, dd):
    """Probabilities (normalised), picks, carried vectors and unnormalised p of one site."""
    n, chi, d = T.shape
    D = O.batch_displaces(dd, alpha[:, None])[:, 0]                  # [n, dd, dd]
    TD = np.einsum('sad,skd->sak', T, D[:, :, :d])                     # sum_j D[k, j] T[a, j]
    if lam is None:
        p = np.abs(TD[:, 0, :]) ** 2
    else:
        p = (np.abs(TD) ** 2 * (lam ** 2)[None, :, None]).sum(axis=1)
    pn = p / p.sum(axis=1, keepdims=True)
    cdf = np.cumsum(pn, axis=1)
    pick = np.minimum((u[:, None] > cdf).sum(axis=1), dd - 1)
    pre = TD[np.arange(n), :, pick]
    return pn, pick, pre, p[np.arange(n), pick], cdf


CASES = [  # (n, chi, d, dd, last)
    (96, 300, 10, 20, False), (64, 64, 16, 40, False), (40, 1200, 8, 16, False), (48, 600, 6, 12, False),
    (80, 100, 4, 8, False), (50, 1, 10, 20, True), (50, 1, 6, 12, True),
    (32, 2000, 10, 20, False), (40, 500, 9, 18, False), (30, 700, 10, 16, False),
]


@pytest.mark.parametrize("form", ["fp64", "c64"])
@pytest.mark.parametrize("case", CASES, ids=[f"n{c[0]}_chi{c[1]}_d{c[2]}_dd{c[3]}{'_last' if c[4] else ''}"
                                             for c in CASES])
def test_draw_kernel_matches_reference(form, case):
    from bosonsupremacy_amd import runtime, sampling
    ctx = runtime.context()
    n, chi, d, dd, last = case
    rng = np.random.default_rng(hash(case) % 2 ** 32)
    T = (rng.normal(size=(n, chi, d)) + 1j * rng.normal(size=(n, chi, d))) * np.exp(-0.4 * np.arange(d))
    if form == "c64":
        T = T.astype(np.complex64).astype(np.complex128)   # the values the kernel reads
    lam = None if last else np.sort(rng.uniform(0.01, 1, size=chi))[::-1].copy()
    alpha = (rng.normal(size=n) + 1j * rng.normal(size=n)) * 0.4
    u = rng.uniform(size=n)
    pn, pick, pre, p_pick, cdf = _reference(T, alpha, lam, u, dd)

    ctype = torch.complex128 if form == "fp64" else torch.complex64
    T_d = torch.from_numpy(T).to(ctype).to(ctx.device).contiguous()
    a_d = runtime.to_device(alpha.astype(np.complex128))
    W, x = sampling.displacement_basis(dd)
    W_d, x_d = runtime.to_device(W), runtime.to_device(x)
    dwork = torch.empty((2 * n * dd * d,), dtype=torch.float64, device=ctx.device)
    lam_d = runtime.to_device(lam) if lam is not None else None
    u_d = runtime.to_device(u)
    smp = torch.full((n,), -1, dtype=torch.int32, device=ctx.device)
    pre_d = torch.zeros((n, chi), dtype=ctype, device=ctx.device)
    probs = torch.empty((n, dd), dtype=torch.float64, device=ctx.device)
    print(f"draw {form} n={n} chi={chi} d={d} dd={dd} last={last}", flush=True)
    ctx.call("bsmps_sample_site" if form == "fp64" else "bsmps_sample_site_c64", n, chi, d, dd,
             runtime.ptr(T_d), chi * d, runtime.ptr(a_d), 1, runtime.ptr(W_d), runtime.ptr(x_d), runtime.ptr(dwork),
             runtime.ptr(lam_d), runtime.ptr(u_d), runtime.ptr(smp), 1, None if last else runtime.ptr(pre_d),
             runtime.ptr(probs))
    got_pick = runtime.to_host(smp)
    got_p = runtime.to_host(probs)
    assert np.abs(got_p - pn).max() < 1e-11
    same = got_pick == pick
    # a differing draw must sit on a cdf boundary (the probabilities are rounded differently)
    for s in np.where(~same)[0]:
        assert np.min(np.abs(cdf[s] - u[s])) < 1e-10, s
    assert same.mean() > 0.95
    if last:
        return
    got_pre = runtime.to_host(pre_d)[same]
    want = pre[same]
    if form == "fp64":
        assert np.abs(got_pre - want).max() <= 1e-13 * np.abs(want).max()
    else:
        # the FP32 forms carry D_pick t_a / sqrt(p_pick) (Lambda-weighted norm 1)
        want = want / np.sqrt(p_pick[same])[:, None]
        assert np.abs(got_pre - want).max() <= 2e-6 * np.abs(want).max()
        assert np.allclose((np.abs(got_pre) ** 2 * lam ** 2).sum(axis=1), 1.0, rtol=1e-5)
