"""Accuracy statistics of sampled photon numbers (the validation half of the metric).

* Rao-Blackwellised marginals (the estimator of Correlation_photon_number.ipynb, with the
  conditional distributions the sampler already computes): E[n_k] = mean_s sum_n n p_k(n|s),
  E[n_i n_j] = mean_s n_i(s) sum_n n p_j(n|s) for i < j.  With common random numbers two forms
  of the sampler differ only numerically, so these isolate numerical error from sampling noise.
* Total-variation distance between the empirical joint distributions of (n_i, n_j) of two
  sample sets (low-order correlations), maximised over the pairs given.
* Exact photon-number moments of a zero-mean Gaussian state (hbar = 2, xxpp covariance V):
  with N = <a^+ a^T> and M = <a a^T>, N = (Vxx + Vpp + i (Vxp - Vpx)) / 4 - I / 2 and
  M = (Vxx - Vpp + i (Vxp + Vpx)) / 4; E[n_i] = N_ii and, by Wick's theorem,
  Cov(n_i, n_j) = |N_ij|^2 + |M_ij|^2 (i != j), Var(n_i) = N_ii (N_ii + 1) + |M_ii|^2.
"""
import numpy as np
import torch


def rb_first_order(probs):
    """probs [M, n, dd] (device or host) -> E[n_k] [M] (host float64)."""
    p = torch.as_tensor(probs)
    k = torch.arange(p.shape[2], device=p.device, dtype=p.dtype)
    return (p * k).sum(dim=2).mean(dim=1).cpu().numpy()


def rb_second_order(samples, probs):
    """samples [n, M] int, probs [M, n, dd] -> E[n_i n_j] [M, M] (host float64; i < j filled,
    the rest zero)."""
    p = torch.as_tensor(probs)
    s = torch.as_tensor(samples, device=p.device).to(p.dtype)          # [n, M]
    k = torch.arange(p.shape[2], device=p.device, dtype=p.dtype)
    cond = (p * k).sum(dim=2)                                           # [M, n]: E[n_j | past]
    full = (s.T @ cond.T) / s.shape[0]                                  # [i, j] = mean n_i E[n_j|.]
    out = torch.triu(full, diagonal=1)
    return out.cpu().numpy()


def pair_tvd(samples_a, samples_b, dd, pairs):
    """Max over (i, j) in pairs of the total-variation distance between the empirical joint
    distributions of (n_i, n_j) in two sample sets [n, M]."""
    a = torch.as_tensor(samples_a).long()
    b = torch.as_tensor(samples_b, device=a.device).long()
    worst = 0.0
    for i, j in pairs:
        ha = torch.bincount(a[:, i] * dd + a[:, j], minlength=dd * dd).double() / a.shape[0]
        hb = torch.bincount(b[:, i] * dd + b[:, j], minlength=dd * dd).double() / b.shape[0]
        worst = max(worst, 0.5 * float(torch.abs(ha - hb).sum()))
    return worst


def exact_photon_moments(V):
    """(E[n] [M], Cov(n_i, n_j) [M, M]) of the zero-mean Gaussian state with covariance V."""
    V = np.asarray(V, dtype=np.float64)
    M = V.shape[0] // 2
    xx, pp, xp, px = V[:M, :M], V[M:, M:], V[:M, M:], V[M:, :M]
    N = (xx + pp + 1j * (xp - px)) / 4 - np.eye(M) / 2
    Mm = (xx - pp + 1j * (xp + px)) / 4
    mean = N.diagonal().real.copy()
    cov = np.abs(N) ** 2 + np.abs(Mm) ** 2
    cov[np.diag_indices(M)] = mean * (mean + 1) + np.abs(Mm.diagonal()) ** 2
    return mean, cov
