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
* Correlation summaries of sample sets against the ideal state (Pearson r, two-norm, fitted
  slope), for photon numbers and for threshold-detector clicks.
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


# ---------------------------------------------------------------------------------------
# Correlation summaries against the ideal state (analysis/Correlation_photon_number.ipynb
# cells 2-6, analysis/Correlation_photon_click.ipynb cells 1-11)
# ---------------------------------------------------------------------------------------

def sample_photon_correlations(samples):
    """(first order [M], second order [M(M-1)/2]) of photon-number samples [n, M]: the means
    E[n_i] and the covariances E[n_i n_j] - E[n_i] E[n_j] for i < j, in the order of
    itertools.combinations (Correlation_photon_number.ipynb cells 2-3)."""
    s = np.asarray(samples, dtype=np.float64)
    mean = s.mean(axis=0)
    cov = (s.T @ s) / s.shape[0] - np.outer(mean, mean)
    iu = np.triu_indices(s.shape[1], k=1)
    return mean, cov[iu]


def ideal_photon_correlations(V):
    """(first order [M], second order [M(M-1)/2]) of the zero-mean Gaussian state V: E[n_i] and
    the two-mode cumulants Cov(n_i, n_j), i < j (the ideal side of the same cells)."""
    mean, cov = exact_photon_moments(V)
    return mean, cov[np.triu_indices(len(mean), k=1)]


def vacuum_probability(V, modes):
    """P(no photon in any of ``modes``) of the zero-mean Gaussian state V (xxpp, hbar = 2):
    1 / sqrt(det((V_S + I) / 2)) over the reduced covariance V_S of those modes."""
    M = V.shape[0] // 2
    idx = list(modes) + [m + M for m in modes]
    Vs = np.asarray(V, dtype=np.float64)[np.ix_(idx, idx)]
    return 1.0 / np.sqrt(np.linalg.det((Vs + np.eye(len(idx))) / 2.0))


def ideal_click_correlations(V):
    """(first order [M], second order [M(M-1)/2]) of threshold-detector clicks of the state V:
    P(click_i) = 1 - P0_i and the two-mode cumulants E[c_i c_j] - E[c_i] E[c_j] = P0_ij - P0_i P0_j
    (Correlation_photon_click.ipynb cells 1 and 4, thewalrus click_cumulant at orders 1-2)."""
    M = V.shape[0] // 2
    p0 = np.array([vacuum_probability(V, [i]) for i in range(M)])
    second = [vacuum_probability(V, [i, j]) - p0[i] * p0[j] for i in range(M) for j in range(i + 1, M)]
    return 1.0 - p0, np.array(second)


def to_clicks(samples):
    """Photon numbers -> threshold-detector clicks (1 where at least one photon)."""
    return (np.asarray(samples) > 0).astype(np.int8)


def _set_partitions(items):
    if len(items) == 1:
        yield [items]
        return
    first, rest = items[0], items[1:]
    for smaller in _set_partitions(rest):
        for i in range(len(smaller)):
            yield smaller[:i] + [[first] + smaller[i]] + smaller[i + 1:]
        yield [[first]] + smaller


def click_cumulant_sample(clicks):
    """Joint cumulant of the click columns [n, k] from sample averages: the sum over set
    partitions pi of (|pi| - 1)! (-1)^(|pi| - 1) prod_{B in pi} E[prod_{i in B} c_i]
    (Correlation_photon_click.ipynb cell 1)."""
    c = np.asarray(clicks, dtype=bool)
    total = 0.0
    for pi in _set_partitions(list(range(c.shape[1]))):
        term = float(np.prod(np.arange(1, len(pi)))) * (-1.0) ** (len(pi) - 1)
        for block in pi:
            term *= np.mean(np.all(c[:, block], axis=1))
        total += term
    return total


def sample_click_correlations(samples):
    """(first order [M], second order [M(M-1)/2]) of click samples (after to_clicks): click
    rates and two-mode click cumulants."""
    c = to_clicks(samples).astype(np.float64)
    first = c.mean(axis=0)
    # at order two the partition sum is E[c_i c_j] - E[c_i] E[c_j]: one product for all pairs
    second = (c.T @ c) / c.shape[0] - np.outer(first, first)
    return first, second[np.triu_indices(c.shape[1], k=1)]


def correlation_summary(ideal, sampled):
    """Agreement of sampled correlations with the ideal ones, as the notebooks report it:
    Pearson r (np.corrcoef), two-norm of the difference, and the slope and intercept of the
    least-squares line sampled ~ ideal (np.polyfit, degree 1)."""
    x = np.asarray(ideal, dtype=np.float64)
    y = np.asarray(sampled, dtype=np.float64)
    slope, intercept = np.polyfit(x, y, 1)
    return {"pearson_r": float(np.corrcoef(x, y)[0, 1]), "two_norm": float(np.linalg.norm(x - y)),
            "slope": float(slope), "intercept": float(intercept)}
