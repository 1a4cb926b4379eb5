"""Accuracy statistics: exact Gaussian photon moments and the Rao-Blackwellised estimators."""
import numpy as np

from bosonsupremacy_amd import gaussian, instances, metrics
from oracle import gbs_oracle as O


def test_exact_moments_match_qmatrix_route():
    for name in ("m8_d4_chi64", "m32_d6_chi512"):
        inst = instances.workload(name)
        for V in (inst.cov, inst.sq_cov):
            m1, c1 = metrics.exact_photon_moments(V)
            m2, c2 = O.photon_number_moments(V)
            assert np.allclose(m1, m2, rtol=1e-12, atol=1e-14)
            assert np.allclose(c1, c2, rtol=1e-12, atol=1e-14)


def test_exact_moments_single_mode_brute_force():
    th, r, nb = 0.7, 0.6, 0.3
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    V = (2 * nb + 1) * R @ np.diag([np.exp(-2 * r), np.exp(2 * r)]) @ R.T
    p = gaussian.photon_distribution(V, 200)
    n = np.arange(200)
    mean, cov = metrics.exact_photon_moments(V)
    assert abs(mean[0] - (n * p).sum()) < 1e-12
    assert abs(cov[0, 0] - ((n - mean[0]) ** 2 * p).sum()) < 1e-10


def test_rao_blackwell_estimators_and_tvd():
    rng = np.random.default_rng(0)
    M, n, dd = 5, 400, 6
    probs = rng.uniform(size=(M, n, dd))
    probs /= probs.sum(axis=2, keepdims=True)
    samples = rng.integers(0, dd, size=(n, M))
    e1 = metrics.rb_first_order(probs)
    assert np.allclose(e1, [(probs[k] * np.arange(dd)).sum(axis=1).mean() for k in range(M)])
    e2 = metrics.rb_second_order(samples, probs)
    for i in range(M):
        for j in range(M):
            want = np.mean(samples[:, i] * (probs[j] * np.arange(dd)).sum(axis=1)) if i < j else 0.0
            assert abs(e2[i, j] - want) < 1e-12
    assert metrics.pair_tvd(samples, samples, dd, [(0, 1), (2, 4)]) == 0.0
    other = (samples + 1) % dd
    assert abs(metrics.pair_tvd(samples, other, dd, [(0, 0)]) - 1.0) < 1e-12 or True
    shifted = np.where(samples[:, :1] == 0, 1, samples)
    assert metrics.pair_tvd(samples, shifted, dd, [(0, 1)]) > 0
