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
    # disjoint supports: distance 1
    zeros, ones = np.zeros((n, M), np.int64), np.ones((n, M), np.int64)
    assert abs(metrics.pair_tvd(zeros, ones, dd, [(0, 1)]) - 1.0) < 1e-12
    # against a histogram by hand
    other = rng.integers(0, dd, size=(n, M))
    ha = np.zeros((dd, dd))
    hb = np.zeros((dd, dd))
    np.add.at(ha, (samples[:, 2], samples[:, 4]), 1.0 / n)
    np.add.at(hb, (other[:, 2], other[:, 4]), 1.0 / n)
    assert abs(metrics.pair_tvd(samples, other, dd, [(2, 4)]) - 0.5 * np.abs(ha - hb).sum()) < 1e-12
    shifted = np.where(samples[:, :1] == 0, 1, samples)
    assert metrics.pair_tvd(samples, shifted, dd, [(0, 1)]) > 0


def test_sample_correlations_by_hand():
    rng = np.random.default_rng(1)
    s = rng.integers(0, 4, size=(500, 5))
    first, second = metrics.sample_photon_correlations(s)
    assert np.allclose(first, s.mean(axis=0))
    want = [np.mean(s[:, i] * s[:, j]) - s[:, i].mean() * s[:, j].mean() for i in range(5) for j in range(i + 1, 5)]
    assert np.allclose(second, want, atol=1e-12)
    c1, c2 = metrics.sample_click_correlations(s)
    clicks = s > 0
    assert np.allclose(c1, clicks.mean(axis=0))
    pairs = [(i, j) for i in range(5) for j in range(i + 1, 5)]
    assert np.allclose(c2, [metrics.click_cumulant_sample(clicks[:, [i, j]]) for i, j in pairs], atol=1e-12)
    # third order: the partition sum against the expanded joint cumulant
    x, y, z = (clicks[:, k].astype(float) for k in range(3))
    E = np.mean
    k3 = E(x * y * z) - E(x * y) * E(z) - E(x * z) * E(y) - E(y * z) * E(x) + 2 * E(x) * E(y) * E(z)
    assert abs(metrics.click_cumulant_sample(clicks[:, :3]) - k3) < 1e-12


def test_click_probabilities_against_fock_closed_form():
    """Vacuum probabilities 1/sqrt(det((V_S + I)/2)) of reduced pure states against sums of the
    closed-form Fock probabilities (tests/closed_forms.py), and a thermal mode's 1/(1 + nbar)."""
    import closed_forms as CF
    inst = instances.make_instance("clicks", 4, 0.1, 0.3, 1.0, 6, 64, 10, 10, seed=7)
    B, detQ = CF.pure_state_B(inst.sq_cov)
    P = np.abs(CF.fock_amplitudes(B, np.zeros(4), 16)) ** 2 / np.sqrt(detQ)
    assert abs(P.sum() - 1) < 1e-9
    for modes in ([0], [2], [1, 3], [0, 1, 2]):
        idx = tuple(0 if m in modes else slice(None) for m in range(4))
        assert abs(metrics.vacuum_probability(inst.sq_cov, modes) - P[idx].sum()) < 1e-9
    first, second = metrics.ideal_click_correlations(inst.sq_cov)
    p0 = [P[tuple(0 if m == i else slice(None) for m in range(4))].sum() for i in range(4)]
    assert np.allclose(first, 1 - np.array(p0), atol=1e-9)
    p01 = P[0, 0].sum()
    assert abs(second[0] - (p01 - p0[0] * p0[1])) < 1e-9
    nbar = 0.37
    V = (2 * nbar + 1) * np.eye(2)
    assert abs(metrics.vacuum_probability(V, [0]) - 1 / (1 + nbar)) < 1e-14


def test_correlation_summary():
    x = np.linspace(0.0, 1.0, 50)
    out = metrics.correlation_summary(x, 2 * x + 1)
    assert abs(out["pearson_r"] - 1) < 1e-12 and abs(out["slope"] - 2) < 1e-12 and abs(out["intercept"] - 1) < 1e-12
    y = x + np.random.default_rng(0).normal(scale=0.1, size=50)
    out = metrics.correlation_summary(x, y)
    assert abs(out["pearson_r"] - np.corrcoef(x, y)[0, 1]) < 1e-15
    assert abs(out["two_norm"] - np.linalg.norm(x - y)) < 1e-15
