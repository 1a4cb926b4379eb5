"""Host-side product code (no GPU): integer helpers, Gaussian linear algebra, instances."""
import numpy as np
import pytest

from conftest import same_bits
from oracle import gbs_oracle as O
from bosonsupremacy_amd import gaussian, instances, kron, mps, sampling


def test_get_target_and_push_to_end_match_reference(golden_m8):
    g = golden_m8
    for k in range(7):
        num = g[f"kron_num_{k}"].reshape(g[f"kron_num_{k}"].shape[0], -1)
        t = mps.get_target(num)
        assert t.dtype == np.int32
        assert same_bits(t, g[f"target_{k}"])
        assert same_bits(mps.push_to_end(t), g[f"pushed_{k}"])


def test_get_target_edge_cases():
    assert mps.get_target(np.zeros((3, 4), np.int8)).shape == (3, 0)
    assert mps.get_target(np.zeros((0, 4), np.int8)).shape == (0, 0)
    num = np.array([[0, 2, 1], [3, 0, 0], [0, 0, 0]], np.int8)
    assert same_bits(mps.get_target(num), O.get_target(num))
    assert same_bits(mps.push_to_end(mps.get_target(num)), O.push_to_end(O.get_target(num)))


def test_gaussian_gauge_matches_reference(golden_m8):
    g = golden_m8
    Db, S = gaussian.williamson(g["sq_cov"])
    assert same_bits(S, g["S_full"]) and same_bits(Db, g["Db_full"])
    M = g["sq_cov"].shape[0] // 2
    S_cuts = [g[f"kron_S_{k}"] for k in range(M - 1)]
    for k in range(M):
        if k == 0:
            S_r, S_l = g["S_full"], S_cuts[0]
        elif k == M - 1:
            S_r, S_l = S_cuts[k - 1], np.zeros((0, 0))
        else:
            S_r, S_l = S_cuts[k - 1], S_cuts[k]
        U2, sq, U1 = gaussian.get_U2_sq_U1(S_l, S_r)
        assert same_bits(U2, g[f"U2_{k}"]) and same_bits(sq, g[f"sq_{k}"]) and same_bits(U1, g[f"U1_{k}"])
        Sig = gaussian.get_Sigma(U2, sq, U1)
        assert same_bits(Sig.astype(np.complex64), g[f"Sigma64_{k}"])
        Sig2, Z = gaussian.site_kernel(k, M, g["sq_cov"], S_cuts, g["S_full"])
        oSig, oZ = O.site_sigma(k, M, g["sq_cov"], S_cuts)
        assert same_bits(Sig2, oSig) and Z == oZ


def test_williamson_errors():
    with pytest.raises(ValueError):
        gaussian.williamson(np.ones((3, 4)))
    with pytest.raises(ValueError):
        gaussian.williamson(np.array([[1.0, 2.0], [0.0, 1.0]]))
    with pytest.raises(ValueError):
        gaussian.williamson(-np.eye(4))


def test_kron_host_inputs_match_oracle(golden_m8):
    g = golden_m8
    for k in range(7):
        tp, S = kron.cut_spectrum_inputs(g["sq_cov"], k + 1, int(g["d"]))
        otp, oS = O.cut_thermal(g["sq_cov"], k + 1, int(g["d"]))
        assert same_bits(S, g[f"kron_S_{k}"]) and same_bits(S, oS)
        assert same_bits(tp, np.stack(otp))


@pytest.mark.parametrize("name", ["m8_d4_chi64", "m32_d6_chi512"])
def test_instances(name):
    inst = instances.workload(name)
    T = np.sqrt(inst.eta) * inst.U
    cov = O.cov_from_T(inst.r, T) + 2 * inst.nth * np.eye(2 * inst.M)
    assert np.abs(cov - inst.cov).max() < 1e-12
    assert np.abs(inst.U @ inst.U.conj().T - np.eye(inst.M)).max() < 1e-12
    Db, _ = O.williamson(inst.sq_cov)
    assert np.abs(np.diag(Db) - 1).max() < 1e-8          # pure
    assert np.linalg.eigvalsh(inst.cov - inst.sq_cov).min() > -1e-12   # W >= 0
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    assert same_bits(sqrtW, O.thermal_sqrtW(inst.cov, inst.sq_cov))
    assert np.abs(sqrtW @ sqrtW.T - (inst.cov - inst.sq_cov)).max() < 1e-10


def test_displacement_basis_reconstructs_generator():
    for dd in (4, 10, 20):
        W, x = sampling.displacement_basis(dd)
        a = sampling.destroy(dd).real
        assert np.abs(W @ np.diag(x) @ W.T - (a + a.T)).max() < 1e-12
