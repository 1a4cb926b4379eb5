"""The host worker pool computes exactly what a serial run computes."""
import numpy as np

from bosonsupremacy_amd import gaussian, hostpool, instances


def test_pool_matches_serial_m32():
    inst = instances.workload("m32_d6_chi512")
    M = inst.M
    cuts = hostpool.cut_inputs_all(inst.sq_cov, inst.d)
    want = [gaussian.cut_inputs(inst.sq_cov, k + 1, inst.d) for k in range(M - 1)]
    for (tp, S), (tp2, S2) in zip(cuts, want):
        assert np.array_equal(tp, tp2) and np.array_equal(S, S2)
    kS = [S for _, S in cuts]
    S_full = hostpool.williamson_S(inst.sq_cov).result()
    assert np.array_equal(S_full, gaussian.williamson(inst.sq_cov)[1])
    sites = [t.result() for t in hostpool.submit_sites(inst.sq_cov, kS, S_full)]
    for k in range(M):
        Sig, Z = gaussian.site_kernel(k, M, inst.sq_cov, kS, S_full)
        assert np.array_equal(sites[k][0], Sig) and sites[k][1] == Z, k


def test_single_thread_blas_is_reproducible():
    # the decompositions give the same bits whatever BLAS thread count the caller set
    from threadpoolctl import threadpool_limits
    inst = instances.workload("m32_d6_chi512")
    S = gaussian.cut_inputs(inst.sq_cov, 1, inst.d)[1]
    with threadpool_limits(4):
        S4 = gaussian.cut_inputs(inst.sq_cov, 1, inst.d)[1]
    assert np.array_equal(S, S4)
