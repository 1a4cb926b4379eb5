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


def test_site_tables_from_pool_match_serial_and_targets(golden_m8):
    """Label tables built by the host workers equal those built in-process, and their targets
    are the reference's get_target rows (right side shifted past the site's output modes)."""
    from bosonsupremacy_amd import mps
    g = golden_m8
    M, d, chi = 8, int(g["d"]), int(g["chi"])
    kres = [g[f"kron_res16_{k}"].astype(np.float64) for k in range(M - 1)]
    knum = [g[f"kron_num_{k}"] for k in range(M - 1)]
    pooled = [t.result() for t in hostpool.submit_tables(M, d, chi, kres, knum, 1e-10)]
    for k in range(M):
        t = mps.site_tables(k, M, d, chi, kres, knum)
        p = pooled[k]
        assert (t.kind, t.n_a, t.n_b) == (p.kind, p.n_a, p.n_b) and np.array_equal(t.Lambda, p.Lambda)
        for a, b in ((t.left, p.left), (t.right, p.right)):
            for name in ("targets", "den", "perm", "class_count", "count"):
                assert np.array_equal(getattr(a, name), getattr(b, name)), (k, name)
        def padded(a, w):  # zero padding to the reference's row width
            out = np.zeros((a.shape[0], max(w, a.shape[1])), np.int32)
            out[:, :a.shape[1]] = a
            return out
        if k < M - 1:
            keep = kres[k] > 1e-10
            want = g[f"target_{k}"][keep]
            assert np.array_equal(padded(t.left.targets, want.shape[1]), padded(want, 0)), k
        if k > 0:
            want = g[f"target_{k - 1}"]
            shifted = np.where(want != 0, want + (M - k - 1), 0)
            assert np.array_equal(padded(t.right.targets, want.shape[1]), padded(shifted, 0)), k


def test_right_label_runs_partition_the_labels(golden_m8):
    """Runs of right labels are contiguous, cover every label once, respect the requested count
    and balance the m^3 2^m work estimate."""
    from bosonsupremacy_amd import mps
    g = golden_m8
    M, d, chi = 8, int(g["d"]), int(g["chi"])
    kres = [g[f"kron_res16_{k}"].astype(np.float64) for k in range(M - 1)]
    knum = [g[f"kron_num_{k}"] for k in range(M - 1)]
    for k in range(M):
        t = mps.site_tables(k, M, d, chi, kres, knum)
        n = len(t.right.perm)
        assert mps.right_label_runs(t, d, 1) == [(0, n)]
        for want in (2, 3, 7, 1000):
            runs = mps.right_label_runs(t, d, want)
            assert runs[0][0] == 0 and runs[-1][1] == n and len(runs) <= min(want, n)
            assert all(a < b for a, b in runs) and all(runs[i][1] == runs[i + 1][0] for i in range(len(runs) - 1))
            # per-run class counts add up to the site's
            tot = sum(np.bincount(t.right.count[t.right.perm[a:b]], minlength=len(t.right.class_count))
                      for a, b in runs)
            assert np.array_equal(tot, t.right.class_count), (k, want)
        if n >= 16:  # two runs: each within one label's work of half the total
            lcc = t.left.class_count
            def w(q):
                return sum(lcc[p] * max(1, ((j + p + q) // 2) ** 3) * 2.0 ** ((j + p + q) // 2)
                           for j in range(d) for p in range(len(lcc)) if (j + p + q) % 2 == 0 and lcc[p])
            per = np.array([w(q) for q in t.right.count[t.right.perm]])
            (a0, b0), (a1, b1) = mps.right_label_runs(t, d, 2)
            assert abs(per[a0:b0].sum() - per[a1:b1].sum()) <= 2 * per.max() + 1e-9 * per.sum(), k
