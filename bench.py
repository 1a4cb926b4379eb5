"""Throughput of the MPS construction + sampling path on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload m32_d6_chi512]

One step is one pass of the path over one synthetic instance: Schmidt spectra of all M-1 cuts,
the M site tensors, and n_samples photon-number samples.  Every input (covariances, site
tensors, random streams) is resident on the GPU or generated there before the sampling
kernels start.  Prints one JSON line:

  value            MPS sites/s (sites / construction time, construction = spectra + tensors)
  samples_per_s    full M-site samples per second
  roofline         the dominant kernel (by GPU time) against its MI355X peak
  kernels          every kernel family: launches, mean duration, achieved rate
  cpu_baseline     the FP64 CPU oracle timed on a bounded sample of the same workload
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 vector and FP64 matrix (dense)
PEAK_HBM_GBS = 8000.0
PEAK_FP32_MFMA_TFLOPS = 157.3  # MI355X dense FP32 matrix
PEAK_BF16_MFMA_TFLOPS = 2500.0  # MI355X dense BF16 matrix


def hafnian_flops(m):
    """Real flops of the reference recursion for a 2m x 2m hafnian that reach the result:
    structurally-zero polynomial terms are skipped, and so are coefficients above the degree
    that can still reach the leaf value (m-1 for pivot-pair entries (2p+1, 2p), m-2 for the
    others).  This is the arithmetic the GPU kernels execute: per child-matrix term two
    complex products and two adds (16), per g-update term a product and an add (8)."""
    n = m
    memo = {}

    def nz(level, D, G):
        if level == n:
            return G >= n
        key = (level, D, G)
        if key not in memo:
            memo[key] = nz(level + 1, D, G) or nz(level + 1, min(n, 2 * D + 1), min(n, G + D + 1))
        return memo[key]

    def rec(level, D, G):
        if level == n:
            return 0
        s = 2 * n - 2 * level
        D2, G2 = min(n, 2 * D + 1), min(n, G + D + 1)
        f = rec(level + 1, D, G) if nz(level + 1, D, G) else 0
        if nz(level + 1, D2, G2):
            nc = (s - 2) * (s - 3) // 2
            npair = n - level - 1  # child entries (2q+1, 2q)

            def pc(cap):
                return sum(1 for t in range(1, min(D2, cap) + 1) for u in range(t)
                           if u <= D and t - 1 - u <= D)
            pe = sum(1 for t in range(1, n + 1) for u in range(t) if u <= G and t - 1 - u <= D)
            f += 16 * ((nc - npair) * pc(n - 2) + npair * pc(n - 1)) + 8 * pe + rec(level + 1, D2, G2)
        return f
    return rec(0, 0, 0) if m >= 3 else {0: 0, 1: 0, 2: 22}[m]


def site_hafnian_work_by_site(kres, knum, d, err_tol=1e-10, site_d=None, site_chi=None):
    """[(hafnian count, flops, largest size)] of every site tensor, from the label photon counts.
    ``site_d`` / ``site_chi`` = per-site cutoffs and per-cut ranks of the learned-cutoff form."""
    M = len(kres) + 1
    out = []
    fl = {}
    for k in range(M):
        if k == M - 1:
            la = np.zeros(1, int)
        else:
            la = np.asarray(knum[k]).reshape(len(kres[k]), -1)[np.asarray(kres[k]) > err_tol].sum(1)
            if site_chi is not None:
                la = la[:site_chi[k]]
        lb = np.zeros(1, int) if k == 0 else np.asarray(knum[k - 1]).reshape(len(knum[k - 1]), -1).sum(1)
        if k > 0 and site_chi is not None:
            lb = lb[:site_chi[k - 1]]
        ca, cb = np.bincount(la), np.bincount(lb)
        cnt, flops, big = 0, 0.0, 0
        for j in range(d if site_d is None else site_d[k]):
            for p in range(len(ca)):
                for q in range(len(cb)):
                    n = j + p + q
                    if n % 2 or ca[p] == 0 or cb[q] == 0:
                        continue
                    c = int(ca[p]) * int(cb[q])
                    if n // 2 not in fl:
                        fl[n // 2] = hafnian_flops(n // 2)
                    cnt += c
                    flops += c * fl[n // 2]
                    big = max(big, n)
        out.append((cnt, flops, big))
    return out


def probe_sampling(inst, ko, probe, n, tuning):
    """Time one site's sampling step at full size: the contraction of n carried vectors with
    a real Gamma_k (FP64, FP32, BF16x6 and FP16x3 forms) and the displacement + draw, then
    extrapolate the chain's samples/s as n / ((M - 1) x per-site time)."""
    import torch
    from bosonsupremacy_amd import runtime, sampling
    ctx = runtime.context()
    dev = ctx.device
    k, g_host, lam_k = probe
    chi_p, chi, d = g_host.shape
    G = g_host.to(dev)
    # Lambda vectors padded to the Gamma extents (the kernels read chi_p and chi entries)
    res_prev = np.asarray(ko[0][k - 1], dtype=np.float64)[:chi_p]
    lp = np.zeros(chi_p)
    lp[:len(res_prev)] = np.sqrt(res_prev)
    lam_prev = runtime.to_device(lp)
    lk = np.zeros(chi)
    lk[:min(chi, len(lam_k))] = np.asarray(lam_k, dtype=np.float64)[:chi]
    lam = runtime.to_device(lk)
    rng = torch.Generator(device=dev).manual_seed(0)
    pre = torch.randn((n, chi_p), dtype=torch.complex128, device=dev, generator=rng)
    pre /= torch.linalg.vector_norm(pre, dim=1, keepdim=True)
    T = torch.empty((n, chi, d), dtype=torch.complex128, device=dev)

    def timed(fn, reps=2):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps / 1e3

    t64 = timed(lambda: ctx.call("bsmps_site_contract", n, chi_p, chi, d, runtime.ptr(pre), runtime.ptr(lam_prev),
                                 runtime.ptr(G), runtime.ptr(T)))
    G32 = G.to(torch.complex64)
    del G
    pre32 = pre.to(torch.complex64)
    T32 = torch.empty((n, chi, d), dtype=torch.complex64, device=dev)
    t32 = timed(lambda: ctx.call("bsmps_site_contract_c64", n, chi_p, chi, d, runtime.ptr(pre32),
                                 runtime.ptr(lam_prev), runtime.ptr(G32), runtime.ptr(T32)))
    # BF16x6: Gamma split once per site (not timed, as in the sampler); the carried vectors are
    # split once per step (timed with the contraction)
    planes = torch.empty((ctx.lib.bsmps_bf16x6_plane_bytes(chi_p, chi, d) // 2,), dtype=torch.bfloat16, device=dev)
    ctx.call("bsmps_split_bf16x6", chi_p, chi, d, runtime.ptr(G32), runtime.ptr(lam_prev), runtime.ptr(planes))
    rows = torch.empty((ctx.lib.bsmps_bf16x6_row_plane_bytes(n, chi_p) // 2,), dtype=torch.bfloat16, device=dev)

    def x6():
        ctx.call("bsmps_split_rows_bf16x6", n, chi_p, runtime.ptr(pre32), runtime.ptr(rows))
        ctx.call("bsmps_site_contract_bf16x6", n, chi_p, chi, d, runtime.ptr(rows), runtime.ptr(planes),
                 runtime.ptr(T32))
    t6 = timed(x6)
    del planes, rows
    # FP16x3: the same with two scaled fp16 parts
    planes = torch.empty((ctx.lib.bsmps_fp16x3_plane_bytes(chi_p, chi, d) // 2,), dtype=torch.float16, device=dev)
    colscale = torch.empty(((chi * d + 63) // 64 * 64,), dtype=torch.float32, device=dev)
    ctx.call("bsmps_split_fp16x3", chi_p, chi, d, runtime.ptr(G32), runtime.ptr(lam_prev), runtime.ptr(planes),
             runtime.ptr(colscale))
    del G32
    rows = torch.empty((ctx.lib.bsmps_fp16x3_row_plane_bytes(n, chi_p) // 2,), dtype=torch.float16, device=dev)
    rowscale = torch.empty((n,), dtype=torch.float32, device=dev)

    def h3():
        ctx.call("bsmps_split_rows_fp16x3", n, chi_p, runtime.ptr(pre32), runtime.ptr(rows), runtime.ptr(rowscale))
        ctx.call("bsmps_site_contract_fp16x3", n, chi_p, chi, d, runtime.ptr(rows), runtime.ptr(rowscale),
                 runtime.ptr(planes), runtime.ptr(colscale), runtime.ptr(T32))
    t3 = timed(h3)
    del planes, rows
    W, x = sampling.displacement_basis(inst.dd)
    Wd, xd = runtime.to_device(W), runtime.to_device(x)
    alpha = (torch.randn((n, inst.M), dtype=torch.complex128, device=dev, generator=rng) * 0.3).contiguous()
    u = torch.rand((n,), dtype=torch.float64, device=dev, generator=rng)
    smp = torch.empty((n, 1), dtype=torch.int32, device=dev)
    nxt = torch.empty((n, chi), dtype=torch.complex128, device=dev)
    dwork = torch.empty((2 * n * inst.dd * d,), dtype=torch.float64, device=dev)
    ts = timed(lambda: ctx.call("bsmps_sample_site", n, chi, d, inst.dd, runtime.ptr(T), chi * d,
                                runtime.ptr(alpha), inst.M, runtime.ptr(Wd), runtime.ptr(xd), runtime.ptr(dwork),
                                runtime.ptr(lam),
                                runtime.ptr(u), runtime.ptr(smp), 1, runtime.ptr(nxt), None))
    forms = {k: tuning.get(k, "blocks") for k in ("contract_fp64", "contract_fp32")}
    f64 = (6.0 if forms["contract_fp64"] == "gauss" else 8.0) * n * chi_p * chi * d  # MFMA flops executed
    f32 = (6.0 if forms["contract_fp32"] == "gauss" else 8.0) * n * chi_p * chi * d
    useful = 8.0 * n * chi_p * chi * d  # complex-GEMM flops
    out = {"site": k, "n": n, "gamma_shape": [chi_p, chi, d], "contract_fp64_s": t64, "contract_fp32_s": t32,
           "contract_bf16x6_s": t6, "contract_fp16x3_s": t3, "sample_site_s": ts, "contract_forms": forms,
           "contract_fp64_tflops": f64 / t64 / 1e12, "contract_fp32_tflops": f32 / t32 / 1e12,
           "contract_bf16x6_complex_gemm_tflops": useful / t6 / 1e12,
           "contract_bf16x6_mfma_utilisation": 6 * useful / t6 / 1e12 / PEAK_BF16_MFMA_TFLOPS,
           "contract_fp16x3_complex_gemm_tflops": useful / t3 / 1e12,
           "contract_fp16x3_mfma_utilisation": 3 * useful / t3 / 1e12 / PEAK_BF16_MFMA_TFLOPS,
           "chain_samples_per_s_fp64": n / ((inst.M - 1) * (t64 + ts)),
           "chain_samples_per_s_fp32": n / ((inst.M - 1) * (t32 + ts)),
           "chain_samples_per_s_bf16x6": n / ((inst.M - 1) * (t6 + ts)),
           "chain_samples_per_s_fp16x3": n / ((inst.M - 1) * (t3 + ts))}
    log(f"sampling probe at site {k}: {out}")
    return out


def hafnian_only(args):
    """Batched hafnians of one size on synthetic targets drawn from a random 64 x 64 Sigma,
    timed with HIP events over ``--steps`` calls (for profiling the hafnian kernels alone; the
    library's default launch shapes, no autotuning run, so a profile holds only these calls)."""
    import torch
    from bosonsupremacy_amd import runtime
    n = args.hafnian_size
    ctx = runtime.context()
    rng = np.random.default_rng(0)
    ld = 64
    S = runtime.to_device((rng.normal(size=(ld, ld)) + 1j * rng.normal(size=(ld, ld))) * 0.3)
    count = args.hafnian_count or max(1000, int(3e10 / hafnian_flops(n // 2)))
    tg = runtime.to_device(rng.integers(0, ld, size=(count, n)).astype(np.int32))
    out = torch.empty((count,), dtype=torch.complex128, device=ctx.device)

    def run():
        ctx.call("bsmps_hafnian_targets", runtime.ptr(S), ld, runtime.ptr(tg), count, n, None, runtime.ptr(out))
    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(args.steps):
        run()
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / args.steps
    print(json.dumps({"hafnian_size": n, "count": count, "ms_per_call": ms,
                      "tflops": hafnian_flops(n // 2) * count / (ms * 1e-3) / 1e12,
                      "launch_shape": "split" if n >= 12 else "lane" if n <= 8 else "group:64"}))


PARTIAL_RUN_FLOPS = 2.7e14  # hafnian flops per library call of a partial-run site (~45 s)


def partial_run(args):
    """Part of an instance too large for one session (the M=144, chi=10^4 workload): the
    Schmidt spectra of every cut, the exact hafnian work of every site, and the listed sites
    built at full size with Gamma streamed to pinned host memory; the whole chain's
    construction time is extrapolated from the measured hafnian rate."""
    import torch
    from bosonsupremacy_amd import _lib, autotune, instances, kron, mps, runtime
    inst = instances.workload(args.workload)
    sites = [int(x) for x in args.partial_sites.split(",")]
    runtime.context()
    tuning = autotune.load_or_tune()
    log("tuned")
    t0 = time.perf_counter()
    ko = kron.kron_all(inst.sq_cov, inst.d, inst.chi)
    torch.cuda.synchronize()
    t_kron = time.perf_counter() - t0
    log(f"spectra of {inst.M - 1} cuts: {t_kron:.2f} s")
    work = site_hafnian_work_by_site(ko[0], ko[1], inst.d)
    # the learned-cutoff plan of the whole chain: what the exact discarded-weight check lets it drop
    from bosonsupremacy_amd import learned
    site_d, site_chi, plan = learned.plan_sites(learned.load_or_train(), inst.sq_cov, inst.cov, ko[0],
                                                inst.d, inst.chi)
    lwork = site_hafnian_work_by_site(ko[0], ko[1], inst.d, site_d=site_d, site_chi=site_chi)
    plan.update({"site_cutoffs": [int(x) for x in site_d], "cut_ranks": [int(x) for x in site_chi],
                 "work_by_site": [{"hafnians": int(w[0]), "flops": float(w[1]), "largest_size": int(w[2])}
                                  for w in lwork],
                 "total_flops": float(sum(w[1] for w in lwork)),
                 "plain_total_flops": float(sum(w[1] for w in work))})
    log(f"learned plan: mean cutoff {plan['mean_cutoff']:.2f}, mean rank {plan['mean_rank']:.0f}, "
        f"hafnian flops {plan['total_flops']:.3e} of {plan['plain_total_flops']:.3e}")
    form = {}
    if args.partial_learned:  # build the listed sites in the learned-cutoff form instead
        work = lwork
        form = {"site_d": site_d, "site_chi": site_chi}
    rows = []
    timer = KernelTimer()
    timer.enabled = True
    sample_probe = None
    for k in sites:
        if work[k][2] > _lib.MAX_HAFNIAN_SIZE:
            rows.append({"site": k, "hafnians": work[k][0], "largest_size": work[k][2], "built": False})
            continue
        t0 = time.perf_counter()
        # a site of more than ~45 s of hafnians runs in several calls over its right labels,
        # with a progress line after each
        n_runs = int(np.ceil(work[k][1] / PARTIAL_RUN_FLOPS))

        def on_run(site, i, n, t0=t0):
            torch.cuda.synchronize()
            log(f"site {site} run {i + 1}/{n}: {time.perf_counter() - t0:.1f} s")
        G, L, _ = mps.build_mps(inst.sq_cov, inst.d, inst.chi, kron=ko, timer=timer, to_host=True, sites=[k],
                                right_runs={k: n_runs}, on_run=on_run, **form)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        g = G[k]
        rows.append({"site": k, "hafnians": work[k][0], "flops": work[k][1], "largest_size": work[k][2],
                     "built": True, "seconds": dt, "runs": max(1, n_runs), "tflops": work[k][1] / dt / 1e12,
                     "gamma_shape": list(g.shape), "gamma_bytes": g.numel() * 16, "gamma_pinned": bool(g.is_pinned())})
        log(f"site {k}: {work[k][0]:.3e} hafnians (largest {work[k][2]}), {dt:.2f} s")
        if args.partial_sample and k == inst.M // 2:
            sample_probe = (k, g, L[k])
        del G, g
    built = [r for r in rows if r["built"]]
    rate = sum(r["flops"] for r in built) / sum(r["seconds"] for r in built)
    supported = [w for w in work if w[2] <= _lib.MAX_HAFNIAN_SIZE]
    sampling_rate = None
    if sample_probe is not None:
        sampling_rate = probe_sampling(inst, ko, sample_probe, args.partial_sample, tuning)
    out = {"workload": args.workload, "partial": True, "M": inst.M, "d": inst.d, "chi": inst.chi,
           "form": "learned" if args.partial_learned else "plain",
           "sampling_probe": sampling_rate, "learned_plan": plan,
           "kron_s": t_kron, "sites": rows,
           "measured_rate_tflops": rate / 1e12,
           "total_hafnians": int(sum(w[0] for w in work)), "total_flops": float(sum(w[1] for w in work)),
           "work_by_site": [{"hafnians": int(w[0]), "flops": float(w[1]), "largest_size": int(w[2])} for w in work],
           "sites_beyond_max_size": [k for k, w in enumerate(work) if w[2] > _lib.MAX_HAFNIAN_SIZE],
           "extrapolated_construction_s_supported_sites": t_kron + sum(w[1] for w in supported) / rate}
    print(json.dumps(out))


def site_hafnian_work(kres, knum, d, err_tol=1e-10):
    """(hafnian count, flops) of all site tensors, from the label photon counts."""
    M = len(kres) + 1
    total_n, total_f = 0, 0.0
    fl = {}
    for k in range(M):
        if k == M - 1:
            la = np.zeros(1, int)
        else:
            la = np.asarray(knum[k]).reshape(len(kres[k]), -1)[np.asarray(kres[k]) > err_tol].sum(1)
        lb = np.zeros(1, int) if k == 0 else np.asarray(knum[k - 1]).reshape(len(knum[k - 1]), -1).sum(1)
        ca, cb = np.bincount(la), np.bincount(lb)
        for j in range(d):
            for p in range(len(ca)):
                for q in range(len(cb)):
                    n = j + p + q
                    if n % 2 or ca[p] == 0 or cb[q] == 0:
                        continue
                    cnt = int(ca[p]) * int(cb[q])
                    if n // 2 not in fl:
                        fl[n // 2] = hafnian_flops(n // 2)
                    total_n += cnt
                    total_f += cnt * fl[n // 2]
    return total_n, total_f


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(workload, n_samples):
    """HBM bytes per site_contract launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    runs (profiles/<workload>_traffic.json, scaled by a 1 GiB copy), when they cover this
    workload and sample count; None otherwise."""
    path = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("n_samples") != n_samples:
        return None
    return 1024.0 * (t["fetch_kib_per_dispatch"] * t["fetch_factor"] + t["write_kib_per_dispatch"] * t["write_factor"])


class KernelTimer:
    """HIP events on the library's stream (torch's current stream) around each launch."""

    def __init__(self):
        import torch
        self.torch = torch
        self.pending = []
        self.enabled = False

    def __call__(self, name):
        timer = self

        class _Ctx:
            def __enter__(self_inner):
                if timer.enabled:
                    self_inner.a = timer.torch.cuda.Event(enable_timing=True)
                    self_inner.a.record()

            def __exit__(self_inner, *exc):
                if timer.enabled:
                    b = timer.torch.cuda.Event(enable_timing=True)
                    b.record()
                    timer.pending.append((name, self_inner.a, b))
        return _Ctx()

    def summary(self):
        self.torch.cuda.synchronize()
        out = {}
        for name, a, b in self.pending:
            out.setdefault(name, []).append(a.elapsed_time(b) * 1e-3)
        return out


def cpu_baseline(inst, kron_out, seconds=20.0):
    """The FP64 oracle (NumPy, single process) on a bounded sample of the workload:
    site tensors for a random subset of one mid-chain site's left labels (all right labels,
    all j), scaled to whole sites; and sampling of a small batch."""
    from oracle import gbs_oracle as O
    kres, knum, kS = kron_out
    M, d, chi = inst.M, inst.d, inst.chi
    k = M // 2
    res = np.asarray(kres[k])
    nL = int((res > 1e-10).sum())
    rng = np.random.default_rng(0)
    n_right = len(knum[k - 1])
    frac_rows = max(1, min(int(0.01 * nL), int(1e5 // (n_right * d))))  # ~10^5 oracle hafnians
    rows = np.sort(rng.choice(nL, size=frac_rows, replace=False))
    sub_res = list(kres)
    sub_num = list(knum)
    sub_res[k] = res[rows]
    sub_num[k] = np.asarray(knum[k]).reshape(len(res), -1)[rows]
    t0 = time.perf_counter()
    O.build_site(k, M, inst.sq_cov, sub_res, sub_num, kS, d, chi)
    t_site = (time.perf_counter() - t0) * nL / frac_rows
    sites_per_s = 1.0 / t_site
    return {
        "value": sites_per_s, "unit": "sites/s", "cores": 1, "kind": "port",
        "sample": f"oracle build_site on {frac_rows}/{nL} left labels of site {k} "
                  f"(all {len(knum[k - 1])} right labels, all j), scaled to a whole site",
    }


def cpu_sampling_rate(inst, G, L, n=50):
    from oracle import gbs_oracle as O
    from bosonsupremacy_amd import runtime, sampling
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    oG = [runtime.to_host(g) for g in G]
    rng = np.random.default_rng(0)
    normals = rng.normal(size=(2 * inst.M, n))
    uniforms = rng.uniform(size=(inst.M, n))
    t0 = time.perf_counter()
    O.sampling(oG, np.stack(L[:-1], axis=1), sqrtW, inst.dd, normals, uniforms)
    return n / (time.perf_counter() - t0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="m32_d6_chi512")
    ap.add_argument("--samples", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-cutoff form")
    ap.add_argument("--no-forms", action="store_true",
                    help="time the plain step only (no precision / learned forms, no CPU baseline)")
    ap.add_argument("--partial-sites", default=None,
                    help="comma-separated sites: build only these (Gamma streamed to pinned host) and "
                         "extrapolate the whole construction (for instances too large for one run)")
    ap.add_argument("--partial-learned", action="store_true",
                    help="with --partial-sites: build the sites in the learned-cutoff form")
    ap.add_argument("--partial-sample", type=int, default=0,
                    help="with --partial-sites including M//2: time that site's sampling step for this many samples")
    ap.add_argument("--batched", type=int, default=0,
                    help="run the batched form over this many random instances (graph-captured sampling)")
    ap.add_argument("--hafnian-size", type=int, default=0,
                    help="only time batched hafnians of this size on synthetic targets (the launch shape "
                         "the autotuner picks; for profiling the hafnian kernels alone)")
    ap.add_argument("--hafnian-count", type=int, default=0, help="hafnians per call with --hafnian-size")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device three times (scales rocprofv3 FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": x.numel() * 8, "copies": 3}))
        return
    if args.hafnian_size:
        hafnian_only(args)
        return
    if args.no_forms:
        args.no_learned = True
        args.no_cpu_baseline = True
    if args.partial_sites:
        partial_run(args)
        return
    if args.batched:
        import torch
        from bosonsupremacy_amd import autotune, batched, instances, runtime
        runtime.context()
        tuning = autotune.load_or_tune()
        insts = instances.random_batch(args.batched, seed=1000, n_samples=args.samples or 10 ** 4)
        model = None
        if not args.no_learned:
            from bosonsupremacy_amd import learned
            model = learned.load_or_train()
        out = batched.run_batch(insts, n_samples=args.samples or 10 ** 4, verbose=True, learned_model=model)
        out.update({"metric": "instances/s (construction + one graph-replayed sampling pass, FP64 or FP32 form)",
                    "workload": f"{args.batched} random instances, M~U{{16..96}}, d~U{{4..8}}, chi~U{{128..1024}}",
                    "tuning": {str(k): v for k, v in tuning.items()}})
        print(json.dumps(out))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path runs on a "
              f"single GPU with no collectives)", file=sys.stderr)
        sys.exit(2)

    import torch
    from bosonsupremacy_amd import batched, instances, kron, mps, runtime, sampling

    inst = instances.workload(args.workload)
    n_samples = args.samples or inst.n_samples
    ctx = runtime.context()
    from bosonsupremacy_amd import autotune
    tuning = autotune.load_or_tune()
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    timer = KernelTimer()

    def step():
        t0 = time.perf_counter()
        # spectra of every cut + every site tensor (host FP64 work pipelined with the GPU)
        G, L, kron_out = mps.build_mps(inst.sq_cov, inst.d, inst.chi, timer=timer)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        s = sampling.Sampler(G, L, sqrtW, inst.dd)
        normals, uniforms = sampling.random_inputs(inst.M, n_samples, seed=1)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        samples = s.run(normals, uniforms, timer=timer)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        return (t1 - t0, t3 - t2, t3 - t0), (kron_out, G, L, samples)

    for w in range(args.warmup):
        step()
        log(f"warmup step {w + 1} done")
    torch.cuda.synchronize()
    timer.enabled = True
    times = []
    last = None
    t_start = time.perf_counter()
    for i in range(args.steps):
        tt, last = step()
        times.append(tt)
        log(f"timed step {i + 1}: construction {tt[0]:.3f} s, sampling {tt[1]:.3f} s")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    t_build = float(np.mean([t[0] for t in times]))
    t_samp = float(np.mean([t[1] for t in times]))
    kern = timer.summary()

    kron_out, G, L, samples = last
    n_haf, haf_flops = site_hafnian_work(kron_out[0], kron_out[1], inst.d)
    chi, d = inst.chi, inst.d
    # per site_contract launch: MFMA flops of the contraction form in use (2x2 real blocks: 8 per
    # complex multiply-add; Gauss: three real products, 6)
    form64 = tuning.get("contract_fp64", "blocks")
    gemm_flops = (6.0 if form64 == "gauss" else 8.0) * n_samples * chi * chi * d
    kernels = {}
    for name, ds in kern.items():
        per_step = len(ds) / args.steps
        mean = float(np.mean(ds))
        entry = {"launches_per_step": per_step, "mean_ms": mean * 1e3, "total_ms_per_step": mean * per_step * 1e3}
        if name == "site_contract":
            entry["tflops"] = gemm_flops / mean / 1e12
        if name == "gamma_site":
            entry["tflops"] = haf_flops / (mean * per_step) / 1e12
            entry["hafnians_per_step"] = n_haf
        if name == "sample_site":
            # algorithmic bytes per (sample, site): read T (chi d complex), write pre (chi complex),
            # read a uniform and alpha, write the draw
            byts = n_samples * (chi * d * 16 + chi * 16 + 8 + 16 + 4)
            entry["gbs"] = byts / mean / 1e9
        kernels[name] = entry
    dom = max(kernels, key=lambda k: kernels[k]["total_ms_per_step"])
    if dom == "site_contract":
        ach = kernels[dom]["tflops"]
        roof = {"kernel": dom, "bound": "mfma", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / PEAK_FP64_TFLOPS, "traffic": measured_traffic(args.workload, n_samples),
                "algorithmic": (f"{'6' if form64 == 'gauss' else '8'}*n*chi^2*d = {gemm_flops:.3e} FP64 MFMA "
                                f"flops per launch (contraction form: {form64})"),
                "algorithmic_bytes": 16.0 * (n_samples * chi + chi * chi * d + n_samples * chi * d)}
    elif dom == "gamma_site":
        ach = kernels[dom]["tflops"]
        roof = {"kernel": dom, "bound": "fp64-vector", "achieved": ach, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": ach / PEAK_FP64_TFLOPS, "traffic": None}
    else:
        ach = kernels[dom].get("gbs", 0.0)
        roof = {"kernel": dom, "bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": ach / PEAK_HBM_GBS, "traffic": None}

    # precision forms on common random numbers: throughput and first-order marginal error
    forms = {}
    normals, uniforms = sampling.random_inputs(inst.M, n_samples, seed=2)
    rb = {}
    from bosonsupremacy_amd import metrics
    pairs = [(i, i + 1) for i in range(inst.M - 1)] + [(0, inst.M - 1), (inst.M // 4, 3 * inst.M // 4)]
    ref = {}
    for prec in (() if args.no_forms else ("fp64", "fp32", "bf16x3", "bf16x6", "fp16x3")):
        s = sampling.Sampler(G, L, sqrtW, inst.dd, precision=prec)
        s.run(normals, uniforms)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.run(normals, uniforms)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ft = KernelTimer()
        ft.enabled = True
        s.run(normals, uniforms, timer=ft)
        t_contract = sum(ft.summary().get("site_contract", [0.0]))
        smp, probs = s.run(normals, uniforms, return_probs=True)
        if prec != "fp64":
            # along the FP64 sample paths: each uniform moved to the middle of the FP64 draw's
            # cdf interval, so a draw that flips at a cdf boundary (sampling, not arithmetic)
            # does not send the two forms down different paths
            draws_differing = int((smp != ref["samples"]).any(dim=1).sum())
            smp, probs = s.run(normals, ref["mid"], return_probs=True)
            off_path = int((smp != ref["samples"]).any(dim=1).sum())
        rb[prec] = (probs * torch.arange(inst.dd, device=probs.device, dtype=probs.dtype)).sum(dim=2).mean(dim=1)
        e2 = metrics.rb_second_order(smp, probs)
        # complex multiply-adds of all site contractions; MFMA flops executed per one of them:
        # 8 (FP64 / FP32, 2x2 real blocks), 3 x 8 (BF16x3, FP16x3), 6 x 8 (BF16x6)
        cmacs = sum(n_samples * s.chis[i - 1] * s.chis[i] * s.ds[i] for i in range(1, inst.M))
        per, peak = {"fp64": (8, PEAK_FP64_TFLOPS), "fp32": (8, PEAK_FP32_MFMA_TFLOPS),
                     "bf16x3": (24, PEAK_BF16_MFMA_TFLOPS), "bf16x6": (48, PEAK_BF16_MFMA_TFLOPS),
                     "fp16x3": (24, PEAK_BF16_MFMA_TFLOPS)}[prec]  # dense FP16 MFMA peak = BF16
        forms[prec] = {"samples_per_s": n_samples / dt, "lambda_spectrum_rel_err_vs_fp64": 0.0,
                       "contract_ms_per_pass": t_contract * 1e3,
                       "contract_complex_gemm_tflops": 8 * cmacs / t_contract / 1e12,
                       "contract_mfma_utilisation": per * cmacs / t_contract / 1e12 / peak}
        log(f"sampling form {prec}: {n_samples / dt:.0f} samples/s")
        if prec == "fp64":
            ref = {"e2": e2, "samples": smp, "mid": batched.mid_uniforms(smp, probs)}
            # the FP64 MPS against the exact lossy Gaussian state (chi/d truncation + sampling noise)
            mean_x, cov_x = metrics.exact_photon_moments(inst.cov)
            e1 = rb[prec].cpu().numpy()
            cov_rb = e2 - np.triu(np.outer(e1, e1), 1)
            iu = np.triu_indices(inst.M, 1)
            forms[prec]["first_order_marginal_max_rel_err_vs_exact"] = float(np.max(np.abs(e1 - mean_x) / mean_x))
            forms[prec]["photon_covariance_max_abs_err_vs_exact"] = float(np.max(np.abs(cov_rb[iu] - cov_x[iu])))
            forms[prec]["photon_covariance_max_abs_exact"] = float(np.max(np.abs(cov_x[iu])))
        else:
            rel = torch.abs(rb[prec] - rb["fp64"]) / torch.abs(rb["fp64"])
            forms[prec]["first_order_marginal_max_rel_err_vs_fp64"] = float(rel.max())
            big = np.abs(ref["e2"]) > 1e-3 * np.abs(ref["e2"]).max()
            forms[prec]["second_order_marginal_max_rel_err_vs_fp64"] = float(
                np.max(np.abs(e2[big] - ref["e2"][big]) / np.abs(ref["e2"][big])))
            forms[prec]["pair_tvd_vs_fp64"] = metrics.pair_tvd(smp, ref["samples"], inst.dd, pairs)
            forms[prec]["draws_differing_same_uniforms"] = draws_differing
            forms[prec]["off_fp64_path"] = off_path
            forms[prec]["speedup_vs_fp64"] = forms[prec]["samples_per_s"] / forms["fp64"]["samples_per_s"]
        del probs, smp

    # learned-cutoff form (per-site cutoffs and ranks from the MLP, each checked exactly) with
    # the fastest reduced-precision contraction whose first-order marginals stayed within 1e-7 of
    # FP64 above (a tenth of the 1e-6 bar), against the plain FP64 form on the same random numbers
    learned_form = None
    if not args.no_learned:
        from bosonsupremacy_amd import learned
        t0 = time.perf_counter()
        model = learned.load_or_train()
        t_train = time.perf_counter() - t0
        log(f"learned form: network ready in {t_train:.1f} s")

        def learned_construction():
            t0 = time.perf_counter()
            ko = kron.kron_all(inst.sq_cov, inst.d, inst.chi)
            sd, sc, rep = learned.plan_sites(model, inst.sq_cov, inst.cov, ko[0], inst.d, inst.chi)
            G2, L2, _ = mps.build_mps(inst.sq_cov, inst.d, inst.chi, kron=ko, site_d=sd, site_chi=sc)
            torch.cuda.synchronize()
            return time.perf_counter() - t0, G2, L2, rep

        learned_construction()
        t_lc, G2, L2, rep = learned_construction()
        accurate = [p for p in ("fp32", "bf16x6", "fp16x3")
                    if p in forms and forms[p]["first_order_marginal_max_rel_err_vs_fp64"] <= 1e-7]
        lprec = max(accurate, key=lambda p: forms[p]["samples_per_s"]) if accurate else "fp32"
        s2 = sampling.Sampler(G2, L2, sqrtW, inst.dd, precision=lprec)
        s2.run(normals, uniforms)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s2.run(normals, uniforms)
        torch.cuda.synchronize()
        t_ls = time.perf_counter() - t0
        smp, probs = s2.run(normals, uniforms, return_probs=True)
        rb_l = (probs * torch.arange(inst.dd, device=probs.device, dtype=probs.dtype)).sum(dim=2).mean(dim=1)
        e2 = metrics.rb_second_order(smp, probs)
        big = np.abs(ref["e2"]) > 1e-3 * np.abs(ref["e2"]).max()
        rel2 = float(np.max(np.abs(e2[big] - ref["e2"][big]) / np.abs(ref["e2"][big])))
        tvd = metrics.pair_tvd(smp, ref["samples"], inst.dd, pairs)
        del probs, smp
        rel = float((torch.abs(rb_l - rb["fp64"]) / torch.abs(rb["fp64"])).max())
        t_plain_samp = n_samples / forms["fp64"]["samples_per_s"]
        learned_form = dict(rep)
        learned_form.update({
            "precision": f"{lprec} contraction",
            "training_s": t_train,
            "construction_s": t_lc,
            "construction_speedup_vs_plain": t_build / t_lc,
            "samples_per_s": n_samples / t_ls,
            "sampling_speedup_vs_plain_fp64": t_plain_samp / t_ls,
            "end_to_end_speedup_vs_plain_fp64": (t_build + t_plain_samp) / (t_lc + t_ls),
            "first_order_marginal_max_rel_err_vs_plain_fp64": rel,
            "second_order_marginal_max_rel_err_vs_plain_fp64": rel2,
            "pair_tvd_vs_plain_fp64": tvd,
        })
        del G2, L2, s2

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(inst, kron_out)
        cpu["samples_per_s"] = cpu_sampling_rate(inst, G, L)
        cpu["samples_sample"] = "oracle sampling of 50 full samples"

    sites_per_s = inst.M / t_build
    out = {
        "metric": "MPS sites/s at chi (construction), with samples/s",
        "value": sites_per_s,
        "unit": "sites/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (fp64)",
        "data": "synthetic: seeded Haar unitary, closed-form pure/classical split",
        "config": {"workload": args.workload, "M": inst.M, "d": inst.d, "chi": inst.chi, "dd": inst.dd,
                   "n_samples": n_samples, "precision_form": "fp64"},
        "samples_per_s": n_samples / t_samp,
        "construction_s": t_build,
        "sampling_s": t_samp,
        "hafnians_per_instance": n_haf,
        "roofline": roof,
        "kernels": kernels,
        "sampling_forms": forms,
        "learned_form": learned_form,
        "tuning": {str(k): v for k, v in tuning.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
