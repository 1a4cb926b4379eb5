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


def device_copy_gbs(nbytes=4 << 30, reps=5):
    """GB/s (bytes read + written) of torch's device-to-device copy of nbytes: what a plain
    streaming kernel with a half-read, half-write mix reaches on this GPU."""
    import torch
    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        y.copy_(x)
    b.record()
    torch.cuda.synchronize()
    gbs = 2 * nbytes * reps / (a.elapsed_time(b) / 1e3) / 1e9
    del x, y
    torch.cuda.empty_cache()
    return gbs


def probe_sampling(inst, ko, probe, n, tuning):
    """Time one site's sampling step at full size: the contraction of n carried vectors with
    a real Gamma_k in the FP64, FP32, BF16x6 and FP16x3 forms, prepared and contracted exactly as
    the sampler does (split into its two photon-number parity blocks when the tensor has them), and
    the displacement + draw; then extrapolate the chain's samples/s as n / ((M - 1) x per-site
    time)."""
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
    pre = pre.reshape(-1).contiguous()

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

    forms = {key: tuning.get(key, "blocks") for key in ("contract_fp64", "contract_fp32")}
    full_macs = n * chi_p * chi * d
    out = {"site": k, "n": n, "gamma_shape": [chi_p, chi, d], "contract_forms": forms}
    per_mac = {"fp64": 6.0 if forms["contract_fp64"] == "gauss" else 8.0,
               "fp32": 6.0 if forms["contract_fp32"] == "gauss" else 8.0, "bf16x6": 48.0, "fp16x3": 24.0}
    peak = {"fp64": PEAK_FP64_TFLOPS, "fp32": PEAK_FP32_MFMA_TFLOPS, "bf16x6": PEAK_BF16_MFMA_TFLOPS,
            "fp16x3": PEAK_BF16_MFMA_TFLOPS}  # dense FP16 MFMA peak = BF16
    T = None
    times = {}
    for prec in ("fp64", "fp32", "bf16x6", "fp16x3"):
        ctype = torch.complex128 if prec == "fp64" else torch.complex64
        # the site's operands are prepared once per site (not timed, as in the sampler); the
        # carried vectors' split and the parity gathers and scatters are timed with the products
        f = sampling._prepare_site(ctx, prec, k, G if prec == "fp64" else G.to(torch.complex64), chi_p, chi, lam_prev)
        work = sampling._Work(ctx, prec, n, chi_p, chi * d, ctype)
        pre_f = pre if prec == "fp64" else pre.to(torch.complex64)
        T = torch.empty((n * chi * d,), dtype=ctype, device=dev)
        t = timed(lambda: sampling.contract_site(ctx, prec, f, n, pre_f, lam_prev, T, work))
        macs = sum(n * b.rows * b.cols for b in f.blocks) if f.blocks is not None else full_macs
        times[prec] = t
        out[f"contract_{prec}_s"] = t
        out[f"contract_{prec}_parity_split"] = f.blocks is not None
        out[f"contract_{prec}_executed_macs"] = macs
        out[f"contract_{prec}_effective_complex_gemm_tflops"] = 8.0 * full_macs / t / 1e12
        out[f"contract_{prec}_mfma_utilisation"] = per_mac[prec] * macs / t / 1e12 / peak[prec]
        del f, work, pre_f
        if prec == "fp64":
            T64 = T
        torch.cuda.empty_cache()
    W, x = sampling.displacement_basis(inst.dd)
    Wd, xd = runtime.to_device(W), runtime.to_device(x)
    alpha = (torch.randn((n, inst.M), dtype=torch.complex128, device=dev, generator=rng) * 0.3).contiguous()
    u = torch.rand((n,), dtype=torch.float64, device=dev, generator=rng)
    smp = torch.empty((n, 1), dtype=torch.int32, device=dev)
    nxt = torch.empty((n, chi), dtype=torch.complex128, device=dev)
    dwork = torch.empty((2 * n * inst.dd * d,), dtype=torch.float64, device=dev)
    ts = timed(lambda: ctx.call("bsmps_sample_site", n, chi, d, inst.dd, runtime.ptr(T64), chi * d,
                                runtime.ptr(alpha), inst.M, runtime.ptr(Wd), runtime.ptr(xd), runtime.ptr(dwork),
                                runtime.ptr(lam),
                                runtime.ptr(u), runtime.ptr(smp), 1, runtime.ptr(nxt), None))
    out["sample_site_s"] = ts
    for prec in ("fp64", "fp32", "bf16x6", "fp16x3"):
        out[f"chain_samples_per_s_{prec}"] = n / ((inst.M - 1) * (times[prec] + ts))
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


def reference_hafnian_flops(size):
    """Algorithmic FP64 flops of one hafnian of the given size in the reference recursion
    (MPS_cpu.py:117-187) with every loop run in full: F(2n) = 8 C(n) + 8 2^n - 2 with
    C(n) = sum_{l<n} 2^l n(n+1)/2 (1 + (2n-2l-2)(2n-2l-3)); F(4) = 22 (three products and two
    adds), F(2) = F(0) = 0, odd sizes 0 (a complex product counts 6, an add 2, a multiply-add 8)."""
    if size % 2 or size <= 2:
        return 0
    if size == 4:
        return 22
    n = size // 2
    C = sum(2 ** l * n * (n + 1) // 2 * (1 + (2 * n - 2 * l - 2) * (2 * n - 2 * l - 3)) for l in range(n))
    return 8 * C + 8 * 2 ** n - 2


def site_hafnian_work(kres, knum, d, err_tol=1e-10, pruned=False):
    """(hafnian count, executed flops, algorithmic flops) of all site tensors, from the label
    photon counts.  Executed flops skip structurally zero and dead terms (hafnian_flops);
    algorithmic flops are the reference recursion's full loops (reference_hafnian_flops).
    ``pruned``: only the kept labels of cut k-1 enter site k (mps.build_mps(pruned=True))."""
    M = len(kres) + 1
    total_n, total_f, total_a = 0, 0.0, 0.0
    fl = {}
    for k in range(M):
        if k == M - 1:
            la = np.zeros(1, int)
        else:
            la = np.asarray(knum[k]).reshape(len(kres[k]), -1)[np.asarray(kres[k]) > err_tol].sum(1)
        if k == 0:
            lb = np.zeros(1, int)
        else:
            lb = np.asarray(knum[k - 1]).reshape(len(knum[k - 1]), -1)
            if pruned:
                lb = lb[np.asarray(kres[k - 1]) > err_tol]
            lb = lb.sum(1)
        ca, cb = np.bincount(la), np.bincount(lb)
        for j in range(d):
            for p in range(len(ca)):
                for q in range(len(cb)):
                    n = j + p + q
                    if n % 2 or ca[p] == 0 or cb[q] == 0:
                        continue
                    cnt = int(ca[p]) * int(cb[q])
                    if n // 2 not in fl:
                        fl[n // 2] = (hafnian_flops(n // 2), reference_hafnian_flops(n))
                    total_n += cnt
                    total_f += cnt * fl[n // 2][0]
                    total_a += cnt * fl[n // 2][1]
    return total_n, total_f, total_a


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profiled_traffic(workload, kernel):
    """HBM bytes per call of ``kernel`` on ``workload`` from a committed rocprofv3 summary
    (profiles/<workload>_traffic.json: FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the same
    counters on a copy of known size), or None.  Reported as coming from that profile, not
    measured in this run (counters need their own rocprofv3 pass)."""
    path = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    entry = t.get("kernels", {}).get(kernel)
    if entry is None:
        return None
    return {"bytes_per_call": entry["bytes_per_call"], "source": os.path.relpath(path, ROOT),
            "command": t.get("command")}


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


def host_record():
    """The host the CPU baseline ran on."""
    import platform
    rec = {"cpu_count": os.cpu_count(), "numpy": np.__version__, "machine": platform.machine()}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    rec["cpu_model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        rec["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        from threadpoolctl import threadpool_info
        rec["blas"] = [{"api": i.get("internal_api"), "threads": i.get("num_threads")} for i in threadpool_info()
                       if i.get("user_api") == "blas"]
    except Exception:
        pass
    return rec


def _oracle_site_rows(args):
    """Worker: the oracle's site k restricted to the given left labels (single-threaded BLAS)."""
    from oracle import gbs_oracle as O
    k, M, sq_cov, kres, knum, kS, d, chi, rows = args
    res = np.asarray(kres[k])
    sub_res, sub_num = list(kres), list(knum)
    sub_res[k] = res[rows]
    sub_num[k] = np.asarray(knum[k]).reshape(len(res), -1)[rows]
    t0 = time.perf_counter()
    O.build_site(k, M, sq_cov, sub_res, sub_num, kS, d, chi)
    return time.perf_counter() - t0


def cpu_pool(workers=16):
    """Worker processes for the all-cores CPU baseline, forked before this process initialises
    the GPU (they only ever run NumPy), started at once so the fork happens now."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    n = max(1, min(workers, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else workers))
    ex = ProcessPoolExecutor(max_workers=n, mp_context=mp.get_context("fork"), initializer=_cpu_worker_init)
    for f in [ex.submit(os.getpid) for _ in range(n)]:
        f.result()
    return ex, n


def _cpu_worker_init():
    """One BLAS thread per worker: the processes themselves use the cores."""
    try:
        from threadpoolctl import threadpool_limits
        threadpool_limits(1)
    except Exception:
        pass


def cpu_baseline(inst, kron_out, G, L, pool):
    """The FP64 oracle (this repository's NumPy restatement of the reference's CPU scripts) on
    bounded samples of the workload, in the two modes of the reference's CPU path:

    * one process, as MPS_cpu.py / sampling_cpu.py run (elementwise NumPy on one core): one
      mid-chain site restricted to a random subset of its left labels (all right labels, all j),
      scaled to whole sites, and 40 full samples;
    * all host cores: ``workers`` processes (at most the cores this process may use), each on
      its own subset of the same site's left labels with single-threaded BLAS, their union scaled
      to whole sites (sites are independent given the spectra, so this is how the CPU path
      parallelises)."""
    from oracle import gbs_oracle as O
    from bosonsupremacy_amd import runtime, sampling
    ex, pool_n = pool
    kres, knum, kS = kron_out
    M, d, chi = inst.M, inst.d, inst.chi
    k = M // 2
    nL = int((np.asarray(kres[k]) > 1e-10).sum())
    n_right = int((np.asarray(kres[k - 1]) > 1e-10).sum())
    per = max(1, min(nL // 20, int(3e5 // max(1, n_right * d))))  # ~3e5 oracle hafnian entries per worker
    rng = np.random.default_rng(0)
    picks = rng.permutation(nL)[:per * (pool_n + 1)]
    one = np.sort(picks[:per])
    t_one = _oracle_site_rows((k, M, inst.sq_cov, kres, knum, kS, d, chi, one))
    out = {"value": per / nL / t_one, "unit": "sites/s", "cores": 1, "kind": "port",
           "sample": f"oracle site {k} on {per} of its {nL} left labels (all right labels, all j), one "
                     f"process, scaled to a whole site", "host": host_record()}
    chunks = [np.sort(picks[per * (i + 1):per * (i + 2)]) for i in range(pool_n)]
    t0 = time.perf_counter()
    list(ex.map(_oracle_site_rows, [(k, M, inst.sq_cov, kres, knum, kS, d, chi, c) for c in chunks]))
    wall = time.perf_counter() - t0
    out["all_cores"] = {"value": per * pool_n / nL / wall, "unit": "sites/s", "cores": pool_n,
                        "sample": f"{pool_n} processes (one BLAS thread each) x {per} left labels of site {k}, "
                                  f"wall time of the pool, scaled to a whole site"}
    # sampling: one process (the reference's sampling_cpu.py loop)
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    oG = [runtime.to_host(g) for g in G]
    chi_max = max(g.shape[1] for g in oG)
    padded = []
    for g in oG:   # the oracle walks [chi, chi, d] tensors
        p = np.zeros((chi_max, chi_max, g.shape[2]), dtype=np.complex128)
        p[:g.shape[0], :g.shape[1]] = g
        padded.append(p)
    lam = np.zeros((chi_max, M - 1))
    for i in range(M - 1):
        lam[:len(L[i]), i] = L[i]
    n = 40
    normals = rng.normal(size=(2 * M, n))
    uniforms = rng.uniform(size=(M, n))
    t0 = time.perf_counter()
    O.sampling(padded, lam, sqrtW, inst.dd, normals, uniforms)
    out["samples_per_s"] = n / (time.perf_counter() - t0)
    out["samples_sample"] = f"oracle sampling of {n} full samples, one process"
    return out


def m144_probe(tuning, n_samples=4096):
    """The M = 144, chi = 10^4 workload at full size where one session allows it: the spectra of
    all 143 cuts, the exact hafnian work of every site (pruned plain form), the middle site built
    at full size, and one sampling step through it of ``n_samples`` samples in FP64 and FP16x3.
    Chain figures are projections from these measurements and are labelled so."""
    import torch
    from bosonsupremacy_amd import instances, kron, mps, runtime
    inst = instances.workload("m144_d10_chi10000")
    t0 = time.perf_counter()
    ko = kron.kron_all(inst.sq_cov, inst.d, inst.chi)
    torch.cuda.synchronize()
    t_kron = time.perf_counter() - t0
    k = inst.M // 2
    t0 = time.perf_counter()
    G, L, _ = mps.build_mps(inst.sq_cov, inst.d, inst.chi, kron=ko, sites=[k], pruned=True)
    torch.cuda.synchronize()
    t_site = time.perf_counter() - t0
    work = site_hafnian_work_by_site(ko[0], ko[1], inst.d,
                                     site_chi=[int((np.asarray(r) > 1e-10).sum()) for r in ko[0]])
    rate = work[k][1] / t_site
    probe = probe_sampling(inst, ko, (k, G[k], L[k]), n_samples, tuning)
    del G
    total_flops = sum(w[1] for w in work)
    proj_construction = t_kron + total_flops / rate
    return {
        "workload": "m144_d10_chi10000 (M=144, d=10, chi=10^4, dd=20)",
        "spectra_s": t_kron,
        "middle_site": {"site": k, "seconds": t_site, "hafnians": int(work[k][0]), "executed_flops": work[k][1],
                        "largest_size": int(work[k][2]), "executed_tflops": rate / 1e12},
        "sampling_step": probe,
        "extrapolated": {
            "note": "projections from the measured middle site and sampling step, not a chain run; the "
                    "hafnian-heavy sites near the chain ends run at a higher rate than the middle site",
            "construction_s_at_middle_site_rate": proj_construction,
            "samples_per_s_fp64": probe["chain_samples_per_s_fp64"],
            "samples_per_s_fp16x3": probe["chain_samples_per_s_fp16x3"]},
        "hafnian_flops_all_sites": total_flops,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="m64_d8_chi2000")
    ap.add_argument("--samples", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-cutoff form")
    ap.add_argument("--no-m144-probe", action="store_true", help="skip the M=144, chi=10^4 full-size leg")
    ap.add_argument("--no-forms", action="store_true",
                    help="time the plain step only (no precision / learned forms, no CPU baseline)")
    ap.add_argument("--partial-sites", default=None,
                    help="comma-separated sites: build only these (Gamma streamed to pinned host) and "
                         "extrapolate the whole construction (for instances too large for one run)")
    ap.add_argument("--partial-learned", action="store_true",
                    help="with --partial-sites: build the sites in the learned-cutoff form")
    ap.add_argument("--partial-sample", type=int, default=0,
                    help="with --partial-sites including M//2: time that site's sampling step for this many samples")
    ap.add_argument("--batched-eager", action="store_true",
                    help="with --batched: run each sampling pass launch by launch instead of as a HIP graph")
    ap.add_argument("--batched", type=int, default=0,
                    help="run the batched form over this many random instances (graph-captured sampling)")
    ap.add_argument("--hafnian-size", type=int, default=0,
                    help="only time batched hafnians of this size on synthetic targets (the launch shape "
                         "the autotuner picks; for profiling the hafnian kernels alone)")
    ap.add_argument("--hafnian-count", type=int, default=0, help="hafnians per call with --hafnian-size")
    ap.add_argument("--chain-sites", default=None,
                    help="FIRST-LAST: construct-and-sample sites FIRST..LAST of --workload site-major (chain.Chain), "
                         "FP64 drawing and FP16x3 replaying its draws; resumes from --chain-resume")
    ap.add_argument("--chain-samples", type=int, default=200)
    ap.add_argument("--chain-forms", default="fp64,fp16x3",
                    help="precisions walked by the chain, the first draws (e.g. fp64,fp16x3,bf16x6,fp32)")
    ap.add_argument("--chain-resume", default=None, help="checkpoint written by the previous part of the chain")
    ap.add_argument("--chain-save", default=None, help="where to write this part's checkpoint")
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
    if args.chain_sites:
        chain_run(args)
        return
    if args.no_forms:
        args.no_learned = True
        args.no_cpu_baseline = True
        args.no_m144_probe = True
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
        out = batched.run_batch_merged(insts, n_samples=args.samples or 10 ** 4, verbose=True, learned_model=model,
                                       graphs=not args.batched_eager)
        out.update({"metric": "instances/s (merged construction + one graph-replayed sampling pass per form)",
                    "workload": f"{args.batched} random instances, M~U{{16..96}}, d~U{{4..8}}, chi~U{{128..1024}}",
                    "tuning": {str(k): v for k, v in tuning.items()}})
        print(json.dumps(out))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path runs on a "
              f"single GPU with no collectives)", file=sys.stderr)
        sys.exit(2)
    standard_run(args)


# samples per step of each workload in the default bench (one batch of the workload's samples;
# the M = 64 step with all 10^5 samples would take 30 s x 25 steps)
SAMPLES_PER_STEP = {"m64_d8_chi2000": 10 ** 4}
MARGINAL_BAR = 1e-6


def form_flops(s, n):
    """Complex multiply-adds of one pass's site contractions over the live ranks of sampler s (the
    full products, zero parity blocks included)."""
    return sum(n * s.rows[i] * s.chis[i] * s.ds[i] for i in range(1, s.M))


def executed_macs(s, n):
    """Complex multiply-adds the contractions of one pass of sampler s execute: the two parity
    blocks of every site that splits, the full live block of one that does not."""
    total = 0
    for i in range(1, s.M):
        f = s.forms[i]
        if f.blocks is not None:
            total += sum(n * b.rows * b.cols for b in f.blocks)
        else:
            total += n * s.rows[i] * s.chis[i] * s.ds[i]
    return total


def chain_run(args):
    """Part of a site-major construct-and-sample run (bosonsupremacy_amd.chain): the spectra of
    every cut, the learned-cutoff plan (its ranks and cutoffs are walked as a third form when they
    differ from the pruned plain form's), then sites FIRST..LAST built at full size one at a time,
    every sample advanced through each before the next is built.  Progress per site on stderr;
    one JSON line per part with per-site seconds and the marginal errors so far."""
    import torch
    from bosonsupremacy_amd import autotune, chain, instances, kron, learned, runtime
    first, last = (int(x) for x in args.chain_sites.split("-"))
    inst = instances.workload(args.workload)
    runtime.context()
    autotune.load_or_tune()
    t0 = time.perf_counter()
    ko = kron.kron_all(inst.sq_cov, inst.d, inst.chi)
    torch.cuda.synchronize()
    t_kron = time.perf_counter() - t0
    log(f"spectra of {inst.M - 1} cuts: {t_kron:.2f} s")
    kept = [int((np.asarray(r) > 1e-10).sum()) for r in ko[0]]
    sd, sc, plan = learned.plan_sites(learned.load_or_train(), inst.sq_cov, inst.cov, ko[0], inst.d, inst.chi)
    same_plan = list(sd) == [inst.d] * inst.M and list(sc) == kept
    log(f"learned plan {'equals' if same_plan else 'differs from'} the pruned plain form "
        f"(mean cutoff {plan['mean_cutoff']:.2f}, mean rank {plan['mean_rank']:.0f})")
    if args.chain_resume:   # every part of a chain walks the forms its first part chose
        forms = chain.Chain.forms_of(args.chain_resume)
    else:
        forms = [chain.Form(p, p) for p in args.chain_forms.split(",")]
        if not same_plan and not args.no_learned:
            forms.append(chain.Form("learned_fp16x3", "fp16x3", site_chi=sc, site_d=sd))
    c = chain.Chain(inst.sq_cov, inst.cov, inst.d, inst.chi, inst.dd, args.chain_samples, seed=11, forms=forms,
                    kron=ko, chunk=args.chain_samples)
    if args.chain_resume:
        c.restore(args.chain_resume)
        log(f"resumed at site {c.next_site}")
    work = site_hafnian_work_by_site(ko[0], ko[1], inst.d, site_chi=kept)
    runs = {k: int(np.ceil(work[k][1] / PARTIAL_RUN_FLOPS)) for k in range(first, last + 1)
            if work[k][1] > PARTIAL_RUN_FLOPS}
    t_part = time.perf_counter()

    def on_run(site, i, n):
        log(f"site {site}: run {i + 1}/{n} of its right labels enqueued")

    def on_site(k, sec):
        log(f"site {k}: {sec:.2f} s ({work[k][0]:.3e} hafnians, largest {work[k][2]}), "
            f"part at {time.perf_counter() - t_part:.0f} s")
    c.run(sites=range(first, last + 1), on_site=on_site, right_runs=runs, on_run=on_run)
    if args.chain_save:
        c.save(args.chain_save)
    out = {"workload": args.workload, "chain_part": [first, last], "samples": args.chain_samples,
           "forms": [f.name for f in forms], "learned_plan_equals_pruned": same_plan,
           "learned_plan": {"mean_rank": plan["mean_rank"], "mean_cutoff": plan["mean_cutoff"],
                            "cuts_with_rank_below_kept": int(sum(a < b for a, b in zip(sc, kept)))},
           "spectra_s": t_kron,
           "site_seconds": {int(k): float(c.site_seconds[k]) for k in range(first, last + 1)},
           "site_work": {int(k): {"hafnians": int(work[k][0]), "executed_flops": float(work[k][1]),
                                  "largest_size": int(work[k][2])} for k in range(first, last + 1)},
           "part_seconds": time.perf_counter() - t_part, "next_site": c.next_site, "errors": c.errors()}
    if c.next_site == inst.M:
        out["first_order_rb_fp64"] = c.first_order(0).tolist()
    print(json.dumps(out))


def standard_run(args):
    import torch
    from bosonsupremacy_amd import autotune, instances, kron, metrics, mps, runtime, sampling

    t_begin = time.perf_counter()
    pool = None if args.no_cpu_baseline else cpu_pool()   # forked before the GPU is initialised
    inst = instances.workload(args.workload)
    n_samples = args.samples or SAMPLES_PER_STEP.get(args.workload, inst.n_samples)
    runtime.context()
    log(f"workload {args.workload}: M={inst.M} d={inst.d} chi={inst.chi}, {n_samples} samples per step")
    tuning = autotune.load_or_tune()
    log("launch tuning ready")
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    timer = KernelTimer()
    state = {}

    def step():
        t0 = time.perf_counter()
        # spectra of every cut + every site tensor over the kept labels (pruned plain form; host
        # FP64 work pipelined with the GPU)
        G, L, kron_out = mps.build_mps(inst.sq_cov, inst.d, inst.chi, timer=timer, pruned=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        s = sampling.Sampler(G, L, sqrtW, inst.dd)
        normals, uniforms = sampling.random_inputs(inst.M, n_samples, seed=1)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        samples = s.run(normals, uniforms, timer=timer)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        state.update(kron_out=kron_out, G=G, L=L, s=s)
        return t1 - t0, t3 - t2

    for w in range(args.warmup):
        step()
        log(f"warmup step {w + 1} done")
    torch.cuda.synchronize()
    timer.enabled = True
    times = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        times.append(step())
        log(f"timed step {i + 1}: construction {times[-1][0]:.3f} s, sampling {times[-1][1]:.3f} s")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    timer.enabled = False
    t_build = float(np.mean([t[0] for t in times]))
    t_samp = float(np.mean([t[1] for t in times]))
    kern = timer.summary()
    kron_out, G, L, s0 = state["kron_out"], state["G"], state["L"], state["s"]

    # ---- per-kernel rates from the HIP events of the timed steps -----------------------------
    n_haf, haf_exec, haf_alg = site_hafnian_work(kron_out[0], kron_out[1], inst.d, pruned=True)
    cmacs = form_flops(s0, n_samples)
    draw_bytes = sum(n_samples * (16 * s0.chis[i] * s0.ds[i] + (16 * s0.chis[i] if i < inst.M - 1 else 0)
                                  + 8 + 16 + 4) for i in range(inst.M))
    kernels = {}
    for name, ds in kern.items():
        per_step = len(ds) / args.steps
        mean = float(np.mean(ds))
        tot = mean * per_step
        e = {"launches_per_step": per_step, "mean_ms": mean * 1e3, "total_ms_per_step": tot * 1e3}
        if name == "site_contract":
            # the full products over the live ranks; the parity split executes the fraction below
            e["tflops"] = 8 * cmacs / tot / 1e12
            e["flops_per_step"] = 8.0 * cmacs
            e["executed_macs_fraction"] = executed_macs(s0, n_samples) / cmacs
        if name == "gamma_site":
            e.update({"hafnians_per_step": n_haf, "algorithmic_flops_per_step": haf_alg,
                      "executed_flops_per_step": haf_exec, "tflops": haf_alg / tot / 1e12,
                      "executed_tflops": haf_exec / tot / 1e12})
        if name == "sample_site":
            e["gbs"] = draw_bytes / tot / 1e9
            e["algorithmic_bytes_per_step"] = draw_bytes
        kernels[name] = e
    dom = max(kernels, key=lambda k: kernels[k]["total_ms_per_step"])
    if dom == "gamma_site":
        ach = kernels[dom]["tflops"]
        roof = {"kernel": "gamma_site (hafnians)", "bound": "fp64-vector", "achieved": ach, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": ach / PEAK_FP64_TFLOPS,
                "algorithmic": "sum over hafnians of F(size), F(2n) = 8 C(n) + 8 2^n - 2 (reference recursion, "
                               f"full loops): {haf_alg:.4e} flops per step in {kernels[dom]['launches_per_step']:.0f} "
                               "calls",
                "executed_tflops": kernels[dom]["executed_tflops"],
                "note": "bit-exact NumPy complex arithmetic (separate product, FMA and add) caps FP64 vector use at "
                        "8/12 of peak; executed flops skip structurally zero and dead polynomial terms"}
        tr = profiled_traffic(args.workload, "gamma_site")
    elif dom == "site_contract":
        ach = kernels[dom]["tflops"]
        roof = {"kernel": "site_contract", "bound": "mfma", "achieved": ach, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": ach / PEAK_FP64_TFLOPS,
                "algorithmic": f"8 n chi_(k-1) chi_k d over the live ranks: {8 * cmacs:.4e} flops per step"}
        tr = profiled_traffic(args.workload, "site_contract")
    else:
        ach = kernels[dom].get("gbs", 0.0)
        roof = {"kernel": dom, "bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": ach / PEAK_HBM_GBS}
        tr = profiled_traffic(args.workload, dom)
    roof["traffic"] = tr["bytes_per_call"] if tr else None
    if tr:
        roof["traffic_source"] = f"{tr['source']} (rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE per call, not this run)"
        rate = tr["bytes_per_call"] / (kernels[dom]["mean_ms"] / 1e3) / 1e9
        copy_gbs = device_copy_gbs()
        roof["hbm"] = {"rate_GBps": rate, "peak_GBps": PEAK_HBM_GBS, "frac_of_peak": rate / PEAK_HBM_GBS,
                       "device_copy_GBps": copy_gbs, "frac_of_device_copy": rate / copy_gbs,
                       "note": "profiled HBM bytes per call / this run's mean call time; device_copy = a 4 GiB "
                               "device-to-device copy (half read, half write) timed in this run on the same GPU"}
    log(f"timed steps done at {time.perf_counter() - t_begin:.0f} s")

    # ---- precision forms along the FP64 sample paths -----------------------------------------
    forms = {}
    normals, uniforms = sampling.random_inputs(inst.M, n_samples, seed=2)
    ref = {}
    pairs = [(i, i + 1) for i in range(inst.M - 1)] + [(0, inst.M - 1), (inst.M // 4, 3 * inst.M // 4)]
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
        gauss = {p: 6 if tuning.get("contract_" + p) == "gauss" else 8 for p in ("fp64", "fp32")}  # MFMA flops per MAC
        per, peak = {"fp64": (gauss["fp64"], PEAK_FP64_TFLOPS), "fp32": (gauss["fp32"], PEAK_FP32_MFMA_TFLOPS),
                     "bf16x3": (24, PEAK_BF16_MFMA_TFLOPS), "bf16x6": (48, PEAK_BF16_MFMA_TFLOPS),
                     "fp16x3": (24, PEAK_BF16_MFMA_TFLOPS)}[prec]  # dense FP16 MFMA peak = BF16
        emacs = executed_macs(s, n_samples)
        f = {"samples_per_s": n_samples / dt, "contract_ms_per_pass": t_contract * 1e3,
             "contract_effective_complex_gemm_tflops": 8 * cmacs / t_contract / 1e12,
             "contract_executed_macs_fraction": emacs / cmacs,
             "contract_mfma_utilisation": per * emacs / t_contract / 1e12 / peak}
        if prec == "fp64":
            e1 = metrics.rb_first_order(probs)
            ref = {"e1": e1, "e2": metrics.rb_second_order(smp, probs), "samples": smp,
                   "mid": sampling_mid(smp, probs)}
            # the FP64 MPS against the ideal lossy Gaussian state (the analysis notebooks' figures)
            smp_h = runtime.to_host(smp)
            f["vs_ideal_photon_numbers"] = {
                "first_order": metrics.correlation_summary(metrics.exact_photon_moments(inst.cov)[0], e1),
                "second_order": metrics.correlation_summary(metrics.ideal_photon_correlations(inst.cov)[1],
                                                            metrics.sample_photon_correlations(smp_h)[1])}
            f["vs_ideal_clicks"] = {
                "first_order": metrics.correlation_summary(metrics.ideal_click_correlations(inst.cov)[0],
                                                           metrics.sample_click_correlations(smp_h)[0]),
                "second_order": metrics.correlation_summary(metrics.ideal_click_correlations(inst.cov)[1],
                                                            metrics.sample_click_correlations(smp_h)[1])}
        else:
            draws_differing = int((smp != ref["samples"]).any(dim=1).sum())
            smp, probs = s.run(normals, ref["mid"], return_probs=True)
            e1 = metrics.rb_first_order(probs)
            e2 = metrics.rb_second_order(smp, probs)
            big = np.abs(ref["e2"]) > 1e-3 * np.abs(ref["e2"]).max()
            f.update({
                "first_order_marginal_max_rel_err_vs_fp64": float(np.max(np.abs(e1 - ref["e1"]) / np.abs(ref["e1"]))),
                "second_order_marginal_max_rel_err_vs_fp64": float(np.max(np.abs(e2[big] - ref["e2"][big])
                                                                          / np.abs(ref["e2"][big]))),
                "pair_tvd_vs_fp64": metrics.pair_tvd(smp, ref["samples"], inst.dd, pairs),
                "draws_differing_same_uniforms": draws_differing,
                "off_fp64_path": int((smp != ref["samples"]).any(dim=1).sum()),
                "speedup_vs_fp64": (n_samples / dt) / forms["fp64"]["samples_per_s"]})
            f["within_marginal_bar"] = bool(max(f["first_order_marginal_max_rel_err_vs_fp64"],
                                                f["second_order_marginal_max_rel_err_vs_fp64"]) <= MARGINAL_BAR)
        forms[prec] = f
        log(f"sampling form {prec}: {n_samples / dt:.0f} samples/s")
        del probs, smp, s

    # ---- learned-cutoff form against the pruned plain FP64 form --------------------------------
    learned_form = None
    if not args.no_learned:
        learned_form = learned_leg(inst, sqrtW, n_samples, normals, ref, forms, t_build, t_samp)
    log(f"forms done at {time.perf_counter() - t_begin:.0f} s")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(inst, kron_out, G, L, pool)
        pool[0].shutdown()
        log(f"cpu baseline done at {time.perf_counter() - t_begin:.0f} s")
    del G, L, s0
    m144 = None
    if not args.no_m144_probe:
        if time.perf_counter() - t_begin < 400:
            m144 = m144_probe(tuning)
        else:
            m144 = {"skipped": "the default run had used 400 s before this leg"}
        log(f"M=144 probe done at {time.perf_counter() - t_begin:.0f} s")

    out = {
        "metric": "MPS sites/s at chi (construction), with samples/s",
        "value": inst.M / t_build,
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
                   "samples_per_step": n_samples, "precision_form": "fp64",
                   "construction": "pruned plain method (kept labels of both cuts)"},
        "samples_per_s": n_samples / t_samp,
        "construction_s": t_build,
        "sampling_s": t_samp,
        "hafnians_per_instance": n_haf,
        "roofline": roof,
        "kernels": kernels,
        "sampling_forms": forms,
        "learned_form": learned_form,
        "m144_probe": m144,
        "tuning": {str(k): v for k, v in tuning.items()},
        "cpu_baseline": cpu,
        "bench_wall_s": time.perf_counter() - t_begin,
    }
    print(json.dumps(out))


def sampling_mid(samples, probs):
    from bosonsupremacy_amd import batched
    return batched.mid_uniforms(samples, probs)


def learned_leg(inst, sqrtW, n_samples, normals, ref, forms, t_build, t_samp):
    """The learned-cutoff form (per-site cutoffs and per-cut ranks predicted by the MLP, each
    judged by the exact discarded-weight bound) with the fastest reduced-precision contraction
    whose marginals stayed within a tenth of the bar, against the pruned plain FP64 form.  The
    network's own saving is the hafnian and contraction work of its plan below the pruned plan's
    (the pruned form already keeps every label of Schmidt weight above 1e-10 and nothing else).

    It is run twice: at the default tolerances (1e-10 of Schmidt weight per cut, the most a
    truncation may drop when the marginals must stay within 1e-6, since a truncation moves the
    sampled conditionals by the square root of the weight it drops) and at the tolerance the
    network is trained for (1e-7), where it has work to remove and costs ~1e-5 in the marginals."""
    import torch
    from bosonsupremacy_amd import kron, learned, metrics, mps, sampling
    t0 = time.perf_counter()
    model = learned.load_or_train()
    t_train = time.perf_counter() - t0
    log(f"learned form: network ready in {t_train:.1f} s")
    accurate = [p for p in ("fp32", "bf16x6", "fp16x3")
                if p in forms and forms[p]["first_order_marginal_max_rel_err_vs_fp64"] <= 0.1 * MARGINAL_BAR
                and forms[p]["second_order_marginal_max_rel_err_vs_fp64"] <= MARGINAL_BAR]
    lprec = max(accurate, key=lambda p: forms[p]["samples_per_s"]) if accurate else "fp32"
    big = np.abs(ref["e2"]) > 1e-3 * np.abs(ref["e2"]).max()

    def construction(tau):
        t0 = time.perf_counter()
        ko = kron.kron_all(inst.sq_cov, inst.d, inst.chi)
        sd, sc, rep = learned.plan_sites(model, inst.sq_cov, inst.cov, ko[0], inst.d, inst.chi, tau_chi=tau)
        G2, L2, _ = mps.build_mps(inst.sq_cov, inst.d, inst.chi, kron=ko, site_d=sd, site_chi=sc)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, G2, L2, rep, ko, sd, sc

    def leg(tau, warm):
        if warm:
            construction(tau)
        t_lc, G2, L2, rep, ko, sd, sc = construction(tau)
        kept = [int((np.asarray(r) > 1e-10).sum()) for r in ko[0]]
        w_plain = site_hafnian_work_by_site(ko[0], ko[1], inst.d, site_chi=kept)
        w_learn = site_hafnian_work_by_site(ko[0], ko[1], inst.d, site_d=sd, site_chi=sc)
        s2 = sampling.Sampler(G2, L2, sqrtW, inst.dd, precision=lprec)
        s2.run(normals, ref["mid"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s2.run(normals, ref["mid"])
        torch.cuda.synchronize()
        t_ls = time.perf_counter() - t0
        smp, probs = s2.run(normals, ref["mid"], return_probs=True)
        e1 = metrics.rb_first_order(probs)
        e2 = metrics.rb_second_order(smp, probs)
        out = dict(rep)
        plain_gemm = sum(kept[i - 1] * kept[i] * inst.d for i in range(1, inst.M - 1)) + kept[-1] * inst.d
        learn_gemm = sum(sc[i - 1] * sc[i] * sd[i] for i in range(1, inst.M - 1)) + sc[-1] * sd[-1]
        r1 = float(np.max(np.abs(e1 - ref["e1"]) / np.abs(ref["e1"])))
        r2 = float(np.max(np.abs(e2[big] - ref["e2"][big]) / np.abs(ref["e2"][big])))
        out.update({
            "tau_schmidt_weight_per_cut": tau,
            "construction_s": t_lc,
            "construction_speedup_vs_pruned_plain": t_build / t_lc,
            "network_saving": {
                "hafnian_flops_pruned_plain": float(sum(w[1] for w in w_plain)),
                "hafnian_flops_learned": float(sum(w[1] for w in w_learn)),
                "hafnian_work_removed_frac": 1 - sum(w[1] for w in w_learn) / max(1.0, sum(w[1] for w in w_plain)),
                "contraction_work_removed_frac": 1 - learn_gemm / plain_gemm,
                "cuts_with_rank_below_kept": int(sum(1 for a, b in zip(sc, kept) if a < b)),
                "sites_with_cutoff_below_d": int(sum(1 for x in sd if x < inst.d))},
            "samples_per_s": n_samples / t_ls,
            "sampling_speedup_vs_plain_fp64": t_samp / t_ls,
            "end_to_end_speedup_vs_plain_fp64": (t_build + t_samp) / (t_lc + t_ls),
            "first_order_marginal_max_rel_err_vs_plain_fp64": r1,
            "second_order_marginal_max_rel_err_vs_plain_fp64": r2,
            "within_bar": bool(max(r1, r2) <= MARGINAL_BAR),
            "off_fp64_path": int((smp != ref["samples"]).any(dim=1).sum()),
        })
        log(f"learned form at tau {tau:g}: hafnian work removed "
            f"{out['network_saving']['hafnian_work_removed_frac']:.3f}, marginals {r1:.2e} / {r2:.2e}")
        del G2, L2, s2, probs, smp
        return out

    out = {"precision": f"{lprec} contraction", "training_s": t_train,
           "at_bar": leg(learned.TAU_CHI, True),
           "at_training_tolerance": leg(learned.TAU_CHI_TRAIN, False)}
    return out


if __name__ == "__main__":
    main()
