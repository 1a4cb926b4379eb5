"""Site tensors Gamma_k / Lambda_k (L3) -- drop-in for MPS_cpu.py.

Per site the host computes the site's Gaussian-unitary kernel Sigma and normalisation Z
(MPS_cpu.py:395, 481-483) and a few small per-label tables (targets, denominators, photon
classes); every Gamma entry -- target assembly, gather, hafnian and epilogue -- is computed
on the GPU by ``bsmps_gamma_site``.  The reference functions are available with their
signatures: ``Sigma_select``, ``hafnian``, ``A_elem`` (GPU-backed), ``get_target`` and
``push_to_end`` (host integer helpers).

Differences from the reference, by design: complex128 Gamma and float64 Lambda (the
reference stores complex64/float32, MPS_cpu.py:397-398).

CLI (same flags and files as the reference)::

    python -m bosonsupremacy_amd.mps --d D --chi CHI --dir ROOT/
"""
import argparse
import contextlib
import ctypes
import os
import time

import numpy as np
import torch
from scipy.special import factorial

from . import _lib, gaussian, hostpool, runtime

ERR_TOL = 1e-10  # MPS_cpu.py:390


# ---------------------------------------------------------------------------------------
# Host integer helpers (MPS_cpu.py:39-47, 314-332)
# ---------------------------------------------------------------------------------------

def get_target(num):
    """Occupation rows [B, L] -> 1-based repeated mode indices [B, max photons], zero padded
    (MPS_cpu.py:314-332)."""
    num = np.asarray(num)
    B, L = num.shape
    counts = num.sum(axis=1).astype(np.int64)
    width = int(counts.max()) if B else 0
    target = np.zeros((B, width), dtype=np.int32)
    if B == 0 or width == 0:
        return target
    flat_counts = num.astype(np.int64).ravel()
    values = np.repeat(np.tile(np.arange(1, L + 1, dtype=np.int32), B), flat_counts)
    rows = np.repeat(np.arange(B), counts)
    starts = np.repeat(np.cumsum(counts) - counts, counts)
    target[rows, np.arange(len(rows)) - starts] = values
    return target


def push_to_end(array):
    """Right-align the non-zero prefix of each row (MPS_cpu.py:39-47 on get_target output)."""
    array = np.asarray(array)
    B, n = array.shape
    counts = (array != 0).sum(axis=1)
    out = np.zeros_like(array)
    rows, cols = np.nonzero(array)
    out[rows, cols + (n - counts)[rows]] = array[rows, cols]
    return out


# ---------------------------------------------------------------------------------------
# GPU-backed reference functions (MPS_cpu.py:19-37, 117-140, 334-355)
# ---------------------------------------------------------------------------------------

def _as_complex_dev(a):
    return runtime.to_device(np.asarray(a, dtype=np.complex128) if not isinstance(a, torch.Tensor) else a,
                             dtype=torch.complex128)


def Sigma_select(Sigma, target):
    """out[l, r, c] = Sigma[target[l, r], target[l, c]] (direct_mps_kernels.cu:22-39)."""
    ctx = runtime.context()
    S = _as_complex_dev(Sigma)
    T = runtime.to_device(np.asarray(target, dtype=np.int32))
    B, n = T.shape
    out = torch.empty((B, n, n), dtype=torch.complex128, device=ctx.device)
    ctx.call("bsmps_sigma_select", runtime.ptr(S), S.shape[0], runtime.ptr(T), B, n, runtime.ptr(out))
    return runtime.to_host(out)


def hafnian(A, loops=None):
    """Batched hafnian of A [B, n, n] (MPS_cpu.py:117-187), complex128.

    ``loops`` (optional, [B, n] or [n]): loop weights, for the loop hafnian (a vertex may be
    matched to itself); the reference has none, and None is its plain hafnian."""
    ctx = runtime.context()
    Ad = _as_complex_dev(A)
    B, n, n2 = Ad.shape
    if n != n2:
        raise ValueError("Matrix must be square")
    out = torch.empty((B,), dtype=torch.complex128, device=ctx.device)
    if loops is None:
        ctx.call("bsmps_hafnian_dense", runtime.ptr(Ad), B, n, runtime.ptr(out))
    else:
        Ld = _as_complex_dev(np.broadcast_to(runtime.host(loops).astype(np.complex128), (B, n)).copy())
        ctx.call("bsmps_loop_hafnian_dense", runtime.ptr(Ad), runtime.ptr(Ld), B, n, runtime.ptr(out))
    return runtime.to_host(out)


def A_elem(Sigma, target, denominator, max_memory_in_gb=0.5, loops=None):
    """haf(Sigma[t][:, t]) / denominator for each target row (MPS_cpu.py:334-355).

    Returns (values, haf_time, sigma_time) like the reference; the gather is fused into the
    hafnian kernel, so sigma_time is 0.  ``max_memory_in_gb`` is accepted for signature
    compatibility; the GPU path never materialises the gathered matrices.  ``loops``
    (optional, one weight per row of Sigma, gathered like Sigma): the loop hafnian instead."""
    ctx = runtime.context()
    S = _as_complex_dev(Sigma)
    T = runtime.to_device(np.ascontiguousarray(target, dtype=np.int32))
    den = runtime.to_device(np.ascontiguousarray(np.broadcast_to(denominator, (T.shape[0],)), dtype=np.float64))
    out = torch.empty((T.shape[0],), dtype=torch.complex128, device=ctx.device)
    t0 = time.time()
    if loops is None:
        ctx.call("bsmps_hafnian_targets", runtime.ptr(S), S.shape[0], runtime.ptr(T), T.shape[0], T.shape[1],
                 runtime.ptr(den), runtime.ptr(out))
    else:
        Ld = _as_complex_dev(loops)
        if Ld.shape != (S.shape[0],):
            raise ValueError("loops must hold one weight per row of Sigma")
        ctx.call("bsmps_loop_hafnian_targets", runtime.ptr(S), S.shape[0], runtime.ptr(Ld), runtime.ptr(T),
                 T.shape[0], T.shape[1], runtime.ptr(den), runtime.ptr(out))
    res = runtime.to_host(out)
    return res, time.time() - t0, 0.0


# ---------------------------------------------------------------------------------------
# Site tensors
# ---------------------------------------------------------------------------------------

class _Labels:
    """Per-label tables of one side of a site: Sigma targets, denominators, photon classes."""

    def __init__(self, num, offset):
        num = np.asarray(num)
        num = num.reshape(num.shape[0], -1)
        self.count = num.sum(axis=1).astype(np.int32)
        tg = get_target(num)
        if offset:
            tg = np.where(tg != 0, tg + offset, 0).astype(np.int32)
        self.targets = tg
        # factorial of each occupation through a table of the same scipy values
        table = np.array(factorial(np.arange(int(num.max(initial=0)) + 1)), dtype=np.float64)
        self.den = np.sqrt(np.prod(table[num], axis=1)).astype(np.float64)
        self.perm = np.argsort(self.count, kind='stable').astype(np.int32)
        self.class_count = np.bincount(self.count).astype(np.int32) if len(self.count) else np.zeros(1, np.int32)

    @classmethod
    def empty(cls):
        self = cls.__new__(cls)
        self.count = np.zeros(1, np.int32)
        self.targets = np.zeros((1, 0), np.int32)
        self.den = np.ones(1)
        self.perm = np.zeros(1, np.int32)
        self.class_count = np.ones(1, np.int32)
        return self


class SiteTables:
    """Host tables of one site tensor: kind, extents, Lambda and the per-label tables of both
    sides.  Built by ``site_tables`` with NumPy only, so a host worker process can build them."""

    __slots__ = ("k", "kind", "n_a", "n_b", "Lambda", "lam", "left", "right", "sqrt_fact")


def site_tables(k, M, d, chi, kron_res, kron_num, err_tol=ERR_TOL, ranks=None):
    """SiteTables of site k (MPS_cpu.py:402-520: which labels of cuts k-1 and k enter, their
    targets and denominators, Lambda_k).  ``ranks`` = (rank of cut k-1, rank of cut k) keeps
    only the leading Schmidt vectors (learned-cutoff form)."""
    r_prev, r_cur = ranks if ranks is not None else (None, None)
    t = SiteTables()
    t.k = k
    if k == M - 1:
        t.kind = _lib.SITE_LAST
        t.left = _Labels.empty()
        t.lam = np.ones(1)
        t.n_a = chi if ranks is None else 1
        t.Lambda = np.zeros(t.n_a)
    else:
        t.kind = _lib.SITE_FIRST if k == 0 else _lib.SITE_MIDDLE
        res = np.asarray(kron_res[k], dtype=np.float64)
        keep = res > err_tol
        num_k = np.asarray(kron_num[k]).reshape(len(res), -1)[keep]
        lam = np.sqrt(res[keep])
        if r_cur is not None:
            num_k, lam = num_k[:r_cur], lam[:r_cur]
        t.left = _Labels(num_k, 0)
        t.lam = lam
        t.n_a = chi if ranks is None else len(lam)
        t.Lambda = np.zeros(t.n_a)
        t.Lambda[:len(lam)] = lam
    if k == 0:
        t.right = _Labels.empty()
    else:
        num_pre = np.asarray(kron_num[k - 1])
        num_pre = num_pre.reshape(num_pre.shape[0], -1)
        if r_prev is not None:
            num_pre = num_pre[:r_prev]
        t.right = _Labels(num_pre, M - k - 1)  # Sigma of site k >= 1 has 2 (M - k) modes
    t.n_b = chi if ranks is None else len(t.right.count)
    t.sqrt_fact = np.array([np.sqrt(factorial(j)) for j in np.arange(d)], dtype=np.float64)
    return t


def right_label_runs(t, d, n_runs):
    """Split the right labels of SiteTables ``t`` (in photon-class order, the order of
    ``t.right.perm``) into at most ``n_runs`` contiguous runs of about equal hafnian work, as
    [(first, end)] positions in that order.  The work of one right label of class q is the sum,
    over photon numbers j < d and left classes p with j + p + q even, of the left class size
    times m^3 2^m for the hafnian of size 2m = j + p + q."""
    rcls = t.right.count[t.right.perm]
    n = len(rcls)
    if n_runs <= 1 or n <= 1:
        return [(0, n)]
    lcc = np.asarray(t.left.class_count, dtype=np.float64)
    per_class = np.zeros(len(t.right.class_count))
    for q in range(len(per_class)):
        for j in range(d):
            for p in range(len(lcc)):
                s = j + p + q
                if s % 2 == 0 and lcc[p]:
                    m = s // 2
                    per_class[q] += lcc[p] * max(1, m ** 3) * 2.0 ** m
    cum = np.cumsum(per_class[rcls])
    cuts = np.searchsorted(cum, cum[-1] * np.arange(1, n_runs) / n_runs, side="right")
    edges = np.unique(np.concatenate([[0], np.clip(cuts, 1, n - 1), [n]]))
    return [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:])]


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a if a.size else np.zeros(1, np.int32)


class SiteCall:
    """One site tensor's library call: its ``bsmps_gamma_site_args``, the zeroed output Gamma
    (device, [n_b][n_a][d]), Lambda_k, and the host and device tables the arguments point to
    (kept alive with the call)."""

    __slots__ = ("k", "args", "gamma", "Lambda", "tables", "uploads", "lcc", "rcc")


def site_call(k, M, sq_cov, kron_res, kron_num, kron_S, d, chi, S_full=None, err_tol=ERR_TOL, sigma_z=None,
              ranks=None, tables=None):
    """The SiteCall of site k (MPS_cpu.py:402-520): Sigma and Z of the site (host FP64), its label
    tables uploaded in one pinned copy, and Gamma allocated on the device.  Arguments as for
    site_tensor; nothing runs on the GPU yet."""
    ctx = runtime.context()
    dev = ctx.device
    Sigma, Z = sigma_z if sigma_z is not None else gaussian.site_kernel(k, M, sq_cov, kron_S, S_full)
    t = tables if tables is not None else site_tables(k, M, d, chi, kron_res, kron_num, err_tol, ranks)
    Mp = Sigma.shape[0] // 2
    if k > 0 and Mp != M - k:
        raise ValueError(f"site {k}: Sigma has {Mp} modes, expected {M - k}")
    left, right = t.left, t.right
    # every table of the site in one pinned staging buffer and one asynchronous copy, which
    # overlaps the previous site's kernels
    up = runtime.upload_many(
        [np.ascontiguousarray(Sigma, dtype=np.complex128), t.sqrt_fact, _i32(left.targets), _i32(left.perm),
         _i32(right.targets), _i32(right.perm), left.den, right.den, t.lam])
    (sig_d, sf_d, lt_d, lp_d, rt_d, rp_d, lden_d, rden_d, lam_d) = up
    c = SiteCall()
    c.k, c.tables, c.uploads, c.Lambda = k, t, up, t.Lambda
    c.lcc = np.ascontiguousarray(left.class_count, dtype=np.int32)
    c.rcc = np.ascontiguousarray(right.class_count, dtype=np.int32)
    c.args = _lib.GammaSiteArgs(
        kind=t.kind, d=d, chi=t.n_a, ld=2 * Mp, sigma=sig_d.data_ptr(), z=Z, inv_z=1.0 / Z,
        sqrt_fact=sf_d.data_ptr(),
        n_left=len(left.count), w_left=left.targets.shape[1], left_targets=lt_d.data_ptr(),
        left_den=lden_d.data_ptr(), lambda_=lam_d.data_ptr(), left_perm=lp_d.data_ptr(),
        left_class_count_host=c.lcc.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n_left_class=len(c.lcc),
        n_right=len(right.count), w_right=right.targets.shape[1], right_targets=rt_d.data_ptr(),
        right_den=rden_d.data_ptr(), right_perm=rp_d.data_ptr(),
        right_class_count_host=c.rcc.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n_right_class=len(c.rcc))
    c.gamma = torch.zeros((t.n_b, t.n_a, d), dtype=torch.complex128, device=dev)
    return c


def gamma_sites(calls, timer=None):
    """Enqueue the site tensors of ``calls`` (SiteCalls of one instance or of several) in one
    ``bsmps_gamma_sites`` call: one hafnian launch per photon total covers every listed site."""
    calls = list(calls)
    if not calls:
        return
    n = len(calls)
    arr = (_lib.GammaSiteArgs * n)(*[c.args for c in calls])
    outs = (ctypes.c_void_p * n)(*[runtime.ptr(c.gamma) for c in calls])
    with (timer("gamma_sites") if timer else contextlib.nullcontext()):
        runtime.context().call("bsmps_gamma_sites", n, arr, outs)


def site_tensor(k, M, sq_cov, kron_res, kron_num, kron_S, d, chi, S_full=None, err_tol=ERR_TOL, timer=None,
                sigma_z=None, ranks=None, tables=None, right_runs=1, on_run=None):
    """Gamma_k as a device tensor [chi, chi, d] complex128 and Lambda_k (float64 [chi]).

    Follows the three branches of MPS_cpu.py:402-520; Lambda is zero for the last site.
    ``sigma_z`` = (Sigma, Z) of the site when already computed (gaussian.site_kernel);
    ``tables`` = its SiteTables when already built (site_tables).

    ``right_runs`` > 1 computes the site in that many library calls, each over a run of the
    right labels (right_label_runs), and calls ``on_run(i, n_runs)`` after enqueuing run i --
    for progress reports on sites that take minutes.  Every entry is the same hafnian either
    way, so the result is bit-identical.

    ``ranks`` = (rank of cut k-1, rank of cut k) truncates the bond labels to the leading
    Schmidt vectors (the learned-cutoff form; ``d`` is then this site's own cutoff): Gamma is
    [rank_{k-1}][rank_k][d] and Lambda_k has rank_k entries (one row / one column at the
    chain ends)."""
    ctx = runtime.context()
    c = site_call(k, M, sq_cov, kron_res, kron_num, kron_S, d, chi, S_full, err_tol, sigma_z, ranks, tables)
    right = c.tables.right
    args = c.args
    rp_ptr = args.right_perm
    runs = right_label_runs(c.tables, d, right_runs)
    for i, (r0, r1) in enumerate(runs):
        if len(runs) > 1:  # this run's labels: a slice of the class-ordered permutation
            rcc = np.bincount(right.count[right.perm[r0:r1]], minlength=len(right.class_count)).astype(np.int32)
            args.n_right = r1 - r0
            args.right_perm = rp_ptr + 4 * r0
            args.right_class_count_host = rcc.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        with (timer("gamma_site") if timer else contextlib.nullcontext()):
            ctx.call("bsmps_gamma_site", ctypes.byref(args), runtime.ptr(c.gamma))
        if on_run is not None:
            on_run(i, len(runs))
    return c.gamma, c.Lambda


class HostStream:
    """Streams device tensors to pinned host memory on a copy stream (default priority),
    overlapping the compute stream: each copy starts when the compute stream has produced the
    tensor, the device memory is released to the allocator once its copy has run, and at most
    ``depth`` copies are in flight (which bounds the device memory held by finished tensors)."""

    def __init__(self, depth=2):
        self.stream = torch.cuda.Stream()
        self.depth = depth
        self.pending = []

    def push(self, t):
        ready = torch.cuda.Event()
        ready.record()
        host = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
        with torch.cuda.stream(self.stream):
            self.stream.wait_event(ready)
            host.copy_(t, non_blocking=True)
            done = torch.cuda.Event()
            done.record(self.stream)
        t.record_stream(self.stream)
        self.pending.append(done)
        while len(self.pending) > self.depth:
            self.pending.pop(0).synchronize()
        return host

    def drain(self):
        for ev in self.pending:
            ev.synchronize()
        self.pending = []


def build_mps(sq_cov, d, chi, kron=None, max_dim=10 ** 5, timer=None, site_d=None, site_chi=None,
              to_host=False, sites=None, right_runs=None, on_run=None, pruned=False, consume=None):
    """All site tensors: (list of device Gamma_k, list of host Lambda_k), plus the kron
    outputs used.  ``kron`` = (res, num, S) lists as kron.kron_all returns them; when it is
    None the spectra are computed too.

    ``pruned`` builds the plain method without the work a zero Lambda multiplies away: each
    Gamma_k is [r_{k-1}][r_k][d] over the kept labels of both cuts (Schmidt weight > 1e-10),
    where the reference (MPS_cpu.py:461-520) fills [chi][chi][d] and computes the rows of every
    label of cut k-1, kept or not.  Its entries are the same values as the plain tensor's live
    block.
    ``site_d`` (per-site photon cutoffs <= d) and ``site_chi`` (per-cut ranks) select the
    learned-cutoff form (bosonsupremacy_amd.learned.plan_sites); neither, and not ``pruned``,
    is the plain method.
    ``to_host`` streams every Gamma_k to pinned host memory as soon as it is computed (the
    returned Gammas are then pinned host tensors and the device copies are released);
    ``sites`` builds only the listed sites (None entries elsewhere).
    ``right_runs`` = {site: number of runs} computes those sites in runs of right labels,
    calling ``on_run(site, i, n_runs)`` after each (site_tensor).
    ``consume(k, Gamma_k, Lambda_k)`` is called as soon as each site is built; the site tensor
    is then not kept (its list entry is None), so only one is held at a time -- the site-major
    construct-and-sample pipeline (chain.Chain).

    Pipelined: the host FP64 work of every site (Bloch-Messiah, Sigma) is queued on the
    host worker pool as soon as the cut symplectics are known -- before the Schmidt spectra
    run on the GPU -- and each site's hafnian launch is enqueued, asynchronously, as soon as
    its Sigma and label tables are ready, so host and GPU work overlap."""
    from . import kron as kron_mod
    M = len(sq_cov) // 2
    S_full_f = hostpool.williamson_S(sq_cov)
    if kron is None:
        cuts = hostpool.cut_inputs_all(sq_cov, d)
        kS = [S for _, S in cuts]
        site_f = hostpool.submit_sites(sq_cov, kS, S_full_f.result())
        spec = kron_mod.kron_cuts([tp for tp, _ in cuts], chi, max_dim, timer=timer)
        kron = ([r for r, _ in spec], [n for _, n in spec], kS)
    else:
        site_f = hostpool.submit_sites(sq_cov, kron[2], S_full_f.result())
    kres, knum, kS = kron
    Gammas, Lambdas = [], []
    shaped = site_d is not None or site_chi is not None or pruned
    if shaped:
        site_d = list(site_d) if site_d is not None else [d] * M
        if site_chi is None:
            site_chi = [int((np.asarray(r) > ERR_TOL).sum()) for r in kres]
    streamer = HostStream() if to_host else None
    wanted = set(range(M)) if sites is None else set(sites)
    # the label tables of every site are built by the host workers while the GPU runs the
    # previous sites (a site whose tables are not ready yet builds them here instead)
    tab_f = hostpool.submit_tables(M, d, chi, kres, knum, ERR_TOL, site_d if shaped else None,
                                   site_chi if shaped else None) if sites is None else None
    for k in range(M):
        if k not in wanted:
            Gammas.append(None)
            Lambdas.append(None)
            continue
        tables = tab_f[k].result_or_run() if tab_f is not None else None
        runs = {}
        if right_runs and right_runs.get(k, 1) > 1:
            runs = {"right_runs": right_runs[k],
                    "on_run": (lambda i, n, k=k: on_run(k, i, n)) if on_run is not None else None}
        if shaped:
            ranks = (site_chi[k - 1] if k > 0 else None, site_chi[k] if k < M - 1 else None)
            G, L = site_tensor(k, M, sq_cov, kres, knum, kS, site_d[k], chi, timer=timer,
                               sigma_z=site_f[k].result(), ranks=ranks, tables=tables, **runs)
        else:
            G, L = site_tensor(k, M, sq_cov, kres, knum, kS, d, chi, timer=timer, sigma_z=site_f[k].result(),
                               tables=tables, **runs)
        if streamer is not None:
            G = streamer.push(G)
        if consume is not None:
            consume(k, G, L)
            G = None
        Gammas.append(G)
        Lambdas.append(L)
    if streamer is not None:
        streamer.drain()
    return Gammas, Lambdas, kron


def site_calls(sq_cov, d, chi, kron=None, max_dim=10 ** 5, pruned=False, site_d=None, site_chi=None):
    """The SiteCalls of every site of one instance, ready for ``gamma_sites`` (alone or merged
    with other instances'): the Schmidt spectra run on the GPU (unless ``kron`` is given), the
    host FP64 work and the label tables in the worker pool, and each site's tables are uploaded;
    no hafnian runs yet.  ``pruned`` / ``site_d`` / ``site_chi`` as in build_mps.  Returns
    (calls, kron)."""
    from . import kron as kron_mod
    M = len(sq_cov) // 2
    S_full_f = hostpool.williamson_S(sq_cov)
    if kron is None:
        cuts = hostpool.cut_inputs_all(sq_cov, d)
        kS = [S for _, S in cuts]
        site_f = hostpool.submit_sites(sq_cov, kS, S_full_f.result())
        spec = kron_mod.kron_cuts([tp for tp, _ in cuts], chi, max_dim)
        kron = ([r for r, _ in spec], [n for _, n in spec], kS)
    else:
        site_f = hostpool.submit_sites(sq_cov, kron[2], S_full_f.result())
    kres, knum, kS = kron
    shaped = site_d is not None or site_chi is not None or pruned
    if shaped:
        site_d = list(site_d) if site_d is not None else [d] * M
        if site_chi is None:
            site_chi = [int((np.asarray(r) > ERR_TOL).sum()) for r in kres]
    tab_f = hostpool.submit_tables(M, d, chi, kres, knum, ERR_TOL, site_d if shaped else None,
                                   site_chi if shaped else None)
    calls = []
    for k in range(M):
        ranks = (site_chi[k - 1] if k > 0 else None, site_chi[k] if k < M - 1 else None) if shaped else None
        calls.append(site_call(k, M, sq_cov, kres, knum, kS, site_d[k] if shaped else d, chi,
                               sigma_z=site_f[k].result(), ranks=ranks, tables=tab_f[k].result_or_run()))
    return calls, kron


def _save_atomic(path, arr):
    """np.save through a temporary name and a rename, so a run killed mid-write never leaves a
    truncated file under the final name (resume trusts every file it finds)."""
    tmp = path + ".partial.npy"
    np.save(tmp, arr)
    os.replace(tmp, path)


def first_missing_site(path, M):
    """The first site k whose Gamma_k.npy (or, for k < M-1, Lambda_k.npy) is not in ``path``:
    where a resumed run starts (MPS_cpu.py:523-525 saves both after each site)."""
    for k in range(M):
        if not os.path.exists(path + f"Gamma_{k}.npy"):
            return k
        if k < M - 1 and not os.path.exists(path + f"Lambda_{k}.npy"):
            return k
    return M


def main(argv=None):
    parser = argparse.ArgumentParser(description="MPS site tensors (drop-in for MPS_cpu.py)")
    parser.add_argument('--d', type=int, required=True)
    parser.add_argument('--chi', type=int, required=True)
    parser.add_argument('--dir', type=str, required=True)
    parser.add_argument('--restart', action='store_true',
                        help='recompute every site (default: resume at the first site without its files)')
    parser.add_argument('--reference-dtypes', action='store_true',
                        help='save Gamma as complex64 and Lambda as float32, as MPS_cpu.py:397-398 does '
                             '(default: complex128 / float64)')
    args = parser.parse_args(argv)
    d, chi = args.d, args.chi
    path = args.dir + f"d_{d}_chi_{chi}/"
    sq_cov = np.load(args.dir + "sq_cov.npy")
    cov = np.load(args.dir + "cov.npy")
    M = len(cov) // 2
    kres = [np.load(path + f'res_{k}.npy').astype(np.float64) for k in range(M - 1)]
    knum = [np.load(path + f'num_{k}.npy') for k in range(M - 1)]
    kS = [np.load(path + f'S_{k}.npy') for k in range(M - 1)]
    start = 0 if args.restart else first_missing_site(path, M)
    if start:
        print(f'resuming at site {start} (sites 0..{start - 1} found in {path})')
    site_f = hostpool.submit_sites(sq_cov, kS, hostpool.williamson_S(sq_cov).result()) if start < M else []
    gdt, ldt = (np.complex64, np.float32) if args.reference_dtypes else (np.complex128, np.float64)
    for k in range(start, M):
        t0 = time.time()
        G, L = site_tensor(k, M, sq_cov, kres, knum, kS, d, chi, sigma_z=site_f[k].result())
        Gh = runtime.to_host(G).astype(gdt, copy=False)
        print('mode: ', k, 'Total {:.3f}'.format(time.time() - t0))
        if k < M - 1:  # Lambda first: a Gamma file marks a finished site
            _save_atomic(path + f"Lambda_{k}.npy", np.asarray(L).astype(ldt, copy=False))
        _save_atomic(path + f"Gamma_{k}.npy", Gh)


if __name__ == "__main__":
    main()
