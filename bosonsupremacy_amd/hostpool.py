"""Host worker pool for the FP64 linear algebra of MPS construction.

The Williamson decomposition of every cut (kron_cpu.py:96-106) and the Bloch-Messiah / Sigma
of every site (MPS_cpu.py:395-414, 439-442, 481-483) are small dense LAPACK problems whose
gauge (mode order and signs) is defined by the exact scipy call sequence, so they stay on
the host in FP64.  They are independent across cuts and across sites, and NumPy/SciPy hold
the GIL for most of their time at these sizes, so they run in a pool of worker processes.
Workers are forked -- before the GPU runtime is initialised when the pool is started by
``runtime.Context`` -- and only ever run NumPy/SciPy and ``gaussian`` (never a GPU call).
Workers keep the parent's BLAS configuration: OpenBLAS results at these sizes depend on its
thread count, so a worker computes exactly what a serial run would (every task executes the
same calls on the same inputs).  If the pool breaks, work continues serially.

``BSMPS_HOST_WORKERS`` sets the pool size (default: min(8, CPU count); 1 = run serially).
"""
import atexit
import multiprocessing
import os
from concurrent.futures import ProcessPoolExecutor
from concurrent.futures.process import BrokenProcessPool

import numpy as np

from . import gaussian

_pool = None


def n_workers():
    n = int(os.environ.get("BSMPS_HOST_WORKERS", "0") or 0)
    return n if n > 0 else min(8, os.cpu_count() or 1)


def _noop():
    return os.getpid()


def pool():
    """The process pool (created, with all its workers, on first use), or None when running
    serially."""
    global _pool
    if _pool is None and n_workers() > 1:
        p = None
        try:
            p = ProcessPoolExecutor(max_workers=n_workers(), mp_context=multiprocessing.get_context("fork"))
            # a worker that never answers (a fork taken while another thread held a lock) must
            # not block the parent: give up on the pool and run serially
            for f in [p.submit(_noop) for _ in range(n_workers())]:
                f.result(timeout=60)
        except Exception:
            if p is not None:
                for proc in list(getattr(p, "_processes", {}).values()):
                    proc.kill()
                p.shutdown(wait=False, cancel_futures=True)
            return None
        _pool = p
        atexit.register(shutdown)
    return _pool


def shutdown():
    global _pool
    if _pool is not None:
        _pool.shutdown(wait=False, cancel_futures=True)
        _pool = None


def _result(fut, fallback):
    """A task's result; on a broken pool, drop it and compute serially."""
    try:
        return fut.result()
    except BrokenProcessPool:
        shutdown()
        return fallback()


class _Task:
    """A pool future with a serial fallback."""

    def __init__(self, fut, fn, *args):
        self._fut, self._fn, self._args = fut, fn, args

    def done(self):
        return self._fut is None or isinstance(self._fut, _Done) or self._fut.done()

    def result_or_run(self):
        """The result if the worker has finished it, else computed here (the worker's copy is
        dropped), so a caller never waits behind the pool's queue."""
        if self._fut is not None and not isinstance(self._fut, _Done) and not self._fut.done():
            self._fut.cancel()
            self._fut = _Done(self._fn(*self._args))
        return self.result()

    def result(self):
        if self._fut is None:
            self._fut = _Done(self._fn(*self._args))
        if isinstance(self._fut, _Done):
            return self._fut.result()
        value = _result(self._fut, lambda: self._fn(*self._args))
        self._fut = _Done(value)
        return value


class _Done:
    """A finished future (serial mode)."""

    def __init__(self, value):
        self._value = value

    def result(self):
        return self._value


def _cut_task(sq_cov, L, cutoff):
    return gaussian.cut_inputs(sq_cov, L, cutoff)


def _site_task(k, M, sq_cov, S_prev, S_cur, S_full):
    cuts = {}
    if S_prev is not None:
        cuts[k - 1] = S_prev
    if S_cur is not None:
        cuts[k] = S_cur
    return gaussian.site_kernel(k, M, sq_cov, cuts, S_full)


def _tables_task(k, M, d, chi, res, num_k, num_prev, err_tol, ranks):
    from . import mps
    kres = {k: res} if res is not None else {}
    knum = {}
    if num_k is not None:
        knum[k] = num_k
    if num_prev is not None:
        knum[k - 1] = num_prev
    return mps.site_tables(k, M, d, chi, kres, knum, err_tol, ranks)


def _williamson_S(V):
    return gaussian.williamson(V)[1]


def _submit(fn, *args):
    p = pool()
    if p is None:
        return _Task(None, fn, *args)
    try:
        return _Task(p.submit(fn, *args), fn, *args)
    except (BrokenProcessPool, RuntimeError):
        shutdown()
        return _Task(None, fn, *args)


def cut_inputs_all(sq_cov, cutoff):
    """[(tp, S)] of cuts k = 0..M-2 (cut k keeps modes k+1..M-1)."""
    M = len(sq_cov) // 2
    sq_cov = np.ascontiguousarray(sq_cov)
    tasks = [_submit(_cut_task, sq_cov, k + 1, cutoff) for k in range(M - 1)]
    return [t.result() for t in tasks]


def williamson_S(V):
    """Task (``.result()``) of the Williamson symplectic of V."""
    return _submit(_williamson_S, np.ascontiguousarray(V))


def submit_tables(M, d, chi, kres, knum, err_tol, site_d=None, site_chi=None):
    """Tasks of the SiteTables of every site k = 0..M-1 (mps.site_tables; per-site cutoffs and
    per-cut ranks when given)."""
    tasks = []
    for k in range(M):
        ranks = None
        dk = d
        if site_d is not None or site_chi is not None:
            ranks = (site_chi[k - 1] if k > 0 else None, site_chi[k] if k < M - 1 else None)
            dk = site_d[k]
        res = kres[k] if k < M - 1 else None
        num_k = knum[k] if k < M - 1 else None
        num_prev = knum[k - 1] if k > 0 else None
        tasks.append(_submit(_tables_task, k, M, dk, chi, res, num_k, num_prev, err_tol, ranks))
    return tasks


def submit_sites(sq_cov, S_cuts, S_full):
    """Tasks (``.result()``) of (Sigma, Z) for every site k = 0..M-1 (gaussian.site_kernel)."""
    M = len(sq_cov) // 2
    sq_cov = np.ascontiguousarray(sq_cov)
    out = []
    for k in range(M):
        S_prev = S_cuts[k - 1] if k >= 1 else None
        S_cur = S_cuts[k] if k <= M - 2 else None
        out.append(_submit(_site_task, k, M, sq_cov, S_prev, S_cur, S_full if k == 0 else None))
    return out
