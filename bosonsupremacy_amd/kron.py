"""Schmidt spectra of all cuts (L2) -- drop-in for kron_cpu.py.

Per cut the host computes the Williamson decomposition of the reduced pure covariance and the
thermal distribution of each normal mode (kron_cpu.py:96-106); the pruned Kronecker loop and
final top-chi (kron_cpu.py:109-125) run on the GPU for all cuts in one launch
(``bsmps_kron_cuts``).

Differences from the reference, by design:
* ``res`` is float64 (the reference casts to float16, kron_cpu.py:127);
* weights that tie exactly are ordered by (photon number of the newest mode, then earlier
  rank); the reference's ``np.argsort`` leaves that order unspecified.

CLI (same flags and files as the reference)::

    python -m bosonsupremacy_amd.kron --d D --chi CHI --dir ROOT/
"""
import argparse
import contextlib
import ctypes
import os

import numpy as np
import torch

from . import gaussian, hostpool, runtime

ERR_TOL = 1e-12    # kron_cpu.py:95
MAX_DIM = 10 ** 5  # kron_cpu.py:145


def cut_spectrum_inputs(sq_cov, L, cutoff):
    """Williamson of modes L..M-1 and their thermal distributions [n, cutoff]
    (kron_cpu.py:96-106)."""
    return gaussian.cut_inputs(sq_cov, L, cutoff)


def kron_cuts(tps, chi, max_dim=MAX_DIM, err_tol=ERR_TOL, device=0, timer=None, no_alloc=False, ctx=None):
    """Top-chi Schmidt weights and labels of several cuts at once on the GPU.

    tps: list of arrays [n_c, d] (thermal distributions of each cut's modes).
    Returns a list of (res float64 [len], num int8 [len, n_c] -- 1-D when n_c == 1).
    ``no_alloc`` runs the library's allocation-free entry point (fixed device scratch; the
    same results).  ``ctx`` (default: the device's process-wide context) is the library
    context to enqueue on; run under ``torch.cuda.stream(s)`` for a context bound to s, so the
    buffers and the copy back are ordered on s.
    """
    ctx = ctx or runtime.context(device)
    n_cuts = len(tps)
    if n_cuts == 0:
        return []
    d = tps[0].shape[1]
    max_modes = max(t.shape[0] for t in tps)
    tp_all = np.zeros((n_cuts, max_modes, d))
    for c, t in enumerate(tps):
        tp_all[c, :t.shape[0]] = t
    n_modes = np.array([t.shape[0] for t in tps], dtype=np.int32)
    dev = ctx.device
    tp_d = runtime.to_device(tp_all)
    res_d = torch.zeros((n_cuts, chi), dtype=torch.float64, device=dev)
    num_d = torch.zeros((n_cuts, chi, max_modes), dtype=torch.int8, device=dev)
    len_d = torch.zeros((n_cuts,), dtype=torch.int32, device=dev)
    with (timer("kron_cuts") if timer else contextlib.nullcontext()):
        ctx.call("bsmps_kron_cuts_noalloc" if no_alloc else "bsmps_kron_cuts", n_cuts, n_modes.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), max_modes, d,
                 runtime.ptr(tp_d), int(max_dim), int(chi), float(err_tol), runtime.ptr(res_d),
                 runtime.ptr(num_d), runtime.ptr(len_d))
    res_h, num_h, len_h = runtime.to_host(res_d), runtime.to_host(num_d), runtime.to_host(len_d)
    out = []
    for c in range(n_cuts):
        k = int(len_h[c])
        nm = int(n_modes[c])
        num = np.ascontiguousarray(num_h[c, :k, :nm])
        if nm == 1:
            num = np.ascontiguousarray(num[:, 0])
        out.append((np.ascontiguousarray(res_h[c, :k]), num))
    return out


def get_cumsum_kron(sq_cov, L, chi=100, max_dim=MAX_DIM, cutoff=6, err_tol=ERR_TOL):
    """Same signature and outputs as kron_cpu.py:95-127 (res in float64)."""
    tp, S = cut_spectrum_inputs(sq_cov, L, cutoff)
    (res, num), = kron_cuts([tp], chi, max_dim, err_tol)
    return res, num, S


def kron_all(sq_cov, d, chi, max_dim=MAX_DIM, err_tol=ERR_TOL, timer=None):
    """All cuts k = 0..M-2 (cut k keeps modes k+1..M-1), as kron_cpu.py:147-153.  The
    per-cut Williamson decompositions run in the host worker pool."""
    M = len(sq_cov) // 2
    cuts = hostpool.cut_inputs_all(sq_cov, d)
    spec = kron_cuts([tp for tp, _ in cuts], chi, max_dim, err_tol, timer=timer)
    return [r for r, _ in spec], [n for _, n in spec], [S for _, S in cuts]


def main(argv=None):
    parser = argparse.ArgumentParser(description="Schmidt spectra of every cut (drop-in for kron_cpu.py)")
    parser.add_argument('--d', type=int, required=True, help='local photon cutoff before displacement')
    parser.add_argument('--chi', type=int, required=True, help='bond dimension')
    parser.add_argument('--dir', type=str, required=True, help='root directory (with trailing slash)')
    parser.add_argument('--reference-dtypes', action='store_true',
                        help='save res as float16 and num as int8, as kron_cpu.py:127 does (default: res float64)')
    args = parser.parse_args(argv)
    path = args.dir + f"d_{args.d}_chi_{args.chi}/"
    sq_cov = np.load(args.dir + "sq_cov.npy")
    cov = np.load(args.dir + "cov.npy")
    M = len(cov) // 2
    os.makedirs(path, exist_ok=True)
    np.save(path + 'active_kron_sites.npy', np.zeros(M - 1, dtype='int32'))
    res, num, S = kron_all(sq_cov, args.d, args.chi)
    for k in range(M - 1):
        print(k, np.sum(res[k]))
        np.save(path + f'res_{k}.npy', res[k].astype(np.float16) if args.reference_dtypes else res[k])
        np.save(path + f'num_{k}.npy', num[k].astype(np.int8))
        np.save(path + f'S_{k}.npy', S[k])


if __name__ == "__main__":
    main()
