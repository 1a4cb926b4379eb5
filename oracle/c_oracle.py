"""ctypes loader for the C hafnian restatement -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libhafnian_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        lib.oracle_hafnian_targets.argtypes = [dp, ctypes.c_int, ip, ctypes.c_long, ctypes.c_int, dp]
        lib.oracle_hafnian_targets.restype = None
        lib.oracle_hafnian_dense.argtypes = [dp, ctypes.c_long, ctypes.c_int, dp]
        lib.oracle_hafnian_dense.restype = None
        lib.oracle_loop_hafnian_targets.argtypes = [dp, ctypes.c_int, dp, ip, ctypes.c_long, ctypes.c_int, dp]
        lib.oracle_loop_hafnian_targets.restype = None
        lib.oracle_loop_hafnian_dense.argtypes = [dp, dp, ctypes.c_long, ctypes.c_int, dp]
        lib.oracle_loop_hafnian_dense.restype = None
        _lib = lib
    return _lib


def hafnian_targets(Sigma, targets):
    """hafnian(Sigma[t][:, t]) for each row t of ``targets`` (MPS_cpu.py:117-187)."""
    lib = _load()
    S = np.ascontiguousarray(Sigma, dtype=np.complex128)
    T = np.ascontiguousarray(targets, dtype=np.int32)
    if T.ndim != 2:
        raise ValueError("targets must be 2-D")
    out = np.zeros(T.shape[0], dtype=np.complex128)
    lib.oracle_hafnian_targets(S.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), S.shape[0],
                               T.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), T.shape[0],
                               T.shape[1], out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def hafnian_dense(A):
    """hafnian of each matrix of A [B, n, n]."""
    lib = _load()
    A = np.ascontiguousarray(A, dtype=np.complex128)
    out = np.zeros(A.shape[0], dtype=np.complex128)
    lib.oracle_hafnian_dense(A.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), A.shape[0],
                             A.shape[1], out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def loop_hafnian_dense(A, loops):
    """Loop hafnian of each matrix of A [B, n, n] with loop weights loops [B, n] (the plain
    recursion of MPS_cpu.py:143-187 extended by one loop polynomial per row; an odd n is padded
    with a vertex of loop weight 1 and no edges)."""
    lib = _load()
    A = np.ascontiguousarray(A, dtype=np.complex128)
    L = np.ascontiguousarray(np.broadcast_to(loops, A.shape[:2]), dtype=np.complex128)
    out = np.zeros(A.shape[0], dtype=np.complex128)
    lib.oracle_loop_hafnian_dense(_dp(A), _dp(L), A.shape[0], A.shape[1], _dp(out))
    return out


def loop_hafnian_targets(Sigma, loops, targets):
    """lhaf(Sigma[t][:, t]) with loop weights loops[t] for each row t of ``targets``."""
    lib = _load()
    S = np.ascontiguousarray(Sigma, dtype=np.complex128)
    L = np.ascontiguousarray(loops, dtype=np.complex128)
    T = np.ascontiguousarray(targets, dtype=np.int32)
    if T.ndim != 2 or L.shape != (S.shape[0],):
        raise ValueError("targets must be 2-D and loops a vector over Sigma's rows")
    out = np.zeros(T.shape[0], dtype=np.complex128)
    lib.oracle_loop_hafnian_targets(_dp(S), S.shape[0], _dp(L), T.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                    T.shape[0], T.shape[1], _dp(out))
    return out
