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
