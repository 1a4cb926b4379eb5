"""NumPy Philox4x32-10 -- TEST INFRASTRUCTURE ONLY (checker for the device RNG).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use this module.
Follows Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC'11),
Philox4x32 with 10 rounds; pinned by the published known-answer vectors (tests/test_oracle.py).
The stream layout (counter = (element lo, element hi, stream, 0), key = seed) is the one
include/bsmps.h documents for bsmps_philox.
"""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    c0, c1, c2, c3 = (np.asarray(x, dtype=np.uint32) for x in (c0, c1, c2, c3))
    k0 = np.asarray(k0, dtype=np.uint32)
    k1 = np.asarray(k1, dtype=np.uint32)
    with np.errstate(over="ignore"):
        for _ in range(10):
            p0 = M0 * c0.astype(np.uint64)
            p1 = M1 * c2.astype(np.uint64)
            hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & MASK).astype(np.uint32)
            hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & MASK).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
            k0 = (k0 + W0).astype(np.uint32)
            k1 = (k1 + W1).astype(np.uint32)
    return c0, c1, c2, c3


def _u53(lo, hi):
    v = (hi.astype(np.uint64) << np.uint64(32)) | lo.astype(np.uint64)
    return (v >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def stream(seed, stream_id, first, n, normal=False):
    e = np.arange(first, first + n, dtype=np.uint64)
    c0 = (e & MASK).astype(np.uint32)
    c1 = (e >> np.uint64(32)).astype(np.uint32)
    c2 = np.full(n, stream_id, dtype=np.uint32)
    c3 = np.zeros(n, dtype=np.uint32)
    seed = np.uint64(seed)
    x0, x1, x2, x3 = philox4x32_10(c0, c1, c2, c3, np.uint32(seed & MASK), np.uint32(seed >> np.uint64(32)))
    u1 = _u53(x0, x1)
    if not normal:
        return u1
    u2 = _u53(x2, x3)
    return np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(6.283185307179586 * u2)
