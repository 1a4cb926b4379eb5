"""The allocation-free Schmidt-spectrum entry point against the allocating one, through the C ABI:
two calls in a row (the second smaller, so the scratch holds the first call's contents), a
workload whose cuts need several launches of the fixed scratch, and the size limit."""
import ctypes

import numpy as np
import pytest
import torch

from conftest import same_bits

pytestmark = pytest.mark.gpu


def _run(entry, tps, chi, max_dim):
    from bosonsupremacy_amd import runtime
    ctx = runtime.context()
    n_cuts, d = len(tps), tps[0].shape[1]
    max_modes = max(t.shape[0] for t in tps)
    tp_all = np.zeros((n_cuts, max_modes, d))
    for c, t in enumerate(tps):
        tp_all[c, :t.shape[0]] = t
    n_modes = np.array([t.shape[0] for t in tps], dtype=np.int32)
    tp_d = runtime.to_device(tp_all)
    res = torch.full((n_cuts, chi), -1.0, dtype=torch.float64, device=ctx.device)
    num = torch.full((n_cuts, chi, max_modes), -1, dtype=torch.int8, device=ctx.device)
    ln = torch.full((n_cuts,), -1, dtype=torch.int32, device=ctx.device)
    ctx.call(entry, n_cuts, n_modes.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), max_modes, d,
             runtime.ptr(tp_d), int(max_dim), int(chi), 1e-12, runtime.ptr(res), runtime.ptr(num), runtime.ptr(ln))
    return runtime.to_host(res), runtime.to_host(num), runtime.to_host(ln)


def _tps(name, cuts=None):
    from bosonsupremacy_amd import instances, kron
    inst = instances.workload(name)
    cuts = range(inst.M - 1) if cuts is None else cuts
    return [kron.cut_spectrum_inputs(inst.sq_cov, k + 1, inst.d)[0] for k in cuts], inst.chi


@pytest.mark.parametrize("order", [("m32_d6_chi512", "m8_d4_chi64")])
def test_noalloc_matches_allocating_entry_twice(order):
    for name in order:   # the second call is smaller and runs on scratch the first one filled
        tps, chi = _tps(name)
        print(f"kron no-alloc: {name}, {len(tps)} cuts", flush=True)
        want = _run("bsmps_kron_cuts", tps, chi, 10 ** 5)
        got = _run("bsmps_kron_cuts_noalloc", tps, chi, 10 ** 5)
        for w, g in zip(want, got):
            assert same_bits(g, w), name


def test_noalloc_many_launches():
    """M = 144, chi = 1e4 cuts at max_dim 1e5: the back-pointer blocks (about 4 GB in all) take
    many launches of the 256 MiB scratch."""
    from bosonsupremacy_amd import _lib
    assert _lib.load().bsmps_kron_noalloc_scratch_bytes() == 256 << 20
    sel, chi = _tps("m144_d10_chi10000", (0, 5, 40, 71, 100, 141, 142))
    print(f"kron no-alloc: {len(sel)} cuts of M = 144", flush=True)
    want = _run("bsmps_kron_cuts", sel, chi, 10 ** 5)
    got = _run("bsmps_kron_cuts_noalloc", sel, chi, 10 ** 5)
    for w, g in zip(want, got):
        assert same_bits(g, w)


def test_noalloc_refuses_cuts_beyond_its_scratch():
    from bosonsupremacy_amd import _lib, runtime
    ctx = runtime.context()
    n_modes = np.array([3, 2000], dtype=np.int32)   # 2000 modes at max_dim 1e5: ~800 MB of back-pointers
    dummy = torch.zeros(8, dtype=torch.float64, device=ctx.device)
    st = ctx.lib.bsmps_kron_cuts_noalloc(ctx.handle, 2, n_modes.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                         2000, 4, runtime.ptr(dummy), 10 ** 5, 16, 1e-12, runtime.ptr(dummy),
                                         runtime.ptr(dummy), runtime.ptr(dummy))
    assert st == _lib.BSMPS_ERR_UNSUPPORTED
    assert b"scratch" in ctx.lib.bsmps_last_error()
