"""Device runtime: one bsmps context per GPU, enqueued on torch's current stream.

PyTorch is used only as device-memory and stream plumbing: buffers are torch tensors, and
their data pointers go to the HIP kernels through the C ABI.
"""
import atexit
import ctypes

import numpy as np
import torch

from . import _lib

_contexts = {}


def _drain_at_exit():
    # Kernels still in flight when the interpreter tears down would race with the release
    # of the tensors they use; wait for them first.
    for ctx in list(_contexts.values()):
        try:
            ctx.synchronize()
        except Exception:
            pass


atexit.register(_drain_at_exit)


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("bosonsupremacy_amd needs a ROCm GPU; none is visible")


class Context:
    """Owns a bsmps_ctx bound to ``device`` and enqueuing on the torch stream current at
    creation (normally torch's default stream, HIP's null stream), so library kernels are
    ordered with torch's allocations, fills, frees and copies of the same buffers.  A context
    created inside ``torch.cuda.stream(s)`` enqueues on s (used for graph capture)."""

    def __init__(self, device=0):
        from . import hostpool
        hostpool.pool()  # fork the host workers before this process initialises the GPU
        _require_gpu()
        lib = _lib.load()
        self.device = torch.device("cuda", device)
        self.stream = torch.cuda.current_stream(self.device)
        handle = ctypes.c_void_p()
        _lib.check(lib.bsmps_create(device, ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(handle)))
        self.handle = handle
        self.lib = lib
        self.settings = {}

    # launch-tuning setters; their last successful calls are kept so another context can
    # be given the same settings (copy_settings)
    _SETTERS = {"bsmps_set_tuning": 0, "bsmps_set_group_lanes": 1, "bsmps_set_split": 1,
                "bsmps_set_split_budget": 0, "bsmps_set_contract_form": 1}

    def call(self, name, *args):
        status = _lib.check(getattr(self.lib, name)(self.handle, *args))
        if name in self._SETTERS:
            if name == "bsmps_set_tuning":  # resets every size's lane count
                self.settings = {k: v for k, v in self.settings.items() if k[0] != "bsmps_set_group_lanes"}
            self.settings[(name,) + tuple(args[:self._SETTERS[name]])] = args
        return status

    def copy_settings(self, other):
        """Apply ``other``'s launch-tuning settings to this context, in the order they were made."""
        for key, args in list(other.settings.items()):
            self.call(key[0], *args)

    def synchronize(self):
        self.call("bsmps_synchronize")

    def set_tuning(self, lane_max_m, group_lanes):
        self.call("bsmps_set_tuning", lane_max_m, group_lanes)

    def close(self):
        if self.handle:
            _lib.check(self.lib.bsmps_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def context(device=0):
    """The process-wide context of ``device`` (created on first use)."""
    ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        _contexts[device] = ctx
    return ctx


def ptr(t):
    """Device pointer of a contiguous device tensor (or None -> NULL).  A host tensor -- pinned
    or not -- is refused: the kernels would read host memory over PCIe, or fault."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("a device tensor is needed here; copy host tensors to the GPU first")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def host(a):
    """NumPy view/copy of an array or tensor (device tensors are copied to the host)."""
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def to_device(a, dtype=None, device=None):
    """numpy (or tensor) -> contiguous device tensor; complex128 stays complex128."""
    device = device or context().device
    if isinstance(a, torch.Tensor):
        t = a.to(device=device, dtype=dtype) if dtype is not None else a.to(device=device)
        return t.contiguous()
    arr = np.ascontiguousarray(a)
    t = torch.from_numpy(arr)
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device).contiguous()


def upload(a, device=None):
    """numpy -> device tensor without waiting for the stream: the array is staged in pinned
    host memory (torch's caching host allocator keeps it alive until the copy has run) and
    copied asynchronously on the current stream."""
    device = device or context().device
    arr = np.ascontiguousarray(a)
    if arr.size == 0:
        return torch.from_numpy(arr).to(device)
    return torch.from_numpy(arr).pin_memory().to(device, non_blocking=True)


_TORCH_DTYPES = {np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32,
                 np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64, np.dtype(np.int8): torch.int8,
                 np.dtype(np.uint8): torch.uint8, np.dtype(np.complex128): torch.complex128,
                 np.dtype(np.complex64): torch.complex64}


def upload_many(arrays, device=None):
    """Several numpy arrays -> device tensors through one pinned staging buffer and one
    asynchronous copy on the current stream (each array 256-byte aligned in it); the tensors
    are views of one device allocation."""
    device = device or context().device
    arrs = [np.ascontiguousarray(a) for a in arrays]
    offs, total = [], 0
    for a in arrs:
        total = (total + 255) // 256 * 256
        offs.append(total)
        total += a.nbytes
    host = torch.empty((max(total, 1),), dtype=torch.uint8, pin_memory=True)
    hv = host.numpy()
    for a, o in zip(arrs, offs):
        hv[o:o + a.nbytes] = a.reshape(-1).view(np.uint8)
    dev = host.to(device, non_blocking=True)
    return [dev[o:o + a.nbytes].view(_TORCH_DTYPES[a.dtype]).view(a.shape) for a, o in zip(arrs, offs)]


def to_host(t):
    return t.detach().cpu().numpy()
