"""ctypes binding of libbsmps.so (include/bsmps.h).

The library is built in-tree by ``python -m bosonsupremacy_amd.build``.  There is no
fallback: if the library or a GPU is missing, every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbsmps.so")

BSMPS_OK = 0
BSMPS_ERR_INVALID = 1
BSMPS_ERR_HIP = 2
BSMPS_ERR_UNSUPPORTED = 3
MAX_HAFNIAN_SIZE = 34

SITE_FIRST, SITE_MIDDLE, SITE_LAST = 0, 1, 2

c_int, c_int64, c_uint64, c_uint32, c_double = (ctypes.c_int, ctypes.c_int64, ctypes.c_uint64,
                                               ctypes.c_uint32, ctypes.c_double)
c_void_p = ctypes.c_void_p
c_size_t = ctypes.c_size_t
c_float = ctypes.c_float
c_int32_p = ctypes.POINTER(ctypes.c_int32)


class BsmpsError(RuntimeError):
    """A libbsmps entry point returned an error status."""

    def __init__(self, status, message):
        super().__init__(f"bsmps error {status}: {message}")
        self.status = status


class GammaSiteArgs(ctypes.Structure):
    """bsmps_gamma_site_args (include/bsmps.h)."""
    _fields_ = [
        ("kind", c_int), ("d", c_int), ("chi", c_int), ("ld", c_int),
        ("sigma", c_void_p), ("z", c_double), ("inv_z", c_double), ("sqrt_fact", c_void_p),
        ("n_left", c_int), ("w_left", c_int), ("left_targets", c_void_p), ("left_den", c_void_p),
        ("lambda_", c_void_p), ("left_perm", c_void_p), ("left_class_count_host", c_int32_p),
        ("n_left_class", c_int),
        ("n_right", c_int), ("w_right", c_int), ("right_targets", c_void_p), ("right_den", c_void_p),
        ("right_perm", c_void_p), ("right_class_count_host", c_int32_p), ("n_right_class", c_int),
    ]


_PROTOTYPES = {
    "bsmps_create": ([c_int, c_void_p, ctypes.POINTER(c_void_p)], c_int),
    "bsmps_destroy": ([c_void_p], c_int),
    "bsmps_synchronize": ([c_void_p], c_int),
    "bsmps_last_error": ([], ctypes.c_char_p),
    "bsmps_version": ([], c_int),
    "bsmps_set_tuning": ([c_void_p, c_int, c_int], c_int),
    "bsmps_set_group_lanes": ([c_void_p, c_int, c_int], c_int),
    "bsmps_set_split": ([c_void_p, c_int, c_int], c_int),
    "bsmps_set_split_budget": ([c_void_p, c_size_t], c_int),
    "bsmps_set_contract_form": ([c_void_p, c_int, c_int], c_int),
    "bsmps_sigma_select": ([c_void_p, c_void_p, c_int, c_void_p, c_int64, c_int, c_void_p], c_int),
    "bsmps_hafnian_dense": ([c_void_p, c_void_p, c_int64, c_int, c_void_p], c_int),
    "bsmps_hafnian_targets": ([c_void_p, c_void_p, c_int, c_void_p, c_int64, c_int, c_void_p, c_void_p], c_int),
    "bsmps_loop_hafnian_dense": ([c_void_p, c_void_p, c_void_p, c_int64, c_int, c_void_p], c_int),
    "bsmps_loop_hafnian_targets": ([c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int64, c_int, c_void_p,
                                    c_void_p], c_int),
    "bsmps_gamma_site": ([c_void_p, ctypes.POINTER(GammaSiteArgs), c_void_p], c_int),
    "bsmps_gamma_sites": ([c_void_p, c_int, ctypes.POINTER(GammaSiteArgs), ctypes.POINTER(c_void_p)], c_int),
    "bsmps_kron_cuts": ([c_void_p, c_int, c_int32_p, c_int, c_int, c_void_p, c_int64, c_int, c_double,
                         c_void_p, c_void_p, c_void_p], c_int),
    "bsmps_kron_cuts_noalloc": ([c_void_p, c_int, c_int32_p, c_int, c_int, c_void_p, c_int64, c_int, c_double,
                                 c_void_p, c_void_p, c_void_p], c_int),
    "bsmps_kron_noalloc_scratch_bytes": ([], c_int64),
    "bsmps_mu_to_alpha": ([c_void_p, c_void_p, c_int, c_void_p, c_int64, c_void_p], c_int),
    "bsmps_batch_displaces": ([c_void_p, c_int, c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p], c_int),
    "bsmps_site_contract": ([c_void_p, c_int64, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p],
                            c_int),
    "bsmps_sample_site": ([c_void_p, c_int64, c_int, c_int, c_int, c_void_p, c_int64, c_void_p, c_int64,
                           c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p],
                          c_int),
    "bsmps_site_contract_c64": ([c_void_p, c_int64, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                 c_void_p], c_int),
    "bsmps_sample_site_c64": ([c_void_p, c_int64, c_int, c_int, c_int, c_void_p, c_int64, c_void_p, c_int64,
                               c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                               c_void_p], c_int),
    "bsmps_split_bf16": ([c_void_p, c_void_p, c_int64, c_void_p, c_void_p], c_int),
    "bsmps_site_contract_bf16x3": ([c_void_p, c_int64, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p], c_int),
    "bsmps_bf16x6_plane_bytes": ([c_int, c_int, c_int], c_int64),
    "bsmps_bf16x6_row_plane_bytes": ([c_int64, c_int], c_int64),
    "bsmps_split_bf16x6": ([c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p], c_int),
    "bsmps_split_rows_bf16x6": ([c_void_p, c_int64, c_int, c_void_p, c_void_p], c_int),
    "bsmps_site_contract_bf16x6": ([c_void_p, c_int64, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p], c_int),
    "bsmps_fp16x3_plane_bytes": ([c_int, c_int, c_int], c_int64),
    "bsmps_fp16x3_row_plane_bytes": ([c_int64, c_int], c_int64),
    "bsmps_split_fp16x3": ([c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "bsmps_split_rows_fp16x3": ([c_void_p, c_int64, c_int, c_void_p, c_void_p, c_void_p], c_int),
    "bsmps_site_contract_fp16x3": ([c_void_p, c_int64, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p], c_int),
    "bsmps_site_contract_cols": ([c_void_p, c_int64, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_int64, c_void_p], c_int),
    "bsmps_site_contract_c64_cols": ([c_void_p, c_int64, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_int64, c_void_p], c_int),
    "bsmps_site_contract_bf16x6_cols": ([c_void_p, c_int64, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                         c_int64, c_void_p], c_int),
    "bsmps_site_contract_fp16x3_cols": ([c_void_p, c_int64, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_int64, c_void_p], c_int),
    "bsmps_philox": ([c_void_p, c_uint64, c_uint32, c_int, c_int64, c_int64, c_int, c_void_p], c_int),
    "bsmps_mlp_forward": ([c_void_p, c_int, c_int, c_int, c_int] + [c_void_p] * 10, c_int),
    "bsmps_mlp_backward": ([c_void_p, c_int, c_int, c_int, c_int] + [c_void_p] * 13, c_int),
    "bsmps_f32_to_bf16": ([c_void_p, c_int64, c_void_p, c_void_p], c_int),
    "bsmps_adam_step": ([c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_float,
                         c_float, c_int], c_int),
    "bsmps_mse_grad": ([c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p], c_int),
}

EXPORTED_SYMBOLS = tuple(_PROTOTYPES)

_lib = None


def load():
    """Load libbsmps.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing; build it with `python -m bosonsupremacy_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (argtypes, restype) in _PROTOTYPES.items():
            fn = getattr(lib, name)
            fn.argtypes = argtypes
            fn.restype = restype
        _lib = lib
    return _lib


def check(status):
    if status != BSMPS_OK:
        msg = load().bsmps_last_error()
        raise BsmpsError(status, msg.decode() if msg else "")
    return status
