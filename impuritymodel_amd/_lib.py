"""ctypes binding of libimped.so (declared in include/imped.h).

The library is built in-tree by ``impuritymodel_amd.build_lib``.  There is no
CPU fallback: loading fails with an ImportError when the library is missing,
and calls raise when no GPU is present.
"""

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libimped.so")

IMPED_OK = 0
IMPED_ERR_INVALID = -1
IMPED_ERR_HIP = -2
IMPED_ERR_CAPACITY = -3
IMPED_ERR_BREAKDOWN = -4
IMPED_ERR_NOCONV = -5

_p = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_u64 = C.c_uint64
_d = C.c_double
_f = C.c_float

# name -> argtypes (all return int)
SIGNATURES = {
    "imped_abi_version": [],
    "imped_set_stream": [_p],
    "imped_device_synchronize": [],
    "imped_timing_enable": [_i32],
    "imped_timing_reset": [],
    "imped_timing_query": [C.c_char_p, _p, _p],
    "imped_set_tuning": [C.c_char_p, _i64],
    "imped_sort_u64": [_p, _i64, _i32, _i32, _p],
    "imped_enumerate_basis": [_i32, _i32, _i32, _p, _i64, _p, _p],
    "imped_ops_create": [_i32, _i32, _i32, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p],
    "imped_ops_destroy": [_p],
    "imped_ops_set_coef": [_p, _p, _i64],
    "imped_ops_set_affine": [_p, _i32, _i64, _p, _p, _i32, _p],
    "imped_vec_create": [_i32, _i32, _p, _p, _p, _p],
    "imped_vec_destroy": [_p],
    "imped_vec_offsets": [_p, _p],
    "imped_vec_get": [_p, _p, _p, _p],
    "imped_vec_apply": [_p, _p, _i32, _p, _p, _d, _i32, _p],
    "imped_vec_from_eigen": [_p, _i32, _p, _p, _d, _p],
    "imped_vec_prune": [_p, _d, _p],
    "imped_batch_solve": [_p, _i32, _p, _d, _i32, _p, _p, _p],
    "imped_batch_create": [_p, _i32, _p, _p],
    "imped_batch_destroy": [_p],
    "imped_batch_set_seeds": [_p, _p, _p, _p],
    "imped_batch_seed_from_vec": [_p, _p],
    "imped_batch_build": [_p, _i32, _d, _i32],
    "imped_batch_set_csr": [_p, _p, _p, _p, _p],
    "imped_batch_sizes": [_p, _p, _p, _p],
    "imped_batch_get_basis": [_p, _p],
    "imped_batch_get_csr": [_p, _p, _p, _p],
    "imped_batch_lanczos": [_p, _p, _i32, _i32, _p, _p, _p, _p],
    "imped_affine_create": [_i32, _p, _p, _p, _p, _i32, _p, _p],
    "imped_affine_destroy": [_p],
    "imped_batch_set_affine": [_p, _p, _p, _p, _i32, _p, _d],
    "imped_batch_seed_affine": [_p, _p, _p, _p, _p, _p, _d, _d, _p],
    "imped_family_xas": [_p, _p, _p, _i32, _i32, _p, _i32, _p, _p, _i32, _i32, _d, _i32, _i32, _d, _i32, _u64, _d,
                         _d, _d, _p, _i32, _d, _p, _p, _p, _p, _p, _p, _p],
    "imped_lanczos_noalloc": [_p, _p, _p, _i64, _p, _i32, _i32, _p, _p, _p, _p, _p],
    "imped_batch_eigen": [_p, _i32, _d, _i32, _u64, _p, _p, _p],
    "imped_batch_cf": [_p, _p, _p, _p, _p, _p, _p, _i32, _p, _i32, _d, _p, _p],
    "imped_mlp_create": [_i32, _i32, _p, _i32, _p],
    "imped_mlp_destroy": [_p],
    "imped_mlp_param_count": [_p, _p],
    "imped_mlp_set_params": [_p, _p],
    "imped_mlp_get_params": [_p, _p],
    "imped_mlp_forward": [_p, _p, _i32, _p],
    "imped_mlp_predict": [_p, _p, _i32, _p, _d, _p, _p],
    "imped_mlp_misfit": [_p, _p, _i32, _p, _d, _p, _p, _p, _p],
    "imped_walkers_create": [_p, _i32, _i32, _p, _p, _p, _p, _d, _d, _d, _u64, _p],
    "imped_walkers_destroy": [_p],
    "imped_walkers_step": [_p, _p, _d, _p, _p],
    "imped_walkers_waiting": [_p, _p, _p, _p],
    "imped_walkers_decide": [_p, _i32, _p, _p],
    "imped_walkers_state": [_p, _p, _p, _p, _p],
    "imped_walkers_served": [_p, _i64, _i64, _p],
    "imped_walkers_set_record_cap": [_p, _i64],
    "imped_walkers_reset_counters": [_p],
    "imped_pinned_alloc": [_i64, _p],
    "imped_pinned_free": [_p],
    "imped_mlp_gradients": [_p, _p, _p, _i32, _p, _p],
    "imped_mlp_set_data": [_p, _p, _p, _i64],
    "imped_mlp_train": [_p, _p, _i32, _i32, _f, _f, _f, _f, _p],
}


class ImpedError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("libimped error %d: %s" % (code, msg))
        self.code = code


def _torch_runtime_first():
    """Let PyTorch's GPU runtime initialize before this library is loaded.

    PyTorch carries its own HIP runtime.  When libimped.so is loaded first
    and PyTorch's runtime then initializes the device, the first HIP call of
    libimped.so finds no device (hipErrorNoDevice); with PyTorch's runtime
    initialized first, both work.  Counting devices does not initialize
    anything; ``torch.cuda.init()`` runs only where a device exists."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.device_count() > 0:
        torch.cuda.init()


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libimped.so is not built (expected at %s); run python -m impuritymodel_amd.build_lib" % LIB_PATH
        )
    _torch_runtime_first()
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.imped_last_error.argtypes = []
    lib.imped_last_error.restype = C.c_char_p
    return lib


lib = _load()


def check(rc, allow=()):
    if rc != IMPED_OK and rc not in allow:
        raise ImpedError(rc, lib.imped_last_error().decode(errors="replace"))
    return rc


def ptr(a):
    """Data pointer of a numpy array (None -> NULL)."""
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"], "arrays passed to libimped must be contiguous"
    return a.ctypes.data_as(C.c_void_p)
