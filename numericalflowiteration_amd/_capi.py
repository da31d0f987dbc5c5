"""ctypes binding of libnufi_hip.so (the C ABI declared in include/nufi_hip.h).

The library is built in-tree by ``numericalflowiteration_amd.build.build()``
(hipcc, gfx950).  There is no CPU fallback: if the library is missing or fails
to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

from .config import Config

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libnufi_hip.so")

OK = 0
ERR_INVALID = 1
ERR_RANGE = 2
ERR_DEVICE = 3
ERR_ALLOC = 4
ERR_UNSUPPORTED = 5
ERR_IO = 6

HISTORY_BINARY = 0
HISTORY_TEXT = 1

TRACE_FAST = 0
TRACE_STRICT = 1
TRACE_FP32C = 2

class HistoryInfo(ctypes.Structure):
    """nufi_history_info of include/nufi_hip.h."""
    _fields_ = [("dim", ctypes.c_int), ("order", ctypes.c_int),
                ("Nx", ctypes.c_uint64), ("Ny", ctypes.c_uint64), ("Nz", ctypes.c_uint64),
                ("Nt", ctypes.c_uint64), ("n_slices", ctypes.c_uint64), ("stride_t", ctypes.c_uint64),
                ("dt", ctypes.c_double), ("n_energies", ctypes.c_uint64)]


_dp = ctypes.POINTER(ctypes.c_double)
_cp = ctypes.POINTER(Config)
_hp = ctypes.POINTER(HistoryInfo)
_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_int = ctypes.c_int

# name -> (restype, argtypes); exactly the functions of include/nufi_hip.h
SIGNATURES = {
    "nufi_version": (ctypes.c_char_p, []),
    "nufi_config_derive": (_int, [_cp]),
    "nufi_device_count": (_int, [ctypes.POINTER(_int)]),
    "nufi_create": (_int, [_cp, _cp, _int, _int, _int, ctypes.POINTER(_vp)]),
    "nufi_destroy": (_int, [_vp]),
    "nufi_last_error": (ctypes.c_char_p, [_vp]),
    "nufi_get_stream": (_int, [_vp, ctypes.POINTER(_vp)]),
    "nufi_synchronize": (_int, [_vp]),
    "nufi_compute_rho": (_int, [_vp, _u64, _u64, _u64]),
    "nufi_download_rho": (_int, [_vp, _dp]),
    "nufi_upload_phi": (_int, [_vp, _u64, _dp]),
    "nufi_compute_metrics": (_int, [_vp, _u64, _u64, _u64]),
    "nufi_download_metrics": (_int, [_vp, _dp]),
    "nufi_step": (_int, [_vp, _u64, _dp]),
    "nufi_run": (_int, [_vp, _u64, _u64, _dp]),
    "nufi_set_run_graph": (_int, [_vp, _int]),
    "nufi_download_history": (_int, [_vp, _u64, _u64, _dp]),
    "nufi_upload_history": (_int, [_vp, _u64, _u64, _dp]),
    "nufi_download_energies": (_int, [_vp, _u64, _u64, _dp]),
    "nufi_poisson_solve": (_int, [_vp, _dp, _dp]),
    "nufi_interpolate": (_int, [_vp, _dp, _dp]),
    "nufi_set_trace_variant": (_int, [_vp, _int]),
    "nufi_set_launch_shape": (_int, [_vp, _int, _int, _int]),
    "nufi_get_launch_shape": (_int, [_vp, ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(_int)]),
    "nufi_autotune": (_int, [_vp, _u64, _int, ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(_int),
                             _dp]),
    "nufi_set_timing": (_int, [_vp, _int]),
    "nufi_trace_timing": (_int, [_vp, ctypes.POINTER(_u64), _dp, _dp]),
    "nufi_adjoint_grad": (_int, [_vp, _dp, _int, _dp]),
    "nufi_rho3d_noalloc": (_int, [_cp, _u64, _vp, _u64, _vp, _vp]),
    "nufi_batch_create": (_int, [_cp, _u64, _int, _int, _int, ctypes.POINTER(_vp)]),
    "nufi_batch_destroy": (_int, [_vp]),
    "nufi_batch_last_error": (ctypes.c_char_p, [_vp]),
    "nufi_batch_run": (_int, [_vp, _u64, _u64, _int]),
    "nufi_batch_synchronize": (_int, [_vp]),
    "nufi_batch_energies": (_int, [_vp, _u64, _u64, _u64, _dp]),
    "nufi_batch_history": (_int, [_vp, _u64, _u64, _u64, _dp]),
    "nufi_batch_set_launch_shape": (_int, [_vp, _int, _int, _int]),
    "nufi_batch_get_launch_shape": (_int, [_vp, ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(_int)]),
    "nufi_batch_autotune": (_int, [_vp, _u64, _int, ctypes.POINTER(_int), ctypes.POINTER(_int),
                                   ctypes.POINTER(_int), _dp]),
    "nufi_sample_field": (_int, [_vp, _u64, _dp, _u64, _dp, _dp]),
    "nufi_sample_f": (_int, [_vp, _u64, _dp, _dp, _u64, _dp]),
    "nufi_phase_flow": (_int, [_vp, _u64, _dp, _dp, _u64, _dp, _dp]),
    "nufi_history_write": (_int, [ctypes.c_char_p, _int, _hp, _dp, _dp]),
    "nufi_history_read_info": (_int, [ctypes.c_char_p, _hp, ctypes.POINTER(_int)]),
    "nufi_history_read": (_int, [ctypes.c_char_p, _hp, _dp, _u64, _dp, _u64]),
    "nufi_save_history": (_int, [_vp, ctypes.c_char_p, _int, _u64]),
    "nufi_load_history": (_int, [_vp, ctypes.c_char_p, ctypes.POINTER(_u64)]),
}


class NufiError(RuntimeError):
    """A failed C-ABI call (the reference's cuda::exception, cuda_runtime.hpp:46-51)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class NufiRangeError(NufiError, IndexError):
    """n > Nt or an out-of-range node range (std::range_error in the reference)."""


class NufiIOError(NufiError, OSError):
    """A history file that cannot be opened, read, written or parsed."""


_lib = None


def library_path() -> str:
    return LIB_PATH


def load():
    """Load libnufi_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP extension first "
            "(python -c 'import numericalflowiteration_amd.build as b; b.build()')")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, ctx=None, batch: bool = False):
    if rc == OK:
        return
    lib = load()
    msg = (lib.nufi_batch_last_error(ctx) if batch else lib.nufi_last_error(ctx)).decode(errors="replace")
    if rc == ERR_RANGE:
        raise NufiRangeError(rc, msg)
    if rc == ERR_IO:
        raise NufiIOError(rc, msg)
    raise NufiError(rc, msg)


def device_count() -> int:
    n = _int(0)
    rc = load().nufi_device_count(ctypes.byref(n))
    if rc != OK:
        return 0
    return n.value
