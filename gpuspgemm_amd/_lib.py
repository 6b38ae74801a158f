"""ctypes binding of the C ABI declared in include/gpuspgemm.h.

The shared library is loaded from the package directory; when it is
missing the import fails loudly (there is no CPU fallback of the product
path).
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libgpuspgemm.so")


class SpgCsr(ctypes.Structure):
    _fields_ = [
        ("nnz", ctypes.c_int32),
        ("m", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("rpt", ctypes.c_void_p),
        ("col", ctypes.c_void_p),
        ("val", ctypes.c_void_p),
    ]


class SpgDeviceInfo(ctypes.Structure):
    _fields_ = [
        ("arch", ctypes.c_char * 64),
        ("compute_units", ctypes.c_int32),
        ("lds_per_block", ctypes.c_int32),
        ("clock_khz", ctypes.c_int32),
        ("global_mem_bytes", ctypes.c_int64),
    ]


SPG_F32, SPG_F64 = 0, 1
STATUS = {
    0: "SPG_OK", 1: "SPG_ERR_INVALID_ARGUMENT", 2: "SPG_ERR_OUT_OF_MEMORY", 3: "SPG_ERR_HIP",
    4: "SPG_ERR_OVERFLOW", 5: "SPG_ERR_UNSUPPORTED_DEVICE", 6: "SPG_ERR_INTERNAL",
}

# name -> (restype, argtypes); every function declared in include/gpuspgemm.h
_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_SIG = {
    "spg_init": (_I32, []),
    "spg_multiply": (_I32, [ctypes.POINTER(SpgCsr), ctypes.POINTER(SpgCsr), ctypes.POINTER(SpgCsr), _I32]),
    "spg_multiply_batched": (_I32, [_I32, ctypes.POINTER(SpgCsr), ctypes.POINTER(SpgCsr),
                                    ctypes.POINTER(SpgCsr), _I32]),
    "spg_multiply_async": (_I32, [ctypes.POINTER(SpgCsr), ctypes.POINTER(SpgCsr), ctypes.POINTER(SpgCsr), _I32,
                                  _P]),
    "spg_multiply_noalloc": (_I32, [ctypes.POINTER(SpgCsr), ctypes.POINTER(SpgCsr), ctypes.POINTER(SpgCsr),
                                    ctypes.c_int64, _I32]),
    "spg_csr_free": (_I32, [ctypes.POINTER(SpgCsr)]),
    "spg_csr_free_async": (_I32, [ctypes.POINTER(SpgCsr), _P]),
    "spg_set_log_level": (_I32, [_I32]),
    "spg_get_nip": (_I32, [ctypes.POINTER(SpgCsr), ctypes.POINTER(SpgCsr), ctypes.POINTER(ctypes.c_int64)]),
    "spg_malloc": (_I32, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]),
    "spg_free": (_I32, [_P]),
    "spg_memcpy": (_I32, [_P, _P, ctypes.c_size_t]),
    "spg_device_sync": (_I32, []),
    "spg_status_str": (ctypes.c_char_p, [_I32]),
    "spg_last_error": (ctypes.c_char_p, []),
    "spg_device_info_get": (_I32, [ctypes.POINTER(SpgDeviceInfo)]),
    "spg_set_tuning": (_I32, [_I32, ctypes.c_char_p]),
    "spg_server_create": (_I32, [ctypes.c_size_t, _I32, ctypes.POINTER(ctypes.c_void_p)]),
    "spg_server_submit": (_I32, [_P, ctypes.POINTER(SpgCsr), ctypes.POINTER(SpgCsr), _I32,
                                 ctypes.POINTER(ctypes.c_int64)]),
    "spg_server_wait": (_I32, [_P, ctypes.c_int64, ctypes.POINTER(SpgCsr), ctypes.POINTER(ctypes.c_double)]),
    "spg_server_stats": (_I32, [_P] + [ctypes.POINTER(ctypes.c_int64)] * 5),
    "spg_server_destroy": (_I32, [_P]),
    "spg_get_tuning": (_I32, [_I32, ctypes.c_char_p, ctypes.c_size_t]),
    "spg_get_variants": (_I32, [_I32, ctypes.c_char_p, ctypes.c_size_t]),
    "spg_set_graphs": (_I32, [_I32]),
    "spg_graph_stats": (_I32, [ctypes.POINTER(ctypes.c_int64)] * 2),
    "spg_set_profiling": (_I32, [_I32]),
    "spg_profile_count": (_I32, []),
    "spg_profile_get": (_I32, [_I32, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int64),
                               ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)]),
}

_lib = None


class SpgError(RuntimeError):
    def __init__(self, status: int, message: str):
        self.status = status
        self.status_name = STATUS.get(status, str(status))
        super().__init__(f"{self.status_name}: {message}")


def available() -> bool:
    """Whether the shared library has been built (loading it needs no GPU)."""
    return os.path.exists(LIB_PATH)


def load() -> ctypes.CDLL:
    """Load libgpuspgemm.so (raises ImportError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing; build it with `python -m gpuspgemm_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIG.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != 0:
        lib = load()
        raise SpgError(status, lib.spg_last_error().decode(errors="replace"))
