"""A few HIP runtime calls through ctypes, for GPU tests that need caller-side
work on the caller's streams (copies queued before or after a library call).
Test plumbing only; the library under test is reached through its own C ABI."""
from __future__ import annotations

import ctypes
import os

_hip = None

H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
        lib = ctypes.CDLL(path)
        P, S = ctypes.c_void_p, ctypes.c_size_t
        for name, args in {
            "hipMalloc": [ctypes.POINTER(P), S], "hipFree": [P],
            "hipHostMalloc": [ctypes.POINTER(P), S, ctypes.c_uint], "hipHostFree": [P],
            "hipMemcpyAsync": [P, P, S, ctypes.c_int, P], "hipMemsetAsync": [P, ctypes.c_int, S, P],
            "hipStreamCreate": [ctypes.POINTER(P)], "hipStreamDestroy": [P],
            "hipStreamSynchronize": [P], "hipDeviceSynchronize": [],
        }.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _hip = lib
    return _hip


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


def malloc(n: int) -> int:
    p = ctypes.c_void_p()
    check(hip().hipMalloc(ctypes.byref(p), n), "hipMalloc")
    return p.value


def host_malloc(n: int) -> int:
    p = ctypes.c_void_p()
    check(hip().hipHostMalloc(ctypes.byref(p), n, 0), "hipHostMalloc")
    return p.value


def copy_async(dst: int, src: int, n: int, kind: int, stream: int = 0) -> None:
    check(hip().hipMemcpyAsync(dst, src, n, kind, stream), "hipMemcpyAsync")


def stream_create() -> int:
    p = ctypes.c_void_p()
    check(hip().hipStreamCreate(ctypes.byref(p)), "hipStreamCreate")
    return p.value


def stream_sync(stream: int) -> None:
    check(hip().hipStreamSynchronize(stream), "hipStreamSynchronize")


def device_sync() -> None:
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
