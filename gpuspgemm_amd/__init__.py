"""MI355X (gfx950) SpGEMM: C = A . B in CSR, keeping the method of the
reference's `proposal` library (flop-count binning, hash-based exact
symbolic count, exclusive-scan row pointers, binary-search accumulation).

Python mirror of the reference's host interface; every call goes through
the C ABI of libgpuspgemm.so (include/gpuspgemm.h):

    proposal(A, B) -> C            proposal::proposal<T>   proposal.cuh:85-124
    get_nip(A, B) -> int           utils::get_nip          nip.cuh:52-78
    HostCSR.load_from_filename     csr.cuh:119-160
    HostCSR.save_to_filename       csr.cuh:162-191
    HostCSR.to_device / DeviceCSR.to_host
                                   CSR::to<Location>()     csr.cuh:193-203
    HostCSR == HostCSR             operator==              csr.cuh:218-239
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import SpgError
from .csr import DeviceCSR, HostCSR, _dtype_code

__all__ = [
    "HostCSR", "DeviceCSR", "SpgError", "proposal", "proposal_batched", "get_nip",
    "device_info", "set_tuning", "get_tuning", "get_variants", "variant_list", "profile", "load_library",
    "Server", "set_graphs", "graph_stats", "graphs", "proposal_async", "set_log_level",
]


def load_library():
    """Load the HIP library (raises ImportError when it is not built)."""
    return _lib.load()


def proposal(A: DeviceCSR, B: DeviceCSR) -> DeviceCSR:
    """C = A . B on the GPU; raises SpgError on failure (e.g. A.n != B.m)."""
    if A.dtype != B.dtype:
        raise TypeError("A and B must have the same value type")
    lib = _lib.load()
    out = _lib.SpgCsr()
    _lib.check(lib.spg_multiply(ctypes.byref(A.raw), ctypes.byref(B.raw), ctypes.byref(out),
                                _dtype_code(A.dtype)))
    return DeviceCSR(out, A.dtype, library_owned=True)


def proposal_async(A: DeviceCSR, B: DeviceCSR, stream: int) -> DeviceCSR:
    """C = A . B ordered on a caller stream (a hipStream_t handle as an int,
    e.g. torch.cuda.Stream().cuda_stream): returns once the fused pass is
    queued; C's contents are complete when the stream reaches the call."""
    if A.dtype != B.dtype:
        raise TypeError("A and B must have the same value type")
    lib = _lib.load()
    out = _lib.SpgCsr()
    _lib.check(lib.spg_multiply_async(ctypes.byref(A.raw), ctypes.byref(B.raw), ctypes.byref(out),
                                      _dtype_code(A.dtype), ctypes.c_void_p(stream)))
    return DeviceCSR(out, A.dtype, library_owned=True)


def set_log_level(level) -> None:
    """Library log level: a name (trace, debug, info, warn, error, critical,
    off) or 0..6; debug prints a line per phase of every multiply to stderr."""
    names = ["trace", "debug", "info", "warn", "error", "critical", "off"]
    lv = names.index(level) if isinstance(level, str) else int(level)
    _lib.check(_lib.load().spg_set_log_level(lv))


def proposal_batched(As: Sequence[DeviceCSR], Bs: Sequence[DeviceCSR]) -> list[DeviceCSR]:
    """C_i = A_i . B_i for every pair, in one pass of the kernels."""
    if len(As) != len(Bs) or not As:
        raise ValueError("need equally many, and at least one, A and B")
    dt = As[0].dtype
    if any(x.dtype != dt for x in list(As) + list(Bs)):
        raise TypeError("all operands must have the same value type")
    lib = _lib.load()
    n = len(As)
    arr_a = (_lib.SpgCsr * n)(*[a.raw for a in As])
    arr_b = (_lib.SpgCsr * n)(*[b.raw for b in Bs])
    arr_c = (_lib.SpgCsr * n)()
    _lib.check(lib.spg_multiply_batched(n, arr_a, arr_b, arr_c, _dtype_code(dt)))
    return [DeviceCSR(_lib.SpgCsr(c.nnz, c.m, c.n, c.rpt, c.col, c.val), dt, library_owned=True)
            for c in arr_c]


def get_nip(A: DeviceCSR, B: DeviceCSR) -> int:
    lib = _lib.load()
    out = ctypes.c_int64()
    _lib.check(lib.spg_get_nip(ctypes.byref(A.raw), ctypes.byref(B.raw), ctypes.byref(out)))
    return int(out.value)


def device_info() -> dict:
    lib = _lib.load()
    info = _lib.SpgDeviceInfo()
    _lib.check(lib.spg_device_info_get(ctypes.byref(info)))
    return {"arch": info.arch.decode(), "compute_units": info.compute_units,
            "lds_per_block": info.lds_per_block, "clock_khz": info.clock_khz,
            "global_mem_bytes": info.global_mem_bytes}


def _text(fn, phase: int) -> str:
    buf = ctypes.create_string_buffer(8192)
    _lib.check(fn(phase, buf, len(buf)))
    return buf.value.decode()


def set_tuning(phase: int, table: str | None) -> None:
    """Bin table of phase 0 (symbolic count) or 1 (fused pass); None resets."""
    lib = _lib.load()
    _lib.check(lib.spg_set_tuning(phase, table.encode() if table else None))


def get_tuning(phase: int) -> str:
    return _text(_lib.load().spg_get_tuning, phase)


def get_variants(phase: int) -> str:
    return _text(_lib.load().spg_get_variants, phase)


def variant_list(phase: int) -> list[dict]:
    """Compiled kernel variants of a phase as dicts: id, wg, rows, lanes
    (per row), table (phase 0) or capacity (phase 1), chunk, slots, and the
    flags max (phase 0's overflow-detecting variant) / global (phase 1's
    global-memory variant), and merged: the merged launch set (0 small
    rows, 1 longer rows) whose bins share one launch, or None."""
    out = []
    for line in get_variants(phase).splitlines():
        parts = line.split()
        v = {"id": int(parts[0]), "max": "max" in parts[1:], "global": "global" in parts[1:], "merged": None}
        for kv in parts[1:]:
            if "=" in kv:
                k, x = kv.split("=")
                v[k] = int(x)
        if "wg" in v:
            v["lanes"] = v["wg"] // v["rows"]
        out.append(v)
    return out


def set_graphs(enabled: bool) -> None:
    """Graph replay of repeated calls (spg_set_graphs): the second call
    with the same launch signature captures each phase into a HIP graph,
    later ones replay it."""
    _lib.check(_lib.load().spg_set_graphs(1 if enabled else 0))


def graph_stats() -> dict:
    """Graphs captured and phase replays since the context was created."""
    cap, rep = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(_lib.load().spg_graph_stats(ctypes.byref(cap), ctypes.byref(rep)))
    return {"captured": cap.value, "replayed": rep.value}


class graphs:
    """Context manager turning graph replay on: `with graphs(): ...`."""

    def __enter__(self):
        set_graphs(True)
        return self

    def __exit__(self, *exc):
        set_graphs(False)


class profile:
    """Context manager collecting per-kernel HIP-event timings:
    `with profile() as p: proposal(A, B)` then `p.records`."""

    def __enter__(self):
        self.lib = _lib.load()
        _lib.check(self.lib.spg_set_profiling(1))
        self.records = []
        return self

    def collect(self) -> list[dict]:
        recs = []
        for i in range(self.lib.spg_profile_count()):
            name = ctypes.create_string_buffer(64)
            rows, grid, ms = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_double()
            _lib.check(self.lib.spg_profile_get(i, name, 64, ctypes.byref(rows), ctypes.byref(grid),
                                                ctypes.byref(ms)))
            recs.append({"name": name.value.decode(), "rows": rows.value, "grid": grid.value, "ms": ms.value})
        self.records = recs
        return recs

    def __exit__(self, *exc):
        self.collect()
        _lib.check(self.lib.spg_set_profiling(0))


class Server:
    """Served form: a request loop in the library (spg_server_*) with
    continuous batching and admission by exact output memory.

        with Server(budget_bytes=8 << 30, max_batch=64) as srv:
            t = srv.submit(dA, dB)          # device-resident operands
            C, latency_ms = srv.wait(t)     # C: DeviceCSR owned by the caller

    Operands are kept referenced until their request is waited for."""

    def __init__(self, budget_bytes: int, max_batch: int = 64):
        self.lib = _lib.load()
        self.handle = ctypes.c_void_p()
        _lib.check(self.lib.spg_server_create(int(budget_bytes), int(max_batch), ctypes.byref(self.handle)))
        self._keep: dict[int, tuple] = {}

    def submit(self, A: DeviceCSR, B: DeviceCSR) -> int:
        if A.dtype != B.dtype:
            raise TypeError("A and B must have the same value type")
        t = ctypes.c_int64()
        _lib.check(self.lib.spg_server_submit(self.handle, ctypes.byref(A.raw), ctypes.byref(B.raw),
                                              _dtype_code(A.dtype), ctypes.byref(t)))
        self._keep[t.value] = (A, B)
        return t.value

    def wait(self, ticket: int) -> tuple[DeviceCSR, float]:
        out = _lib.SpgCsr()
        lat = ctypes.c_double()
        A, _ = self._keep.pop(ticket)
        _lib.check(self.lib.spg_server_wait(self.handle, ticket, ctypes.byref(out), ctypes.byref(lat)))
        return DeviceCSR(out, A.dtype, library_owned=True), lat.value

    def stats(self) -> dict:
        v = [ctypes.c_int64() for _ in range(5)]
        _lib.check(self.lib.spg_server_stats(self.handle, *[ctypes.byref(x) for x in v]))
        return dict(zip(("batches", "served", "max_batch", "deferred", "peak_bytes"), (x.value for x in v)))

    def close(self) -> None:
        if self.handle:
            _lib.check(self.lib.spg_server_destroy(self.handle))
            self.handle = ctypes.c_void_p()
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
