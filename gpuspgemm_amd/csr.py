"""CSR containers and the `.csr` binary format, mirroring utils::CSR.

Reference: src/libutils/include/utils/csr.cuh
  * members nnz, m, n, rpt, col, val (csr.cuh:23-31)
  * load_from_filename / save_to_filename (csr.cuh:119-191): little-endian
    int32 nnz, int32 m, int32 n, int32 rpt[m+1], int32 col[nnz], then
    float64 val[nnz] whatever the value type (float32 narrows on load and
    widens on save); a file shorter or longer than its header says is an
    error
  * to<Location>() (csr.cuh:193-203)
  * operator== (csr.cuh:218-239): equal sizes, identical rpt and col, values
    equal or within atol 1e-8 + rtol 1e-5 * max(|l|, |r|)
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass
class HostCSR:
    m: int
    n: int
    rpt: np.ndarray  # int32 [m+1]
    col: np.ndarray  # int32 [nnz]
    val: np.ndarray  # float64 / float32 [nnz]

    @property
    def nnz(self) -> int:
        return int(self.col.shape[0])

    @property
    def dtype(self):
        return self.val.dtype

    def astype(self, dtype) -> "HostCSR":
        return HostCSR(self.m, self.n, self.rpt, self.col, self.val.astype(dtype))

    # ---- .csr format (csr.cuh:119-191)
    @staticmethod
    def load_from_filename(filename: str, dtype=np.float64) -> "HostCSR":
        try:
            size = os.path.getsize(filename)
            fp = open(filename, "rb")
        except OSError as exc:
            raise RuntimeError(f"Failed to open file {filename}") from exc
        with fp:
            if size < 12:
                raise RuntimeError(f"File {filename} is too small")
            nnz, m, n = (int(x) for x in np.frombuffer(fp.read(12), dtype="<i4"))
            if nnz < 0 or m < 0 or n < 0:
                raise RuntimeError(f"Failed to read file {filename}")
            need = 12 + 4 * (m + 1) + 4 * nnz + 8 * nnz
            if size < need:
                raise RuntimeError(f"Failed to read file {filename}")
            if size > need:
                raise RuntimeError(f"File {filename} is too large")
            rpt = np.frombuffer(fp.read(4 * (m + 1)), dtype="<i4").astype(np.int32)
            col = np.frombuffer(fp.read(4 * nnz), dtype="<i4").astype(np.int32)
            val = np.frombuffer(fp.read(8 * nnz), dtype="<f8").astype(dtype)
        return HostCSR(m, n, rpt, col, val)

    def save_to_filename(self, filename: str) -> None:
        try:
            fp = open(filename, "wb")
        except OSError as exc:
            raise RuntimeError(f"Failed to open file {filename}") from exc
        with fp:
            fp.write(np.array([self.nnz, self.m, self.n], dtype="<i4").tobytes())
            fp.write(np.ascontiguousarray(self.rpt, dtype="<i4").tobytes())
            fp.write(np.ascontiguousarray(self.col, dtype="<i4").tobytes())
            fp.write(np.ascontiguousarray(self.val, dtype="<f8").tobytes())

    def to_device(self) -> "DeviceCSR":
        return DeviceCSR.from_host(self)

    # ---- equality of csr.cuh:218-239
    def __eq__(self, other) -> bool:
        if not isinstance(other, HostCSR):
            return NotImplemented
        if self.nnz != other.nnz or self.m != other.m or self.n != other.n:
            return False
        if not np.array_equal(self.rpt, other.rpt) or not np.array_equal(self.col, other.col):
            return False
        l = self.val.astype(np.float64)
        r = other.val.astype(np.float64)
        ok = (l == r) | (np.abs(l - r) <= 1e-8 + 1e-5 * np.maximum(np.abs(l), np.abs(r)))
        return bool(np.all(ok))


def _dtype_code(dtype) -> int:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return _lib.SPG_F64
    if dt == np.float32:
        return _lib.SPG_F32
    raise TypeError(f"unsupported value type {dt}")


class DeviceCSR:
    """Device CSR owning its arrays (freed on `free()` or garbage collection).

    `library_owned` matrices come from the multiply and are released with
    spg_csr_free; uploaded ones with spg_free."""

    def __init__(self, raw: _lib.SpgCsr, dtype, library_owned: bool):
        self.raw = raw
        self.dtype = np.dtype(dtype)
        self._library_owned = library_owned

    # members of utils::CSR
    nnz = property(lambda self: self.raw.nnz)
    m = property(lambda self: self.raw.m)
    n = property(lambda self: self.raw.n)

    @staticmethod
    def from_host(h: HostCSR) -> "DeviceCSR":
        lib = _lib.load()
        dt = np.dtype(h.val.dtype)
        _dtype_code(dt)
        raw = _lib.SpgCsr(h.nnz, h.m, h.n, None, None, None)
        arrays = (
            ("rpt", np.ascontiguousarray(h.rpt, dtype=np.int32)),
            ("col", np.ascontiguousarray(h.col, dtype=np.int32)),
            ("val", np.ascontiguousarray(h.val, dtype=dt)),
        )
        out = DeviceCSR(raw, dt, library_owned=False)
        for name, arr in arrays:
            ptr = ctypes.c_void_p()
            _lib.check(lib.spg_malloc(ctypes.byref(ptr), arr.nbytes))
            setattr(raw, name, ptr.value)
            if arr.nbytes:
                _lib.check(lib.spg_memcpy(ptr, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes))
        return out

    def to_host(self) -> HostCSR:
        lib = _lib.load()
        rpt = np.empty(self.m + 1, dtype=np.int32)
        col = np.empty(self.nnz, dtype=np.int32)
        val = np.empty(self.nnz, dtype=self.dtype)
        for arr, ptr in ((rpt, self.raw.rpt), (col, self.raw.col), (val, self.raw.val)):
            if arr.nbytes:
                _lib.check(lib.spg_memcpy(arr.ctypes.data_as(ctypes.c_void_p), ptr, arr.nbytes))
        return HostCSR(self.m, self.n, rpt, col, val)

    def free(self) -> None:
        if self.raw is None:
            return
        lib = _lib.load()
        if self._library_owned:
            _lib.check(lib.spg_csr_free(ctypes.byref(self.raw)))
        else:
            for name in ("rpt", "col", "val"):
                p = getattr(self.raw, name)
                if p:
                    _lib.check(lib.spg_free(p))
        self.raw = None

    def free_async(self, stream: int) -> None:
        """Release a library-owned matrix once `stream` reaches this point."""
        if self.raw is None:
            return
        if not self._library_owned:
            raise ValueError("only matrices returned by a multiply are released on a stream")
        _lib.check(_lib.load().spg_csr_free_async(ctypes.byref(self.raw), ctypes.c_void_p(stream)))
        self.raw = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()
