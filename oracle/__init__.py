"""ORACLE -- test infrastructure, not part of the product.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker or the timed CPU baseline.

`spgemm` is a host Gustavson SpGEMM (oracle/spgemm_oracle.c, built by
oracle/Makefile into liboracle.so) restating the reference's `proposal`
result: NIP per row (setup.cuh:48-91), exact nnz per row and the exclusive
scan into row_ptr (sym1.cuh:72-98, binning.cuh:219-224), ascending column
order (sym2.cuh:104-120) and values accumulated from 0 (num.cuh:97-121).
`dense_reference` is an independent brute force for small inputs.

Every function here is pinned by tests/test_oracle.py (dense brute force,
27-point-stencil closed forms, the exactly-8-per-row NIP, kept exact zeros);
none is "parity unpinned".
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-C", _HERE, "-B", "liboracle.so"], check=True,
                   stdout=subprocess.DEVNULL)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    P = ctypes.c_void_p
    I64 = ctypes.c_int64
    lib.oracle_nip_rows.argtypes = [I64, P, P, P, P, ctypes.c_int]
    lib.oracle_nip_rows.restype = I64
    lib.oracle_symbolic.argtypes = [I64, I64, P, P, P, P, P, ctypes.c_int]
    lib.oracle_symbolic.restype = I64
    lib.oracle_numeric.argtypes = [I64, I64, P, P, P, P, P, P, P, P, P, P, ctypes.c_int]
    lib.oracle_numeric.restype = ctypes.c_int
    lib.oracle_time_product.argtypes = [I64, I64, P, P, P, P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
    lib.oracle_time_product.restype = ctypes.c_int
    lib.oracle_hw_threads.argtypes = []
    lib.oracle_hw_threads.restype = ctypes.c_int
    _lib = lib
    return lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p) if a.size else None


@dataclass
class OracleResult:
    rpt: np.ndarray      # int64 [m+1]
    col: np.ndarray      # int32 [nnz]
    val: np.ndarray      # float64 [nnz]
    absval: np.ndarray   # float64 [nnz]: (|A||B|)_ij
    nip: np.ndarray      # int64 [m]

    @property
    def nnz(self) -> int:
        return int(self.rpt[-1])


def hw_threads() -> int:
    return int(_load().oracle_hw_threads())


def nip_rows(A, B, threads: int = 0) -> np.ndarray:
    lib = _load()
    a_rpt = np.ascontiguousarray(A.rpt, dtype=np.int32)
    a_col = np.ascontiguousarray(A.col, dtype=np.int32)
    b_rpt = np.ascontiguousarray(B.rpt, dtype=np.int32)
    nip = np.zeros(A.m, dtype=np.int64)
    lib.oracle_nip_rows(A.m, _ptr(a_rpt), _ptr(a_col), _ptr(b_rpt), _ptr(nip), threads)
    return nip


def spgemm(A, B, threads: int = 0) -> OracleResult:
    """C = A.B on the host in FP64 (float32 operands are widened exactly)."""
    if A.n != B.m:
        raise ValueError("A.n != B.m")
    lib = _load()
    a_rpt = np.ascontiguousarray(A.rpt, dtype=np.int32)
    a_col = np.ascontiguousarray(A.col, dtype=np.int32)
    a_val = np.ascontiguousarray(A.val, dtype=np.float64)
    b_rpt = np.ascontiguousarray(B.rpt, dtype=np.int32)
    b_col = np.ascontiguousarray(B.col, dtype=np.int32)
    b_val = np.ascontiguousarray(B.val, dtype=np.float64)
    nip = np.zeros(A.m, dtype=np.int64)
    lib.oracle_nip_rows(A.m, _ptr(a_rpt), _ptr(a_col), _ptr(b_rpt), _ptr(nip), threads)
    rpt = np.zeros(A.m + 1, dtype=np.int64)
    nnz = lib.oracle_symbolic(A.m, B.n, _ptr(a_rpt), _ptr(a_col), _ptr(b_rpt), _ptr(b_col),
                              _ptr(rpt), threads)
    if nnz < 0:
        raise MemoryError("oracle symbolic phase failed")
    col = np.empty(nnz, dtype=np.int32)
    val = np.empty(nnz, dtype=np.float64)
    absval = np.empty(nnz, dtype=np.float64)
    st = lib.oracle_numeric(A.m, B.n, _ptr(a_rpt), _ptr(a_col), _ptr(a_val), _ptr(b_rpt),
                            _ptr(b_col), _ptr(b_val), _ptr(rpt), _ptr(col), _ptr(val),
                            _ptr(absval), threads)
    if st != 0:
        raise RuntimeError(f"oracle numeric phase failed ({st})")
    return OracleResult(rpt, col, val, absval, nip)


def time_product(A, B, threads: int, warmups: int, runs: int) -> np.ndarray:
    """Wall-clock seconds of `runs` timed host products (count pass, scan,
    sorted columns and values into fresh arrays) after `warmups` untimed
    ones: the CPU baseline of the benchmark."""
    lib = _load()
    a_rpt = np.ascontiguousarray(A.rpt, dtype=np.int32)
    a_col = np.ascontiguousarray(A.col, dtype=np.int32)
    a_val = np.ascontiguousarray(A.val, dtype=np.float64)
    b_rpt = np.ascontiguousarray(B.rpt, dtype=np.int32)
    b_col = np.ascontiguousarray(B.col, dtype=np.int32)
    b_val = np.ascontiguousarray(B.val, dtype=np.float64)
    times = np.zeros(max(runs, 1), dtype=np.float64)
    st = lib.oracle_time_product(A.m, B.n, _ptr(a_rpt), _ptr(a_col), _ptr(a_val), _ptr(b_rpt), _ptr(b_col),
                                 _ptr(b_val), threads, warmups, runs, _ptr(times))
    if st != 0:
        raise MemoryError("host product ran out of memory")
    return times[:runs]


def dense_reference(A, B):
    """Independent brute force for small inputs: dense FP64 product and the
    structural pattern (pattern(A) . pattern(B) > 0, which keeps exact-zero
    sums because the pattern product is a count)."""
    def dense(M, ones=False):
        D = np.zeros((M.m, M.n), dtype=np.float64)
        rows = np.repeat(np.arange(M.m), np.diff(np.asarray(M.rpt, dtype=np.int64)))
        D[rows, np.asarray(M.col)] = 1.0 if ones else np.asarray(M.val, dtype=np.float64)
        return D
    values = dense(A) @ dense(B)
    pattern = (dense(A, True) @ dense(B, True)) > 0
    absvals = np.abs(dense(A)) @ np.abs(dense(B))
    return values, pattern, absvals
