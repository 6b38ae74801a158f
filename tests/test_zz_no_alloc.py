"""spg_multiply_noalloc on the GPU, called through the C ABI: C goes into
caller arrays and scratch comes from the library's fixed device array.
Two calls in a row, the second with a smaller input, each compared with
spg_multiply's result on the same operands (structure bit-exact, values
within the |A||B|-scaled tolerance: FP64 1e-12, FP32 1e-5) and with the
oracle; then the error paths: an output capacity one short of nnz(C), and
a product whose scratch exceeds the fixed array, both refused with
out-of-memory, after which the entry point still works."""
import ctypes

import numpy as np
import pytest

from gpuspgemm_amd import _lib
from gpuspgemm_amd import generators as gen
from gpuspgemm_amd.csr import DeviceCSR, HostCSR, _dtype_code
import oracle
from parity import TOL, compare, gpu_multiply

pytestmark = pytest.mark.gpu

CODE = {name: code for code, name in _lib.STATUS.items()}
OK, OOM = CODE["SPG_OK"], CODE["SPG_ERR_OUT_OF_MEMORY"]


class CallerArrays:
    """C's arrays, allocated by the caller before the call."""

    def __init__(self, m: int, capacity: int, dtype):
        lib = _lib.load()
        self.dtype = np.dtype(dtype)
        self.ptrs = []
        for nbytes in (4 * (m + 1), 4 * max(capacity, 1), self.dtype.itemsize * max(capacity, 1)):
            p = ctypes.c_void_p()
            _lib.check(lib.spg_malloc(ctypes.byref(p), nbytes))
            self.ptrs.append(p.value)
        self.capacity = capacity

    def csr(self) -> _lib.SpgCsr:
        return _lib.SpgCsr(0, 0, 0, *self.ptrs)

    def free(self):
        lib = _lib.load()
        for p in self.ptrs:
            _lib.check(lib.spg_free(p))
        self.ptrs = []


def noalloc(dA: DeviceCSR, dB: DeviceCSR, arrays: CallerArrays, capacity: int):
    """(status, C raw) of one spg_multiply_noalloc call."""
    lib = _lib.load()
    C = arrays.csr()
    st = lib.spg_multiply_noalloc(ctypes.byref(dA.raw), ctypes.byref(dB.raw), ctypes.byref(C),
                                  ctypes.c_int64(capacity), _dtype_code(dA.dtype))
    return st, C


def to_host(C: _lib.SpgCsr, dtype) -> HostCSR:
    # a view of the caller's arrays; DeviceCSR.to_host copies them out
    view = DeviceCSR(_lib.SpgCsr(C.nnz, C.m, C.n, C.rpt, C.col, C.val), dtype, library_owned=False)
    try:
        return view.to_host()
    finally:
        view.raw = None  # the arrays stay the caller's


def same_as_plain(A: HostCSR, C: HostCSR, plain: HostCSR, ref) -> None:
    """C equals spg_multiply's product: structure bit-exact, values within
    the tolerance of the |A||B| scale; both match the oracle."""
    assert C.nnz == plain.nnz
    np.testing.assert_array_equal(C.rpt, plain.rpt)
    np.testing.assert_array_equal(C.col, plain.col)
    diff = np.abs(C.val.astype(np.float64) - plain.val.astype(np.float64))
    scale = np.where(ref.absval > 0, ref.absval, 1.0)
    assert (float((diff / scale).max()) if diff.size else 0.0) <= TOL[np.dtype(C.val.dtype)]
    compare(A, A, C, ref)


def run_pair(A: HostCSR) -> None:
    ref = oracle.spgemm(A, A)
    plain = gpu_multiply(A, A)
    dA = A.to_device()
    arrays = CallerArrays(A.m, ref.nnz, A.val.dtype)
    try:
        st, C = noalloc(dA, dA, arrays, ref.nnz)
        assert st == OK, _lib.load().spg_last_error().decode()
        assert (C.m, C.n, C.nnz) == (A.m, A.n, ref.nnz)
        assert (C.rpt, C.col, C.val) == tuple(arrays.ptrs)
        same_as_plain(A, to_host(C, A.val.dtype), plain, ref)
    finally:
        arrays.free()
        dA.free()


def test_two_calls_second_smaller(gpu_lib):
    """The power-law workload at full size (1,000,000 rows), then a smaller
    power-law product, then a smaller FP32 stencil product: the second and
    third reuse the scratch the first left behind."""
    print("noalloc: powerlaw m=1e6 f64, then m=60000 f64, then stencil g=24 f32", flush=True)
    run_pair(gen.powerlaw(m=1_000_000))
    run_pair(gen.powerlaw(m=60_000, seed=11))
    run_pair(gen.stencil27(g=24, seed=12).astype(np.float32))


def test_capacity_and_scratch_limits(gpu_lib):
    """A capacity one short of nnz(C) fails with out-of-memory and reports
    the nnz needed; a product whose scratch (about 124 bytes per row of A)
    exceeds the fixed array is refused before any launch; the next call
    succeeds."""
    print("noalloc: capacity one short, then 2.5e6-row identity over the scratch limit", flush=True)
    lib = _lib.load()
    A = gen.powerlaw(m=20_000, seed=13)
    ref = oracle.spgemm(A, A)
    dA = A.to_device()
    arrays = CallerArrays(A.m, ref.nnz, np.float64)
    try:
        st, C = noalloc(dA, dA, arrays, ref.nnz - 1)
        assert st == OOM
        assert C.nnz == ref.nnz
        assert "capacity" in lib.spg_last_error().decode()
        m = 2_500_000  # the diagonal: nothing but the per-row scratch is large
        I = HostCSR(m, m, np.arange(m + 1, dtype=np.int32), np.arange(m, dtype=np.int32), np.ones(m))
        dI = I.to_device()
        big = CallerArrays(m, m, np.float64)
        try:
            st, _ = noalloc(dI, dI, big, m)
            assert st == OOM
            assert "fixed scratch" in lib.spg_last_error().decode()
        finally:
            big.free()
            dI.free()
        st, C = noalloc(dA, dA, arrays, ref.nnz)
        assert st == OK, lib.spg_last_error().decode()
        compare(A, A, to_host(C, np.float64), ref)
    finally:
        arrays.free()
        dA.free()
