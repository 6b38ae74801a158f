"""Streams and threads on the GPU: plain multiplies from several host
threads at once, ordering with the caller's default stream (inputs written
by a kernel queued just before the call, outputs still read by a kernel when
they are released and the next multiply runs), the caller-stream entry
point, and the library's debug log.  Caller-side work (long device copies
that delay what follows them on a stream) is queued through the HIP runtime
directly (tests/hip_host.py)."""
import ctypes
import threading

import numpy as np
import pytest

import gpuspgemm_amd as g
from gpuspgemm_amd import generators as gen
from gpuspgemm_amd.csr import DeviceCSR, HostCSR
import oracle
import hip_host as hh
from parity import compare, gpu_multiply

BIG = 1 << 31  # bytes of the delaying device copy (two buffers)

pytestmark = pytest.mark.gpu


def test_threads_multiply_at_once(gpu_lib):
    """Four host threads, each multiplying its own products repeatedly: every
    result matches the oracle."""
    mats = [gen.powerlaw(m=40_000, seed=70 + i) for i in range(2)] + \
           [gen.stencil27(g=20, seed=72), gen.uniform8(m=4096, seed=73).astype(np.float32)]
    refs = [oracle.spgemm(A, A) for A in mats]
    dev = [A.to_device() for A in mats]
    errors = []

    def work(i):
        try:
            for _ in range(6):
                dC = g.proposal(dev[i], dev[i])
                C = dC.to_host()
                dC.free()
                compare(mats[i], mats[i], C, refs[i])
        except Exception as e:  # reported by the main thread
            errors.append((i, repr(e)))

    try:
        ts = [threading.Thread(target=work, args=(i,)) for i in range(len(mats))]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
    finally:
        for d in dev:
            d.free()
    assert not errors, errors


def _slow_copy_on(stream, bufs):
    """Queue a device copy of BIG bytes on `stream`: work queued after it on
    the same stream starts a millisecond or so later."""
    hh.copy_async(bufs[1], bufs[0], BIG, hh.D2D, stream)


@pytest.fixture(scope="module")
def big_buffers(gpu_lib):
    bufs = [hh.malloc(BIG), hh.malloc(BIG)]
    yield bufs
    hh.device_sync()
    for b in bufs:
        hh.hip().hipFree(b)


def test_input_written_just_before_the_call(gpu_lib, big_buffers):
    """A's values are written by a copy queued on the legacy default stream
    behind a long device copy, right before spg_multiply: the multiply
    (ordered on the thread's default stream, which follows the legacy
    stream) reads the new values."""
    A = gen.powerlaw(m=30_000, seed=75)
    new_val = np.ascontiguousarray(A.val * 3.0 - 0.5)
    dA = A.to_device()
    host = hh.host_malloc(new_val.nbytes)
    try:
        ctypes_buf = (ctypes.c_char * new_val.nbytes).from_address(host)
        ctypes.memmove(ctypes_buf, new_val.ctypes.data, new_val.nbytes)
        hh.device_sync()
        _slow_copy_on(0, big_buffers)
        hh.copy_async(dA.raw.val, host, new_val.nbytes, hh.H2D, 0)
        dC = g.proposal(dA, dA)
        C = dC.to_host()
        dC.free()
        A2 = HostCSR(A.m, A.n, A.rpt, A.col, new_val)
        compare(A2, A2, C)
    finally:
        hh.device_sync()
        hh.hip().hipHostFree(host)
        dA.free()


def test_output_read_while_released(gpu_lib, big_buffers):
    """C is released while a copy of its values queued on the legacy stream
    behind a long device copy has not run yet, and products of the same size
    with other values follow at once: the copy still reads C's values (its
    blocks are not handed out again before the queued work is done)."""
    A = gen.powerlaw(m=60_000, seed=76)
    A2 = HostCSR(A.m, A.n, A.rpt, A.col, A.val * 2.0)  # same structure, 4x the values
    ref = oracle.spgemm(A, A)
    dA, dA2 = A.to_device(), A2.to_device()
    host = 0
    try:
        dC = g.proposal(dA, dA)
        n = dC.nnz
        host = hh.host_malloc(8 * n)
        hh.device_sync()
        _slow_copy_on(0, big_buffers)
        hh.copy_async(host, dC.raw.val, 8 * n, hh.D2H, 0)
        dC.free()
        later = [g.proposal(dA2, dA2) for _ in range(3)]
        hh.device_sync()
        got = np.frombuffer((ctypes.c_char * (8 * n)).from_address(host), dtype=np.float64).copy()
        for d in later:
            d.free()
        scale = np.where(ref.absval > 0, ref.absval, 1.0)
        assert float((np.abs(got - ref.val) / scale).max()) <= 1e-12
    finally:
        hh.device_sync()
        if host:
            hh.hip().hipHostFree(host)
        dA.free()
        dA2.free()


def test_async_on_a_caller_stream(gpu_lib, big_buffers):
    """spg_multiply_async on two caller streams, three products each, queued
    behind a long copy on one of them: every C is complete once its stream
    is synchronised, and release on the stream works."""
    mats = [gen.powerlaw(m=50_000, seed=77), gen.stencil27(g=18, seed=78)]
    refs = [oracle.spgemm(A, A) for A in mats]
    dev = [A.to_device() for A in mats]
    streams = [hh.stream_create(), hh.stream_create()]
    try:
        _slow_copy_on(streams[0], big_buffers)
        outs = []
        for rep in range(3):
            for i, st in enumerate(streams):
                outs.append((i, g.proposal_async(dev[i], dev[i], st)))
        for st in streams:
            hh.stream_sync(st)
        for i, dC in outs:
            compare(mats[i], mats[i], dC.to_host(), refs[i])
            dC.free_async(streams[i])
        for st in streams:
            hh.stream_sync(st)
    finally:
        hh.device_sync()
        for st in streams:
            hh.hip().hipStreamDestroy(st)
        for d in dev:
            d.free()


def test_debug_log(gpu_lib, capfd):
    """At debug level every multiply logs its symbolic and fused phases
    (rows per bin, nnz(C), the fail-bin size)."""
    A = gen.powerlaw(m=20_000, seed=79)
    g.set_log_level("debug")
    try:
        C = gpu_multiply(A, A)
    finally:
        g.set_log_level("off")
    err = capfd.readouterr().err
    assert "symbolic phase" in err and f"nnz(C) {C.nnz}" in err and "fail bin" in err, err[-2000:]
    assert "fused pass done" in err
