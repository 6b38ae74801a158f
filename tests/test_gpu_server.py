"""Served form on the GPU (spg_server_*): every result is the same alone
(one plain multiply) and in the mix (submitted among requests of other
sizes, generators and value types, batched and overlapped by the server),
structure bit-exact and values within the oracle tolerance."""
import numpy as np
import pytest

import gpuspgemm_amd as g
from gpuspgemm_amd import generators as gen

from parity import compare

pytestmark = pytest.mark.gpu


def _mix():
    out = [
        gen.uniform8(m=2048, seed=1),
        gen.powerlaw(m=20_000, seed=2),
        gen.stencil27(g=12, seed=3),
        gen.random_csr(300, 300, 0.02, seed=4, empty_rows=0.2),
        gen.powerlaw(m=4000, seed=5).astype(np.float32),
        gen.uniform8(m=1024, seed=6).astype(np.float32),
    ]
    pairs = [(A, A) for A in out]
    a, b = gen.batched_instance(7, m=3000)
    pairs.append((a, b))
    pairs.append((a.astype(np.float32), b.astype(np.float32)))
    return pairs


def _plain(dA, dB):
    dC = g.proposal(dA, dB)
    try:
        return dC.to_host()
    finally:
        dC.free()


def test_results_same_alone_and_in_the_mix(gpu_lib):
    pairs = _mix() * 3  # every product three times, interleaved
    dev = [(A.to_device(), B.to_device()) for A, B in pairs]
    try:
        alone = [_plain(dA, dB) for dA, dB in dev]
        with g.Server(budget_bytes=1 << 30, max_batch=5) as srv:
            tickets = [srv.submit(dA, dB) for dA, dB in dev]
            mixed = []
            for t in tickets:
                dC, ms = srv.wait(t)
                assert ms >= 0
                mixed.append(dC.to_host())
                dC.free()
            st = srv.stats()
        assert st["served"] == len(pairs)
        assert st["batches"] < len(pairs)  # requests were batched
        for (A, B), a, m in zip(pairs, alone, mixed):
            assert np.array_equal(a.rpt, m.rpt) and np.array_equal(a.col, m.col)
            compare(A, B, m)
    finally:
        for dA, dB in dev:
            dA.free()
            dB.free()


def test_admission_by_exact_output_memory(gpu_lib):
    """A budget of about one output admits one request at a time; a budget
    below one output fails the request with out-of-memory."""
    A = gen.powerlaw(m=30_000, seed=11)
    dA = A.to_device()
    try:
        C = _plain(dA, dA)
        need = C.nnz * 12
        budget = need + need // 2
        with g.Server(budget_bytes=budget, max_batch=4) as srv:
            ts = [srv.submit(dA, dA) for _ in range(4)]
            for t in ts:
                dC, _ = srv.wait(t)
                h = dC.to_host()
                dC.free()
                assert np.array_equal(h.col, C.col)
            st = srv.stats()
        # never two outputs at once; a batch of two or more is split
        assert st["served"] == 4 and st["max_batch"] == 1 and st["peak_bytes"] == need
        with g.Server(budget_bytes=need // 2, max_batch=2) as srv:
            t = srv.submit(dA, dA)
            with pytest.raises(g.SpgError, match="OUT_OF_MEMORY"):
                srv.wait(t)
    finally:
        dA.free()


def test_plain_calls_beside_a_server(gpu_lib):
    """Plain multiplies and copies from the calling thread while the server's
    lanes run requests: every result of both is right, and nothing waits
    forever on the library's lock."""
    import oracle
    A = gen.powerlaw(m=20_000, seed=31)
    U = gen.uniform8(m=2048, seed=32)
    ref_a, ref_u = oracle.spgemm(A, A), oracle.spgemm(U, U)
    dA, dU = A.to_device(), U.to_device()
    try:
        with g.Server(budget_bytes=1 << 30, max_batch=4) as srv:
            tickets = [srv.submit(dA, dA) for _ in range(12)]
            for _ in range(6):
                compare(U, U, _plain(dU, dU), ref_u)
            for t in tickets:
                dC, _ = srv.wait(t)
                h = dC.to_host()
                dC.free()
                compare(A, A, h, ref_a)
    finally:
        dA.free()
        dU.free()


def test_served_workload_sample(gpu_lib):
    """The first 32 requests of the served workload, drawn exactly as the
    benchmark draws them (all four generators, FP64 and FP32, rows up to
    2^20), submitted to a server configured like the benchmark's: every
    served result has the structure of its plain result and both match the
    oracle."""
    import oracle
    reqs = [gen.served_request(k) for k in range(32)]
    kinds = {q["kind"] for q in reqs}
    assert kinds == {"uniform8", "powerlaw", "stencil", "batched"}
    assert {q["dtype"] for q in reqs} == {"float64", "float32"}
    assert max(q["m"] for q in reqs) > 800_000
    dev = [(q["A"].to_device(), q["B"].to_device()) for q in reqs]
    try:
        alone = [_plain(dA, dB) for dA, dB in dev]
        with g.Server(budget_bytes=16 << 30, max_batch=32) as srv:
            tickets = [srv.submit(dA, dB) for dA, dB in dev]
            served = []
            for t in tickets:
                dC, _ = srv.wait(t)
                served.append(dC.to_host())
                dC.free()
        for q, a, s in zip(reqs, alone, served):
            print(q["kind"], q["m"], q["dtype"], flush=True)
            assert np.array_equal(a.rpt, s.rpt) and np.array_equal(a.col, s.col)
            ref = oracle.spgemm(q["A"], q["B"])
            compare(q["A"], q["B"], a, ref)
            compare(q["A"], q["B"], s, ref)
    finally:
        for dA, dB in dev:
            dA.free()
            dB.free()
