"""Graph replay on the GPU (spg_set_graphs): the second call with a launch
signature captures each of the two phases into a HIP graph, later calls
replay them.  Every replay multiplies fresh device copies of the operands
with new values (new addresses and contents, the same launches), with eager
work and synchronisations between replays, and is checked on its own:
structure bit-exact against the oracle and values within the oracle
tolerance."""
import numpy as np
import pytest

import gpuspgemm_amd as g
from gpuspgemm_amd import generators as gen
from gpuspgemm_amd.csr import HostCSR

import oracle
from parity import compare

pytestmark = pytest.mark.gpu


@pytest.fixture
def graphs_on(gpu_lib):
    g.set_graphs(True)
    yield
    g.set_graphs(False)


def _scaled(A: HostCSR, s: float) -> HostCSR:
    return HostCSR(A.m, A.n, A.rpt.copy(), A.col.copy(), (A.val * s).astype(A.val.dtype))


def _multiply(A: HostCSR, B: HostCSR) -> HostCSR:
    dA, dB = A.to_device(), B.to_device()
    try:
        dC = g.proposal(dA, dB)
        C = dC.to_host()
        dC.free()
        return C
    finally:
        dA.free()
        dB.free()


def _check_scaled(A, B, C, ref, s):
    """C = (s A) B against the oracle product ref = A B."""
    assert C.nnz == ref.nnz
    np.testing.assert_array_equal(C.rpt.astype(np.int64), ref.rpt)
    np.testing.assert_array_equal(C.col, ref.col)
    diff = np.abs(C.val.astype(np.float64) - s * ref.val)
    scale = np.where(ref.absval > 0, s * ref.absval, 1.0)
    tol = 1e-12 if C.val.dtype == np.float64 else 1e-5
    assert float((diff / scale).max(initial=0.0)) <= tol


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_plain_product_replays(graphs_on, dtype):
    A = gen.powerlaw(m=60_000, seed=21).astype(dtype)
    ref = oracle.spgemm(A, A)
    before = g.graph_stats()
    for r in range(5):
        s = float(r + 1)
        C = _multiply(_scaled(A, s), A)
        _check_scaled(A, A, C, ref, s)
        g.load_library().spg_device_sync()
    after = g.graph_stats()
    # call 1 runs eagerly, call 2 captures both phases, calls 3-5 replay them
    assert after["captured"] - before["captured"] == 2
    assert after["replayed"] - before["replayed"] == 6


def test_interleaved_products_replay(graphs_on):
    """Two products of different shapes alternate; each keeps its own graphs,
    and eager calls (profiling on turns replay off) run between replays."""
    P = gen.powerlaw(m=30_000, seed=22)
    U = gen.uniform8(m=4096, seed=23)
    S = gen.stencil27(g=10, seed=24)
    refs = {id(M): oracle.spgemm(M, M) for M in (P, U)}
    for r in range(4):
        for M in (P, U):
            s = 0.5 + r
            _check_scaled(M, M, _multiply(_scaled(M, s), M), refs[id(M)], s)
        with g.profile():
            compare(S, S, _multiply(S, S))


def test_batched_replays(graphs_on):
    pairs = [gen.batched_instance(i, m=[1200, 300, 5000, 40, 2600, 900][i % 6]) for i in range(18)]
    refs = [oracle.spgemm(a, b) for a, b in pairs]
    before = g.graph_stats()
    for r in range(4):
        s = 1.0 + 0.25 * r
        dev = [(_scaled(a, s).to_device(), b.to_device()) for a, b in pairs]
        try:
            outs = g.proposal_batched([a for a, _ in dev], [b for _, b in dev])
            hosts = [c.to_host() for c in outs]
            for c in outs:
                c.free()
        finally:
            for a, b in dev:
                a.free()
                b.free()
        for (a, b), ref, h in zip(pairs, refs, hosts):
            _check_scaled(a, b, h, ref, s)
        # eager work between replays
        compare(pairs[0][0], pairs[0][1], _multiply(pairs[0][0], pairs[0][1]))
    after = g.graph_stats()
    assert after["replayed"] - before["replayed"] >= 4


def test_server_runs_eagerly_with_replay_on(graphs_on):
    """A server's lanes never capture (their worker threads run beside the
    caller's synchronous copies); with replay on, served results stay
    correct and no graph is captured or replayed for them."""
    A = gen.powerlaw(m=20_000, seed=25)
    ref = oracle.spgemm(A, A)
    dA = A.to_device()
    try:
        before = g.graph_stats()
        with g.Server(budget_bytes=1 << 30, max_batch=3) as srv:
            for _ in range(3):
                ts = [srv.submit(dA, dA) for _ in range(9)]
                for t in ts:
                    dC, _ = srv.wait(t)
                    h = dC.to_host()
                    dC.free()
                    compare(A, A, h, ref)
        assert g.graph_stats() == before
    finally:
        dA.free()
