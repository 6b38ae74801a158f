"""GPU parity: the HIP path (through the C ABI) against the oracle.

Structure (row_ptr, col_idx) must match bit-exactly; values within the
|A||B|-scaled tolerance of tests/parity.py (FP64 1e-12, FP32 1e-5).
"""
import os
import sys

import numpy as np
import pytest

import gpuspgemm_amd as g
from gpuspgemm_amd import generators as gen
from gpuspgemm_amd.csr import HostCSR
import oracle
from parity import compare, gpu_multiply

pytestmark = pytest.mark.gpu

INT_MAX = 2**31 - 1


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def csr_from_rows(rows, n, seed=0, dtype=np.float64):
    """CSR from a list of column lists (sorted, distinct) with U[-1,1] values."""
    rpt = np.zeros(len(rows) + 1, dtype=np.int64)
    rpt[1:] = np.cumsum([len(r) for r in rows])
    col = np.concatenate([np.asarray(sorted(set(r)), dtype=np.int32) for r in rows]) if rows else np.zeros(0, np.int32)
    val = gen.rand_unit(seed, 9, np.arange(col.size)) * 2 - 1
    return HostCSR(len(rows), n, rpt.astype(np.int32), col.astype(np.int32), val.astype(dtype))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_uniform8(gpu_lib, dtype):
    A = gen.uniform8().astype(dtype)
    _say("uniform8", np.dtype(dtype).name)
    st = compare(A, A, gpu_multiply(A, A))
    assert st["nnz"] > 0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_powerlaw_small(gpu_lib, dtype):
    A = gen.powerlaw(m=50_000, seed=11).astype(dtype)
    _say("powerlaw m=50000", np.dtype(dtype).name)
    compare(A, A, gpu_multiply(A, A))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_powerlaw_full(gpu_lib, dtype):
    """The full 10^6-row power-law product in both value types (FP32 operands
    are the narrowed FP64 ones; the oracle runs in FP64 on them)."""
    A = gen.powerlaw().astype(dtype)
    _say("powerlaw m=1e6", np.dtype(dtype).name)
    st = compare(A, A, gpu_multiply(A, A))
    _say(st)


def test_many_tiles_single_product(gpu_lib):
    """5 * 10^6 rows (2442 tiles of 2048 rows): the scan stage reads each
    tile's prefix off one scanned array of tile sums instead of summing the
    earlier tiles per workgroup; structure and values against the oracle."""
    A = gen.uniform8(m=5_000_000, per_row=2, seed=17, n=5_000_000)
    _say("uniform m=5e6, 2 per row")
    st = compare(A, A, gpu_multiply(A, A))
    assert st["nnz"] > 0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_stencil_small(gpu_lib, dtype):
    A = gen.stencil27(g=24).astype(dtype)
    _say("stencil g=24", np.dtype(dtype).name)
    compare(A, A, gpu_multiply(A, A))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_stencil_integer_known_answer(gpu_lib, dtype):
    """Noise-free stencil: C is integer-valued and exactly representable, so
    values are bit-exact whatever the summation order."""
    A = gen.stencil27(g=20, noise=0.0).astype(dtype)
    C = gpu_multiply(A, A)
    ref = oracle.spgemm(A, A)
    np.testing.assert_array_equal(C.col, ref.col)
    np.testing.assert_array_equal(C.val.astype(np.float64), ref.val)


def test_stencil_full_f64(gpu_lib):
    A = gen.stencil27()
    _say("stencil g=96 f64")
    C = gpu_multiply(A, A)
    assert C.nnz == 474**3
    st = compare(A, A, C)
    _say(st)


def test_rectangular_against_dense(gpu_lib):
    A = gen.random_csr(300, 500, 0.03, seed=5)
    B = gen.random_csr(500, 200, 0.05, seed=6)
    C = gpu_multiply(A, B)
    compare(A, B, C)
    dense, pattern, _ = oracle.dense_reference(A, B)
    rows = np.repeat(np.arange(C.m), np.diff(C.rpt))
    got = np.zeros((C.m, C.n), dtype=bool)
    got[rows, C.col] = True
    np.testing.assert_array_equal(got, pattern)
    np.testing.assert_allclose(C.val, dense[rows, C.col], rtol=0, atol=1e-12)


def test_exact_zero_sums_are_kept(gpu_lib):
    A = HostCSR(1, 2, np.array([0, 2], np.int32), np.array([0, 1], np.int32), np.array([1.0, 1.0]))
    B = HostCSR(2, 1, np.array([0, 1, 2], np.int32), np.array([0, 0], np.int32), np.array([1.0, -1.0]))
    C = gpu_multiply(A, B)
    assert C.nnz == 1 and C.col.tolist() == [0] and C.val.tolist() == [0.0]


def test_empty_rows_and_matrices(gpu_lib):
    A = gen.random_csr(400, 300, 0.02, seed=7, empty_rows=0.3)
    B = gen.random_csr(300, 250, 0.02, seed=8, empty_rows=0.3)
    compare(A, B, gpu_multiply(A, B))
    # no rows
    Z = HostCSR(0, 300, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    C = gpu_multiply(Z, B)
    assert (C.m, C.n, C.nnz) == (0, 250, 0) and C.rpt.tolist() == [0]
    # no entries
    E = HostCSR(50, 300, np.zeros(51, np.int32), np.zeros(0, np.int32), np.zeros(0))
    C = gpu_multiply(E, B)
    assert C.nnz == 0 and not C.rpt.any()
    EB = HostCSR(300, 250, np.zeros(301, np.int32), np.zeros(0, np.int32), np.zeros(0))
    C = gpu_multiply(A, EB)
    assert C.nnz == 0 and not C.rpt.any()


def test_dimension_mismatch_is_an_error(gpu_lib):
    A = gen.random_csr(10, 12, 0.2, seed=1)
    B = gen.random_csr(11, 10, 0.2, seed=2)
    dA, dB = A.to_device(), B.to_device()
    with pytest.raises(g.SpgError) as exc:
        g.proposal(dA, dB)
    assert exc.value.status_name == "SPG_ERR_INVALID_ARGUMENT"
    assert "must match" in str(exc.value)


def _heavy_case(seed=3):
    """Rows that take every fallback: one row with ~60000 distinct columns
    (overflows the largest LDS table in the count pass and every LDS
    capacity in the fused pass), one with ~6000 (largest LDS bins), and
    ordinary rows."""
    nb = 200_000
    r0 = np.unique(gen.rand_below(seed, 1, np.arange(20000), np.full(20000, nb)))
    r1 = np.unique(gen.rand_below(seed, 2, np.arange(2000), np.full(2000, nb)))
    rng_rows = [r0.tolist(), r1.tolist()] + [
        np.unique(gen.rand_below(seed, 3 + i, np.arange(4), np.full(4, nb))).tolist() for i in range(500)]
    A = csr_from_rows(rng_rows, nb, seed=seed)
    keys = np.unique(np.arange(nb, dtype=np.int64)[:, None] * 1_000_000
                     + gen.rand_below(seed, 99, np.arange(nb * 3), np.full(nb * 3, 1_000_000)).reshape(nb, 3))
    rows = keys // 1_000_000
    col = (keys % 1_000_000).astype(np.int32)
    rpt = np.zeros(nb + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=nb), out=rpt[1:])
    B = HostCSR(nb, 1_000_000, rpt.astype(np.int32), col,
                gen.rand_unit(seed, 98, np.arange(col.size)) * 2 - 1)
    return A, B


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_heavy_rows_take_global_paths(gpu_lib, dtype):
    A, B = _heavy_case()
    A, B = A.astype(dtype), B.astype(dtype)
    _say("heavy rows", np.dtype(dtype).name)
    st = compare(A, B, gpu_multiply(A, B))
    assert st["nnz"] > 60_000


def _mixed_case():
    A = gen.powerlaw(m=30_000, seed=21)
    return A, A


def _variants(phase):
    return g.variant_list(phase)


SYM_MULTI = [v["id"] for v in _variants(0) if not v["max"]] if g._lib.available() else list(range(13))
NUM_LDS = [v["id"] for v in _variants(1) if not v["global"]] if g._lib.available() else list(range(14))


def _sym_max():
    return next(v["id"] for v in _variants(0) if v["max"])


def _num_global():
    return next(v["id"] for v in _variants(1) if v["global"])


@pytest.mark.parametrize("variant", SYM_MULTI)
def test_every_symbolic_variant(gpu_lib, variant):
    A, B = _mixed_case()
    t = next(v["table"] for v in _variants(0) if v["id"] == variant)
    g.set_tuning(0, f"{variant} {t // 2} {INT_MAX}\n{_sym_max()} {INT_MAX} {INT_MAX}\n")
    try:
        compare(A, B, gpu_multiply(A, B))
    finally:
        g.set_tuning(0, None)


@pytest.mark.parametrize("variant", NUM_LDS)
def test_every_fused_variant(gpu_lib, variant):
    A, B = _mixed_case()
    p = next(v["capacity"] for v in _variants(1) if v["id"] == variant)
    g.set_tuning(1, f"{variant} {INT_MAX} {p}\n{_num_global()} {INT_MAX} {INT_MAX}\n")
    try:
        compare(A, B, gpu_multiply(A, B))
    finally:
        g.set_tuning(1, None)


def test_global_paths_forced(gpu_lib):
    """Every row through the largest-table and global kernels."""
    A = gen.random_csr(600, 800, 0.02, seed=31)
    B = gen.random_csr(800, 700, 0.02, seed=32)
    g.set_tuning(0, f"{_sym_max()} {INT_MAX} {INT_MAX}\n")
    g.set_tuning(1, f"{_num_global()} {INT_MAX} {INT_MAX}\n")
    try:
        compare(A, B, gpu_multiply(A, B))
    finally:
        g.set_tuning(0, None)
        g.set_tuning(1, None)


def test_batched_matches_single(gpu_lib):
    pairs = []
    for i in range(12):
        m = [1500, 40, 3000, 1, 700, 2048, 1024, 1025, 5000, 300, 64, 2500][i]
        pairs.append((gen.random_csr(m, 600, 0.01, seed=40 + i), gen.random_csr(600, 500, 0.01, seed=80 + i)))
    pairs.append((HostCSR(0, 600, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0)),
                  gen.random_csr(600, 500, 0.01, seed=99)))
    dA = [a.to_device() for a, _ in pairs]
    dB = [b.to_device() for _, b in pairs]
    Cs = g.proposal_batched(dA, dB)
    for (A, B), dC in zip(pairs, Cs):
        C = dC.to_host()
        compare(A, B, C)
        dC.free()


def test_repeated_calls_reuse_state(gpu_lib):
    small = gen.random_csr(100, 100, 0.05, seed=1)
    big = gen.powerlaw(m=100_000, seed=2)
    for M in (small, big, small, big):
        compare(M, M, gpu_multiply(M, M))


def test_get_nip(gpu_lib):
    A = gen.uniform8()
    dA = A.to_device()
    assert g.get_nip(dA, dA) == 64 * 4096
    S = gen.stencil27(g=12)
    dS = S.to_device()
    assert g.get_nip(dS, dS) == int(oracle.nip_rows(S, S).sum())
    # the bench's FLOP count: the full power-law product (489 tiles, the
    # last one partial), and a rectangular product with empty rows
    P = gen.powerlaw()
    dP = P.to_device()
    assert g.get_nip(dP, dP) == int(oracle.nip_rows(P, P).sum())
    R = gen.random_csr(3000, 900, 0.01, seed=5, empty_rows=0.3)
    Q = gen.random_csr(900, 700, 0.02, seed=6, empty_rows=0.2)
    dR, dQ = R.to_device(), Q.to_device()
    assert g.get_nip(dR, dQ) == int(oracle.nip_rows(R, Q).sum())


def _batched_pair(i):
    return gen.batched_instance(i, seed=3)


def test_batched_full_size(gpu_lib):
    """The batched workload at full size: 512 instances, rows log-uniform in
    [10^3, 2*10^5], geometric row lengths, a third of rows R-MAT-skewed.
    Every instance's result from the one batched call has the structure of
    its own plain multiply bit for bit and values within the FP64 bound
    (relative to the instance's largest |value|, both runs are the same
    algorithm); eight instances are also checked against the oracle."""
    import multiprocessing as mp
    with mp.get_context("fork").Pool(min(16, os.cpu_count() or 1)) as pool:
        pairs = pool.map(_batched_pair, range(512), chunksize=8)
    dev = [(a.to_device(), b.to_device()) for a, b in pairs]
    try:
        outs = g.proposal_batched([a for a, _ in dev], [b for _, b in dev])
        hosts = [c.to_host() for c in outs]
        for c in outs:
            c.free()
        for i, (da, db) in enumerate(dev):
            dc = g.proposal(da, db)
            single = dc.to_host()
            dc.free()
            h = hosts[i]
            assert h.nnz == single.nnz and np.array_equal(h.rpt, single.rpt) and np.array_equal(h.col, single.col)
            scale = max(1.0, float(np.abs(single.val).max())) if single.nnz else 1.0
            assert float(np.abs(h.val - single.val).max(initial=0.0)) <= 1e-12 * scale
        for i in range(0, 512, 64):
            compare(pairs[i][0], pairs[i][1], hosts[i])
    finally:
        for a, b in dev:
            a.free()
            b.free()
