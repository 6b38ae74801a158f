"""Synthetic workload generators: determinism and the stated properties."""
import numpy as np

from gpuspgemm_amd import generators as gen


def _rows_sorted_distinct(A):
    for i in range(0, A.m, max(1, A.m // 500)):
        c = A.col[A.rpt[i]:A.rpt[i + 1]]
        assert np.all(np.diff(c) > 0) and (c.size == 0 or (c[0] >= 0 and c[-1] < A.n))


def test_counter_rng_is_deterministic_and_uniform():
    a = gen.rand_unit(7, 3, np.arange(100_000))
    b = gen.rand_unit(7, 3, np.arange(100_000))
    np.testing.assert_array_equal(a, b)
    assert abs(a.mean() - 0.5) < 0.01 and a.min() >= 0 and a.max() < 1
    assert not np.array_equal(a, gen.rand_unit(8, 3, np.arange(100_000)))
    k = gen.rand_below(1, 1, np.arange(60_000), np.full(60_000, 6))
    assert set(np.unique(k)) == set(range(6))


def test_uniform8():
    A = gen.uniform8()
    assert (A.m, A.n, A.nnz) == (4096, 4096, 32768)
    assert np.all(np.diff(A.rpt) == 8)
    _rows_sorted_distinct(A)
    assert np.all(np.abs(A.val) <= 1)
    B = gen.uniform8()
    np.testing.assert_array_equal(A.col, B.col)


def test_powerlaw_properties():
    A = gen.powerlaw(m=200_000, seed=1)
    L = np.diff(A.rpt)
    assert L.min() >= 2 and L.max() <= 20000
    assert 2.7 < L.mean() < 3.2
    _rows_sorted_distinct(A)
    rows = np.repeat(np.arange(A.m), L)
    local = np.abs(A.col.astype(np.int64) - rows) <= 64
    assert 0.65 < local.mean() < 0.8


def test_stencil_structure():
    A = gen.stencil27(g=10)
    assert A.m == 1000 and A.nnz == (3 * 10 - 2) ** 3
    diag = A.col.astype(np.int64) == np.repeat(np.arange(A.m), np.diff(A.rpt))
    assert np.all(np.abs(A.val[diag] - 26) <= 1e-3) and np.all(np.abs(A.val[~diag] + 1) <= 1e-3)
    Z = gen.stencil27(g=10, noise=0.0)
    assert set(np.unique(Z.val)) == {-1.0, 26.0}
    _rows_sorted_distinct(A)


def test_batched_instance_properties():
    A, B = gen.batched_instance(5)
    assert A.m == B.m == A.n == B.n and 1000 <= A.m <= 200_000
    L = np.diff(A.rpt)
    assert L.min() >= 1 and 4.5 < L.mean() < 7.5
    _rows_sorted_distinct(A)
    _rows_sorted_distinct(B)
