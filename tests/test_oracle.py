"""Pins the oracle (oracle/spgemm_oracle.c) before it is trusted as the
checker.  The reference ships no golden vectors or fixtures for this path,
so every function is pinned by what the mathematics fixes:

* a dense brute-force product (values, pattern kept by counting) on random
  square and rectangular inputs, where a transposed operand or a swapped
  dimension fails the shape or value check;
* closed forms of the 27-point stencil: NIP and nnz per row (products of
  per-dimension counts), totals (9g-10)^3 and (5g-6)^3, and the integer
  values of the noise-free product;
* the uniform workload's NIP of exactly 64 per row;
* exact-zero sums kept as entries, empty rows and empty matrices;
* planted mistakes (a dropped product, a wrong column, a perturbed value)
  that tests/parity.py's comparison must reject.
"""
import numpy as np
import pytest

import oracle
from gpuspgemm_amd import generators as gen
from gpuspgemm_amd.csr import HostCSR
import parity


def _check_dense(A, B):
    r = oracle.spgemm(A, B)
    dense, pattern, absval = oracle.dense_reference(A, B)
    rows = np.repeat(np.arange(A.m), np.diff(r.rpt))
    got = np.zeros((A.m, B.n), dtype=bool)
    got[rows, r.col] = True
    np.testing.assert_array_equal(got, pattern)
    # columns ascending within each row
    for i in range(A.m):
        c = r.col[r.rpt[i]:r.rpt[i + 1]]
        assert np.all(np.diff(c) > 0)
    np.testing.assert_allclose(r.val, dense[rows, r.col], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(r.absval, absval[rows, r.col], rtol=1e-13, atol=1e-13)
    return r


@pytest.mark.parametrize("seed", range(4))
def test_random_square_against_dense(seed):
    A = gen.random_csr(60, 60, 0.08, seed=seed)
    _check_dense(A, A)


def test_rectangular_against_dense():
    A = gen.random_csr(37, 53, 0.1, seed=11)
    B = gen.random_csr(53, 29, 0.12, seed=12)
    r = _check_dense(A, B)
    assert r.rpt.shape == (38,)
    # B.A has a different shape: a transposed operand cannot pass
    with pytest.raises(ValueError):
        oracle.spgemm(B, B)


def test_non_commuting_product():
    A = gen.random_csr(40, 40, 0.1, seed=21)
    B = gen.random_csr(40, 40, 0.1, seed=22)
    ab = oracle.spgemm(A, B)
    ba = oracle.spgemm(B, A)
    d_ab = oracle.dense_reference(A, B)[0]
    assert not np.allclose(d_ab, oracle.dense_reference(B, A)[0])
    rows = np.repeat(np.arange(40), np.diff(ab.rpt))
    np.testing.assert_allclose(ab.val, d_ab[rows, ab.col], atol=1e-13)
    assert ab.nnz != ba.nnz or not np.array_equal(ab.col, ba.col) or not np.allclose(ab.val, ba.val)


def _per_dim_counts(g, radius):
    x = np.arange(g)
    lo = np.maximum(x - radius, 0)
    hi = np.minimum(x + radius, g - 1)
    return hi - lo + 1


def test_stencil_closed_forms():
    g = 14
    A = gen.stencil27(g=g)
    r = oracle.spgemm(A, A)
    c1 = _per_dim_counts(g, 1)                       # stencil neighbours per dimension
    nip_dim = np.array([c1[max(0, x - 1):min(g, x + 2)].sum() for x in range(g)])
    nnz_dim = _per_dim_counts(g, 2)
    rr = np.arange(g ** 3)
    x, y, z = rr % g, (rr // g) % g, rr // (g * g)
    np.testing.assert_array_equal(r.nip, nip_dim[x] * nip_dim[y] * nip_dim[z])
    np.testing.assert_array_equal(np.diff(r.rpt), nnz_dim[x] * nnz_dim[y] * nnz_dim[z])
    assert r.nip.sum() == (9 * g - 10) ** 3
    assert r.nnz == (5 * g - 6) ** 3


def test_stencil_full_size_totals():
    """96^3 grid: NIP 854^3 and nnz(C) 474^3."""
    A = gen.stencil27()
    nip = oracle.nip_rows(A, A)
    assert nip.sum() == 854 ** 3
    r = oracle.spgemm(A, A)
    assert r.nnz == 474 ** 3


def test_noise_free_stencil_values():
    """C_ij = sum_k A_ik A_kj with 26 on the diagonal and -1 off it: for
    i == j, 676 + (|N(i)| - 1); for j a neighbour of i, |K| - 2 - 52; else |K|,
    with K the common in-grid neighbourhood of i and j."""
    g = 7
    A = gen.stencil27(g=g, noise=0.0)
    r = oracle.spgemm(A, A)
    rows = np.repeat(np.arange(g ** 3), np.diff(r.rpt))
    cols = r.col.astype(np.int64)

    def coords(v):
        return v % g, (v // g) % g, v // (g * g)

    (ix, iy, iz), (jx, jy, jz) = coords(rows), coords(cols)

    def common(a, b):  # in-grid t with |t-a| <= 1 and |t-b| <= 1
        lo = np.maximum(np.maximum(a, b) - 1, 0)
        hi = np.minimum(np.minimum(a, b) + 1, g - 1)
        return np.maximum(hi - lo + 1, 0)

    K = common(ix, jx) * common(iy, jy) * common(iz, jz)
    neighbour = (np.abs(ix - jx) <= 1) & (np.abs(iy - jy) <= 1) & (np.abs(iz - jz) <= 1)
    expect = np.where(rows == cols, 676 + K - 1, np.where(neighbour, K - 2 - 52, K)).astype(np.float64)
    np.testing.assert_array_equal(r.val, expect)


def test_uniform_nip_is_64_per_row():
    A = gen.uniform8()
    assert np.all(np.diff(A.rpt) == 8)
    nip = oracle.nip_rows(A, A)
    assert np.all(nip == 64) and nip.sum() == 262_144


def test_exact_zero_sums_are_entries():
    A = HostCSR(1, 2, np.array([0, 2], np.int32), np.array([0, 1], np.int32), np.array([1.0, 1.0]))
    B = HostCSR(2, 1, np.array([0, 1, 2], np.int32), np.array([0, 0], np.int32), np.array([1.0, -1.0]))
    r = oracle.spgemm(A, B)
    assert r.nnz == 1 and r.col.tolist() == [0] and r.val.tolist() == [0.0] and r.absval.tolist() == [2.0]


def test_empty_rows_and_matrices():
    A = gen.random_csr(50, 40, 0.05, seed=3, empty_rows=0.5)
    _check_dense(A, gen.random_csr(40, 30, 0.05, seed=4, empty_rows=0.5))
    Z = HostCSR(0, 5, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    r = oracle.spgemm(Z, gen.random_csr(5, 5, 0.5, seed=1))
    assert r.nnz == 0 and r.rpt.tolist() == [0]
    E = HostCSR(4, 5, np.zeros(5, np.int32), np.zeros(0, np.int32), np.zeros(0))
    r = oracle.spgemm(E, gen.random_csr(5, 5, 0.5, seed=1))
    assert r.nnz == 0 and not r.rpt.any()


def test_threads_do_not_change_the_result():
    A = gen.powerlaw(m=20_000, seed=3)
    r1 = oracle.spgemm(A, A, threads=1)
    r8 = oracle.spgemm(A, A, threads=8)
    np.testing.assert_array_equal(r1.rpt, r8.rpt)
    np.testing.assert_array_equal(r1.col, r8.col)
    np.testing.assert_array_equal(r1.val, r8.val)


# ---- the comparison used by the GPU parity tests must reject planted mistakes

def _as_host(r, A, B):
    return HostCSR(A.m, B.n, r.rpt.astype(np.int32), r.col.copy(), r.val.copy())


def test_comparison_accepts_the_oracle_itself():
    A = gen.random_csr(80, 80, 0.06, seed=5)
    r = oracle.spgemm(A, A)
    st = parity.compare(A, A, _as_host(r, A, A), ref=r)
    assert st["max_scaled_err"] == 0.0 and st["reference_rule_ok"]


def test_comparison_rejects_planted_mistakes():
    A = gen.random_csr(80, 80, 0.06, seed=5)
    r = oracle.spgemm(A, A)
    # a dropped term in one entry
    C = _as_host(r, A, A)
    k = int(np.argmax(r.absval - np.abs(r.val)))
    C.val[k] = r.val[k] + r.absval[k] * 1e-3
    with pytest.raises(AssertionError):
        parity.compare(A, A, C, ref=r)
    # a wrong column index
    C = _as_host(r, A, A)
    C.col[3] += 1
    with pytest.raises(AssertionError):
        parity.compare(A, A, C, ref=r)
    # a tiny FP64 perturbation beyond 1e-12 of |A||B|
    C = _as_host(r, A, A)
    C.val[7] += r.absval[7] * 1e-10
    with pytest.raises(AssertionError):
        parity.compare(A, A, C, ref=r)
    # the transposed product
    B = gen.random_csr(80, 80, 0.06, seed=6)
    rab = oracle.spgemm(A, B)
    rba = oracle.spgemm(B, A)
    with pytest.raises(AssertionError):
        parity.compare(A, B, _as_host(rba, B, A), ref=rab)
