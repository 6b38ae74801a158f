"""Seeded synthetic CSR generators for the benchmark and test workloads.

Every random draw is a pure function of (seed, stream, index) through a
counter-based splitmix64 mix, so a matrix is reproducible from its seed on
any host and any numpy version (no numpy RNG stream is involved).

Workloads (all square, 0-based, int32 indices, columns sorted and distinct
per row, float64 values; float32 operands are these values narrowed, as the
reference's `.csr` loader narrows them, src/libutils/include/utils/csr.cuh:146-154):

* ``uniform8``   4096 x 4096, exactly 8 distinct uniform columns per row,
                 values U[-1, 1]                                    (seed 0)
* ``powerlaw``   1,000,000 rows; row length min(20000, ceil(1.2 U^-1/2))
                 (Pareto alpha = 2); each column is with p = 0.7 uniform in
                 [i-64, i+64] clipped to the matrix, else uniform over all
                 columns; duplicates are redrawn; values U[-1, 1]   (seed 1)
* ``stencil``    27-point 3D stencil on a g^3 grid (g = 96 by default),
                 26 on the diagonal and -1 off it, plus U[-1e-3, 1e-3] noise
                 (noise = 0 gives the integer-valued known-answer case)
                                                                    (seed 2)
* ``batched``    independent products with rows log-uniform in
                 [1e3, 2e5], geometric(p = 1/6) row lengths, uniform columns,
                 one third of the rows R-MAT-skewed (a=.57, b=c=.19)  (seed 3)
"""
from __future__ import annotations

import math

import numpy as np

from .csr import HostCSR

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_S1 = np.uint64(30)
_S2 = np.uint64(27)
_S3 = np.uint64(31)


def _mix64(z: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser on a uint64 array (wrapping arithmetic)."""
    z = (z ^ (z >> _S1)) * _M1
    z = (z ^ (z >> _S2)) * _M2
    return z ^ (z >> _S3)


def rand_u64(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    """Counter-based 64-bit draws: one independent value per index."""
    with np.errstate(over="ignore"):
        key = _mix64(np.array([(seed * 0x632BE59BD9B4E019 + stream * 0x2545F4914F6CDD1D)
                               & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64))[0]
        z = key + (np.asarray(idx, dtype=np.uint64) + np.uint64(1)) * _GOLDEN
        return _mix64(z)


def rand_unit(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    """Uniform doubles in [0, 1) with 53 random bits."""
    return (rand_u64(seed, stream, idx) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def rand_below(seed: int, stream: int, idx: np.ndarray, bound: np.ndarray) -> np.ndarray:
    """Uniform integers in [0, bound) (bound < 2^32; multiply-shift, bias < 2^-32)."""
    hi = rand_u64(seed, stream, idx) >> np.uint64(32)
    return ((hi * np.asarray(bound, dtype=np.uint64)) >> np.uint64(32)).astype(np.int64)


def _csr_from_sorted_keys(keys: np.ndarray, m: int, n: int) -> tuple[np.ndarray, np.ndarray]:
    rows = (keys // n).astype(np.int64)
    col = (keys % n).astype(np.int32)
    counts = np.bincount(rows, minlength=m)
    rpt = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(counts, out=rpt[1:])
    if rpt[-1] >= 2 ** 31:
        raise OverflowError("nnz does not fit int32")
    return rpt.astype(np.int32), col


def _distinct_rows(seed: int, m: int, n: int, lengths: np.ndarray, draw) -> tuple[np.ndarray, np.ndarray]:
    """Rows of exactly lengths[i] distinct columns; `draw(round, rows, j)` gives candidates.

    Round r draws lengths[i] - have[i] fresh candidates for every deficient row;
    duplicates are dropped and redrawn in the next round.
    """
    lengths = np.asarray(lengths, dtype=np.int64)
    if np.any(lengths > n):
        raise ValueError("row length exceeds column count")
    keys = np.empty(0, dtype=np.int64)
    have = np.zeros(m, dtype=np.int64)
    rnd = 0
    while True:
        need = lengths - have
        rows = np.nonzero(need > 0)[0]
        if rows.size == 0:
            break
        cnt = need[rows]
        row_of = np.repeat(rows, cnt)
        start = np.cumsum(cnt) - cnt
        j = np.arange(row_of.size, dtype=np.int64) - np.repeat(start, cnt) + np.repeat(have[rows], cnt)
        cols = draw(rnd, row_of, j)
        keys = np.unique(np.concatenate([keys, row_of * n + cols]))
        have = np.bincount(keys // n, minlength=m).astype(np.int64)
        rnd += 1
        if rnd > 200:
            raise RuntimeError("distinct-column generation did not converge")
    return _csr_from_sorted_keys(keys, m, n)


def _counter(rnd: int, rows: np.ndarray, j: np.ndarray, m: int) -> np.ndarray:
    # unique draw index per (round, row, position); positions < 2^21
    return ((np.int64(rnd) * m + rows) << 21) + j


def uniform8(m: int = 4096, per_row: int = 8, seed: int = 0, n: int | None = None) -> HostCSR:
    """m x n (n = m by default), exactly `per_row` distinct uniform columns
    per row, values U[-1,1]."""
    n = m if n is None else n
    def draw(rnd, rows, j):
        return rand_below(seed, 1, _counter(rnd, rows, j, m), np.full(rows.shape, n))
    rpt, col = _distinct_rows(seed, m, n, np.full(m, per_row), draw)
    val = rand_unit(seed, 2, np.arange(col.size)) * 2.0 - 1.0
    return HostCSR(m, n, rpt, col, val)


def powerlaw(m: int = 1_000_000, seed: int = 1, max_len: int = 20000,
             x_min: float = 1.2, local: int = 64, p_local: float = 0.7) -> HostCSR:
    """Pareto(alpha=2) row lengths, 70 % near-diagonal / 30 % uniform columns."""
    n = m
    u = 1.0 - rand_unit(seed, 1, np.arange(m))           # (0, 1]
    lengths = np.minimum(np.minimum(max_len, n), np.ceil(x_min / np.sqrt(u))).astype(np.int64)

    def draw(rnd, rows, j):
        idx = _counter(rnd, rows, j, m)
        pick_local = rand_unit(seed, 2, idx) < p_local
        lo = np.maximum(rows - local, 0)
        hi = np.minimum(rows + local, n - 1)
        c_local = lo + rand_below(seed, 3, idx, hi - lo + 1)
        c_far = rand_below(seed, 4, idx, np.full(rows.shape, n))
        return np.where(pick_local, c_local, c_far)

    rpt, col = _distinct_rows(seed, m, n, lengths, draw)
    val = rand_unit(seed, 5, np.arange(col.size)) * 2.0 - 1.0
    return HostCSR(m, n, rpt, col, val)


def stencil27(g: int = 96, noise: float = 1e-3, seed: int = 2) -> HostCSR:
    """27-point stencil on a g^3 grid; row r = (z*g + y)*g + x."""
    m = g ** 3
    r = np.arange(m, dtype=np.int64)
    x, y, z = r % g, (r // g) % g, r // (g * g)
    cols = []
    valid = []
    for dz in (-1, 0, 1):          # ascending column order: dz, dy, dx
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                ok = ((x + dx >= 0) & (x + dx < g) & (y + dy >= 0) & (y + dy < g)
                      & (z + dz >= 0) & (z + dz < g))
                cols.append(r + (dz * g + dy) * g + dx)
                valid.append(ok)
    cols = np.stack(cols, axis=1)
    valid = np.stack(valid, axis=1)
    col = cols[valid].astype(np.int32)                  # row-major: rows in order, cols ascending
    counts = valid.sum(axis=1)
    rpt = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(counts, out=rpt[1:])
    row_of = np.repeat(r, counts)
    val = np.where(col.astype(np.int64) == row_of, 26.0, -1.0)
    if noise:
        val = val + (rand_unit(seed, 1, np.arange(col.size)) * 2.0 - 1.0) * noise
    return HostCSR(m, m, rpt.astype(np.int32), col, val)


def _rmat_cols(seed: int, stream: int, idx: np.ndarray, n: int, b_plus_d: float = 0.24) -> np.ndarray:
    """Column marginal of R-MAT(a=.57,b=.19,c=.19,d=.05): each bit is 1 w.p. b+d."""
    levels = max(1, math.ceil(math.log2(max(n, 2))))
    out = np.zeros(idx.shape, dtype=np.int64)
    pending = np.ones(idx.shape, dtype=bool)
    attempt = 0
    while pending.any():
        cand = np.zeros(int(pending.sum()), dtype=np.int64)
        sub = idx[pending]
        for lvl in range(levels):
            bit = rand_unit(seed, stream + 1 + attempt * 64 + lvl, sub) < b_plus_d
            cand = (cand << 1) | bit.astype(np.int64)
        ok = cand < n
        tgt = np.nonzero(pending)[0]
        out[tgt[ok]] = cand[ok]
        pending[tgt[ok]] = False
        attempt += 1
    return out


def batched_instance(i: int, seed: int = 3, m: int | None = None) -> tuple[HostCSR, HostCSR]:
    """Instance i of the batched workload: independent A_i, B_i (m_i x m_i)."""
    if m is None:
        u = rand_unit(seed, 1, np.array([i]))[0]
        m = int(round(math.exp(math.log(1e3) + u * (math.log(2e5) - math.log(1e3)))))
    mats = []
    for which in range(2):
        s = seed * 1_000_003 + i * 2 + which + 17
        u = 1.0 - rand_unit(s, 2, np.arange(m))
        # geometric on {1,2,...} with p = 1/6 by inversion
        lengths = np.minimum(m, np.maximum(1, np.ceil(np.log(u) / math.log(5.0 / 6.0)))).astype(np.int64)
        rmat_row = rand_unit(s, 3, np.arange(m)) < (1.0 / 3.0)

        def draw(rnd, rows, j, s=s, rmat_row=rmat_row):
            idx = _counter(rnd, rows, j, m)
            c_uni = rand_below(s, 4, idx, np.full(rows.shape, m))
            c_rmat = _rmat_cols(s, 5 + rnd * 4096, idx, m)
            return np.where(rmat_row[rows], c_rmat, c_uni)

        rpt, col = _distinct_rows(s, m, m, lengths, draw)
        val = rand_unit(s, 6, np.arange(col.size)) * 2.0 - 1.0
        mats.append(HostCSR(m, m, rpt, col, val))
    return mats[0], mats[1]


def random_csr(m: int, n: int, density: float, seed: int, dtype=np.float64,
               empty_rows: float = 0.0) -> HostCSR:
    """Small random CSR for parity tests (Bernoulli pattern, U[-1,1] values)."""
    idx = np.arange(m * n, dtype=np.int64)
    keep = rand_unit(seed, 1, idx) < density
    if empty_rows:
        dead = rand_unit(seed, 2, np.arange(m)) < empty_rows
        keep &= ~np.repeat(dead, n)
    keys = idx[keep]
    rpt, col = _csr_from_sorted_keys(keys, m, n)
    val = (rand_unit(seed, 3, np.arange(col.size)) * 2.0 - 1.0).astype(dtype)
    return HostCSR(m, n, rpt, col, val)


def served_request(k: int, seed: int = 4) -> dict:
    """Request k of the served workload: value type 70 % FP64 / 30 % FP32,
    generator drawn uniformly from uniform8 / power-law / stencil / batched
    instance, rows log-uniform in [2^10, 2^20]."""
    r = rand_unit(seed, 1, np.arange(4 * k, 4 * k + 4))
    dtype = np.float64 if r[0] < 0.7 else np.float32
    kind = ["uniform8", "powerlaw", "stencil", "batched"][min(3, int(r[1] * 4))]
    m = int(round(2.0 ** (10 + 10 * r[2])))
    if kind == "uniform8":
        A = uniform8(m=m, seed=100 + k)
        B = A
    elif kind == "powerlaw":
        A = powerlaw(m=m, seed=100 + k)
        B = A
    elif kind == "stencil":
        A = stencil27(g=max(2, int(round(m ** (1 / 3)))), seed=100 + k)
        B = A
    else:
        A, B = batched_instance(k, seed=seed, m=m)
    return {"A": A.astype(dtype), "B": B.astype(dtype), "kind": kind, "m": A.m, "dtype": np.dtype(dtype).name}


WORKLOADS = {
    "uniform8": lambda: uniform8(),
    "powerlaw": lambda: powerlaw(),
    "stencil": lambda: stencil27(),
}
