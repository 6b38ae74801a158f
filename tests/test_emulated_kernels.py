"""The library's kernels and host driver, compiled as host C++ against the
CPU emulation of the HIP API in tests/emu (one thread per GPU lane) and run
under AddressSanitizer: every bin variant with rows that fit it in one
chunk and rows spread over several chunks, the global fallbacks, and
out-of-bounds accesses caught on the CPU.  Results are checked against the
host Gustavson product by tests/emu/emu_main.cpp.  Inputs are small:
emulating a 1024-lane workgroup costs a thread per lane."""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from gpuspgemm_amd import generators as gen
from gpuspgemm_amd.csr import HostCSR

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")
BUILD = os.path.join(EMU, "_build")
EXE = os.path.join(BUILD, "emu_spgemm")
CSRC = os.path.join(ROOT, "gpuspgemm_amd", "csrc")
INT_MAX = 2**31 - 1

import gpuspgemm_amd as g
import oracle

# Compiled variants, read from the library (no GPU needed): symbolic
# (lanes per row, A entries per chunk, table slots) and fused (lanes, chunk,
# capacity).
_SV = g.variant_list(0)
_NV = g.variant_list(1)
SYM = {v["id"]: (v["lanes"], v["chunk"], v["table"]) for v in _SV if not v["max"]}
SYM_MAX = next(v["id"] for v in _SV if v["max"])
NUM = {v["id"]: (v["lanes"], v["chunk"], v["capacity"]) for v in _NV if not v["global"]}
NUM_GLOBAL = next(v["id"] for v in _NV if v["global"])


def build_emu(csrc: str, exe: str) -> None:
    """The emulation build of the kernel and host sources in `csrc`: each
    translation unit compiled by its own g++ process, side by side."""
    srcs = [os.path.join(csrc, f) for f in ("k_rows.hip", "k_sym.hip", "k_symnum.hip", "api.hip")]
    srcs += [os.path.join(EMU, "emu_runtime.cpp"), os.path.join(EMU, "emu_main.cpp")]
    flags = ["-std=c++17", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-pthread",
             "-Wno-unused-result"]
    inc = ["-I", os.path.join(EMU, "include"), "-I", os.path.join(ROOT, "include"), "-I", csrc,
           "-I", os.path.join(ROOT, "gpuspgemm_amd", "apps")]
    objs = [f"{exe}.{i}.o" for i in range(len(srcs))]
    procs = [subprocess.Popen(["g++", *flags, *inc, "-x", "c++", "-c", s, "-o", o]) for s, o in zip(srcs, objs)]
    codes = [p.wait() for p in procs]
    try:
        assert codes == [0] * len(srcs), f"emulation build failed: {codes}"
        subprocess.run(["g++", *flags, *objs, "-o", exe], check=True)
    finally:
        for o in objs:
            if os.path.exists(o):
                os.remove(o)


@pytest.fixture(scope="module")
def emu():
    os.makedirs(BUILD, exist_ok=True)
    tmp = f".{os.getpid()}"
    build_emu(CSRC, EXE + tmp)
    os.replace(EXE + tmp, EXE)  # parallel test workers may be running the previous build
    return EXE


def run_emu(exe, tmp_path, A, B, *extra, tables=None, env=None):
    fa, fb = tmp_path / "a.csr", tmp_path / "b.csr"
    A.save_to_filename(str(fa))
    B.save_to_filename(str(fb))
    args = [exe, str(fa), str(fb), *extra]
    for phase, text in (tables or {}).items():
        f = tmp_path / f"table{phase}.txt"
        f.write_text(text)
        args += ["--sym-table" if phase == 0 else "--num-table", str(f)]
    res = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600,
                         env=None if env is None else {**os.environ, **env})
    assert res.returncode == 0, res.stdout[-4000:]
    assert "structure identical" in res.stdout
    return res.stdout


def call(A, B, *extra, tables=None, env=None):
    """One emulator run for run_emu_all."""
    return A, B, extra, tables, env


def run_emu_all(exe, tmp_path, *calls):
    """Independent emulator runs side by side (each emulated workgroup keeps
    only a few host cores busy), each in a directory of its own; their
    outputs in order."""
    def one(i):
        A, B, extra, tables, env = calls[i]
        d = tmp_path / f"run{i}"
        d.mkdir()
        return run_emu(exe, d, A, B, *extra, tables=tables, env=env)
    with ThreadPoolExecutor(len(calls)) as ex:
        return list(ex.map(one, range(len(calls))))


def squares(m, k, seed):
    """k entries per row of A and of B: NIP = k^2, nnz <= k^2."""
    n = max(4 * k, 64)
    return gen.uniform8(m=m, per_row=k, seed=seed, n=n), gen.uniform8(m=n, per_row=k, seed=seed + 1, n=40 * k)


def chunked(m, k, seed):
    """k entries per row of A, one per row of B: NIP = nnz <= k, A rows span k/EC chunks."""
    n = max(4 * k, 64)
    return gen.uniform8(m=m, per_row=k, seed=seed, n=n), gen.uniform8(m=n, per_row=1, seed=seed + 1, n=n)


def rows_for(lanes):
    return 96 if lanes <= 16 else (12 if lanes <= 256 else 3)


def test_uniform_stencil_powerlaw(emu, tmp_path):
    A = gen.uniform8(m=512)
    S = gen.stencil27(g=5).astype(np.float32)
    P = gen.powerlaw(m=1000, seed=5)
    run_emu_all(emu, tmp_path, call(A, A), call(S, S, "--f32"), call(P, P, "--repeat", "2"))


def test_graph_replay(emu, tmp_path):
    """Graph replay on: each repeat multiplies a fresh copy of A with new
    values (new addresses, same launches).  The first call runs eagerly, the
    second captures both phases, the rest replay them; every result must
    equal the host product."""
    A = gen.powerlaw(m=1000, seed=5)
    R = gen.random_csr(200, 150, 0.04, seed=6, empty_rows=0.3)
    Rb = gen.random_csr(150, 120, 0.05, seed=7)
    P = gen.powerlaw(m=600, seed=9)
    out = run_emu_all(emu, tmp_path, call(A, A, "--repeat", "4", "--graphs"),
                      call(R, Rb, "--repeat", "3", "--graphs"),
                      # the server's lanes run eagerly with replay on
                      call(P, P, "--serve", "6", str(1 << 30), "3", "--graphs"))
    assert out[0].count("structure identical") == 4
    assert "graphs captured 2 replayed 4" in out[0]
    assert out[1].count("structure identical") == 3
    assert "graphs captured 2 replayed 2" in out[1]
    assert out[2].count("structure identical") == 6


def test_merged_launches(emu, tmp_path):
    """A fused table with a bin for every variant of the merged small-row
    launch and four bins of the merged longer-row launch: the rows of each
    set's bins share one launch, a workgroup per unit; every result must
    equal the host product."""
    assert sorted(v["id"] for v in _NV if v["merged"] == 0) == [10, 11, 12, 13, 14, 16]
    assert sorted(v["id"] for v in _NV if v["merged"] == 1) == [4, 5, 6, 7, 8]
    table = ("10 4 4\n14 8 8\n11 12 16\n16 16 16\n12 32 32\n13 64 64\n"
             f"4 96 96\n5 160 128\n6 512 256\n7 {INT_MAX} 1024\n{NUM_GLOBAL} {INT_MAX} {INT_MAX}\n")
    A = gen.powerlaw(m=2500, seed=21)
    F = A.astype(np.float32)
    run_emu_all(emu, tmp_path, call(A, A, tables={1: table}), call(F, F, "--f32", tables={1: table}))


def _alternating_units_case():
    """Rows whose symbolic units alternate LDS layouts inside one workgroup
    of the emulated grid (8 workgroups, one XCD slice each): bins of the
    larger-table launch with variants 6, 5, 6 (as in the tuned defaults)
    holding 1, 1 and 22 units of 4 rows, so that workgroup 0 runs a unit of
    each in turn.  B has 16 entries per row; a row of A with k entries has
    NIP 16k."""
    nb = 400
    # columns drawn from 48: rows of C stay short (cheap fused pass), while
    # the symbolic pass still hashes every product
    brows = [np.unique(gen.rand_below(31, 1 + r, np.arange(16), np.full(16, 48))) for r in range(nb)]
    brows = [np.concatenate([x, np.arange(48, 48 + 16 - x.size)]) for x in brows]  # exactly 16 each
    brpt = np.concatenate([[0], np.cumsum([x.size for x in brows])]).astype(np.int32)
    bcol = np.concatenate(brows).astype(np.int32)
    B = HostCSR(nb, 64, brpt, bcol, gen.rand_unit(32, 1, np.arange(bcol.size)) * 2 - 1)
    lengths = [6] * 4 + [12] * 4 + [24] * 88  # NIP 96 (bin of v6), 192 (v5), 384 (v6)
    arows = [np.unique(gen.rand_below(33, 1 + i, np.arange(k), np.full(k, nb)))[:k] for i, k in enumerate(lengths)]
    arows = [np.concatenate([x, np.setdiff1d(np.arange(nb), x)[:k - x.size]]) for x, k in zip(arows, lengths)]
    arows = [np.sort(x) for x in arows]
    rpt = np.concatenate([[0], np.cumsum([x.size for x in arows])]).astype(np.int32)
    col = np.concatenate(arows).astype(np.int32)
    A = HostCSR(len(arows), nb, rpt, col, gen.rand_unit(34, 1, np.arange(col.size)) * 2 - 1)
    table = "".join(f"{v} {cap} {INT_MAX}\n" for v, cap in
                    [(0, 8), (1, 16), (2, 32), (3, 64), (6, 128), (5, 256), (6, 768), (7, 2048), (8, 4096),
                     (SYM_MAX, INT_MAX)])
    return A, B, table


def test_symbolic_units_with_a_wave_held_back(emu, tmp_path):
    """Symbolic units of alternating LDS layouts (variants 6, 5, 6) in one
    workgroup while wave 1 of that workgroup is held back at every wave
    barrier, so the other waves run a unit or two ahead of it: the
    workgroup barrier between units of different layouts must keep them
    off the slots wave 1 still uses, and the result must equal the host
    product.  (Built without that barrier, the same run gives a wrong count
    or an out-of-bounds read only in some runs, so it is not kept as a
    planted-mistake check.)"""
    A, B, table = _alternating_units_case()
    fa, fb, ft = tmp_path / "a.csr", tmp_path / "b.csr", tmp_path / "sym.txt"
    A.save_to_filename(str(fa))
    B.save_to_filename(str(fb))
    ft.write_text(table)
    res = subprocess.run([emu, str(fa), str(fb), "--sym-table", str(ft)], stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True, timeout=600,
                         env=dict(os.environ, EMU_SLOW_WAVE="1:100"))
    assert res.returncode == 0 and "structure identical" in res.stdout, res.stdout[-3000:]


def test_rectangular_and_empty_rows(emu, tmp_path):
    A = gen.random_csr(120, 90, 0.05, seed=1, empty_rows=0.3)
    B = gen.random_csr(90, 70, 0.06, seed=2, empty_rows=0.2)
    run_emu(emu, tmp_path, A, B)


def test_nip_per_row(emu, tmp_path):
    """NIP per row, read through the symbolic bins: a table of one-wide NIP
    ranges (bin b takes the rows of NIP b + 1) counts the rows of every NIP
    up to 14, so a length credited to the wrong row changes a count.  Rows
    of up to six entries sit beside rows of 700 and 1500 entries, whose
    64-row groups take several rounds of the NIP pass, and the 3000 rows
    end in a partial tile.  The debug log gives the counts, the total and
    the maximum; the oracle's NIP per row gives the expected ones."""
    rng = np.random.default_rng(41)
    n = 4096
    b_len = rng.integers(0, 4, n)
    b_rpt = np.concatenate([[0], np.cumsum(b_len)]).astype(np.int32)
    b_col = np.concatenate([np.sort(rng.choice(64, k, replace=False)) for k in b_len]).astype(np.int32)
    B = HostCSR(n, 64, b_rpt, b_col, rng.uniform(-1, 1, b_col.size))
    m = 3000
    a_len = rng.integers(0, 7, m)
    a_len[rng.random(m) < 0.2] = 0
    a_len[[5, 1900, 2990]] = [700, 1500, 40]
    a_rpt = np.concatenate([[0], np.cumsum(a_len)]).astype(np.int32)
    a_col = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for k in a_len]).astype(np.int32)
    A = HostCSR(m, n, a_rpt, a_col, rng.uniform(-1, 1, a_col.size))
    nip = oracle.nip_rows(A, B)
    caps = list(range(1, 15))
    table = "".join(f"0 {c} {INT_MAX}\n" for c in caps) + f"8 4000 {INT_MAX}\n{SYM_MAX} {INT_MAX} {INT_MAX}\n"
    out = run_emu(emu, tmp_path, A, B, tables={0: table}, env={"SPG_LOG_LEVEL": "debug"})
    want = [int(np.sum(nip == c)) for c in caps]
    want += [int(np.sum((nip > 14) & (nip <= 4000))), int(np.sum(nip > 4000))]
    assert want[-2] >= 2 and sum(want[:14]) > 1500, want
    line = next(x for x in out.splitlines() if "symbolic phase" in x)
    assert f"NIP {int(nip.sum())} (max {int(nip.max())} per row)" in line, line
    got = line.split("rows per symbolic bin [")[1].split("]")[0]
    assert [int(x) for x in got.split()] == want, line


def test_scanned_tile_prefix(emu, tmp_path):
    """The scan stage reads each tile's prefix off one scanned array of tile
    sums, the path of multiplies of 2048 tiles or more, forced here from one
    tile: a product of three tiles and a server's batches of several
    products (each product's prefix starts at its own first tile) must
    equal the host product."""
    env = {"SPG_TILE_PREFIX_MIN": "1"}
    P = gen.powerlaw(m=4500, seed=11)
    Q = gen.powerlaw(m=600, seed=9)
    out = run_emu_all(emu, tmp_path, call(P, P, env=env), call(Q, Q, "--serve", "4", str(1 << 30), "4", env=env))
    assert out[1].count("structure identical") == 4


@pytest.mark.parametrize("variant", sorted(SYM))
def test_symbolic_variant(emu, tmp_path, variant):
    lanes, ec, t = SYM[variant]
    cap = t // 2
    table = {0: f"{variant} {cap} {INT_MAX}\n{SYM_MAX} {INT_MAX} {INT_MAX}\n"}
    m = rows_for(lanes)
    k = int(np.sqrt(cap))
    run_emu_all(emu, tmp_path, call(*squares(m, k, seed=variant), tables=table),
                call(*chunked(m, min(cap, 2 * ec + 1), seed=variant), tables=table))


@pytest.mark.parametrize("variant", sorted(NUM))
def test_fused_variant(emu, tmp_path, variant):
    lanes, ec, p = NUM[variant]
    table = {1: f"{variant} {INT_MAX} {p}\n{NUM_GLOBAL} {INT_MAX} {INT_MAX}\n"}
    m = rows_for(lanes)
    run_emu_all(emu, tmp_path, call(*squares(m, int(np.sqrt(p)), seed=10 + variant), tables=table),
                call(*chunked(m, min(p, 2 * ec + 1), seed=10 + variant), "--f32", tables=table))


WAVE_ROWS = [(0, v) for v, (lanes, _, _) in sorted(SYM.items()) if lanes == 64] + \
            [(1, v) for v, (lanes, _, _) in sorted(NUM.items()) if lanes == 64]


@pytest.mark.parametrize("phase,variant", WAVE_ROWS)
def test_wave_rows_count_entry_starts(emu, tmp_path, phase, variant):
    """Rows of one full wave find each product's A entry by counting entry
    starts in a bitmap: chunks whose products span more than one round
    (symbolic: rows of ~60 entries and NIP near the table size; fused: rows
    of ~120 entries, two chunks, ~2000 products), with B rows of uneven
    length and 40 % of them empty (dropped from the chunk before counting)."""
    if phase == 0:
        _, _, t = SYM[variant]
        table = {0: f"{variant} {t - 1} {INT_MAX}\n{SYM_MAX} {INT_MAX} {INT_MAX}\n"}
        A = gen.random_csr(8, 400, 0.15, seed=60 + variant)
        ncol = 2000
        per_row = 0.8 * (t - 1) / (60 * 0.6)
    else:
        _, _, p = NUM[variant]
        table = {1: f"{variant} {INT_MAX} {p}\n{NUM_GLOBAL} {INT_MAX} {INT_MAX}\n"}
        A = gen.random_csr(8, 400, 0.3, seed=60 + variant)
        ncol = min(p, 400)
        per_row = 2.5 * 64 * 12 / (120 * 0.6)
    B = gen.random_csr(400, ncol, min(0.9, per_row / ncol), seed=70 + variant, empty_rows=0.4)
    run_emu_all(emu, tmp_path, call(A, B, tables=table),
                call(A.astype(np.float32), B.astype(np.float32), "--f32", tables=table))


def _heavy_row_case():
    """A row that overflows the largest LDS table in the count pass and every
    LDS capacity in the fused pass, plus three ordinary rows."""
    nb = 20_000
    r0 = np.unique(gen.rand_below(3, 1, np.arange(9500), np.full(9500, nb)))
    rows = [r0] + [np.unique(gen.rand_below(3, 3 + i, np.arange(3), np.full(3, nb))) for i in range(3)]
    rpt = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    A = HostCSR(len(rows), nb, rpt, col, gen.rand_unit(3, 9, np.arange(col.size)) * 2 - 1)
    keys = np.unique(np.arange(nb, dtype=np.int64)[:, None] * 200_000
                     + gen.rand_below(4, 1, np.arange(nb * 3), np.full(nb * 3, 200_000)).reshape(nb, 3))
    brpt = np.zeros(nb + 1, dtype=np.int64)
    np.cumsum(np.bincount(keys // 200_000, minlength=nb), out=brpt[1:])
    bcol = (keys % 200_000).astype(np.int32)
    B = HostCSR(nb, 200_000, brpt.astype(np.int32), bcol, gen.rand_unit(4, 2, np.arange(bcol.size)) * 2 - 1)
    return A, B


GLOBAL_TABLES = {0: f"{SYM_MAX} {INT_MAX} {INT_MAX}\n", 1: f"{NUM_GLOBAL} {INT_MAX} {INT_MAX}\n"}


def test_global_paths(emu, tmp_path):
    """The heavy-row case first with the default tables and then with every
    row forced to the global kernels."""
    A, B = _heavy_row_case()
    run_emu_all(emu, tmp_path, call(A, B), call(A, B, tables=GLOBAL_TABLES))


def test_noalloc_entry(emu, tmp_path):
    """spg_multiply_noalloc into caller arrays, scratch in the fixed device
    array: a capacity one short of nnz(C) fails with out-of-memory and
    reports the nnz needed, then repeated products match the host; the
    heavy-row case forced to the global kernels carves both global tables
    out of the fixed array too."""
    P = gen.powerlaw(m=1000, seed=5)
    S = gen.stencil27(g=5).astype(np.float32)
    A, B = _heavy_row_case()
    out = run_emu_all(emu, tmp_path, call(P, P, "--noalloc", "--repeat", "2"), call(S, S, "--f32", "--noalloc"),
                      call(A, B, "--noalloc", tables=GLOBAL_TABLES))
    assert out[0].count("structure identical") == 2
    assert "status out of device memory" in out[0]
    assert out[2].count("structure identical") == 1


def test_server_batches_and_admission(emu, tmp_path):
    """Served form on the CPU: the same product submitted many times to a
    server with a small maximum batch and a budget of about two outputs,
    so batches are split, admitted one after another and run on both
    lanes; every result must equal the host product.  A budget below one
    output must fail every request with out-of-memory."""
    A = gen.powerlaw(m=1000, seed=8)
    C = oracle.spgemm(A, A)
    need = C.nnz * 12
    out = run_emu(emu, tmp_path, A, A, "--serve", "5", str(2 * need + 64), "2")
    assert out.count("structure identical") == 5
    assert "served 5" in out
    res = subprocess.run([emu, str(tmp_path / "a.csr"), str(tmp_path / "b.csr"), "--serve", "2", str(need // 2), "2"],
                         stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert res.returncode == 0, res.stdout[-3000:]
    assert res.stdout.count("out of device memory") == 2
