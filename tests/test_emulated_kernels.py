"""The library's kernels and host driver, compiled as host C++ against the
CPU emulation of the HIP API in tests/emu (one thread per GPU lane) and run
under AddressSanitizer on small inputs: every bin variant, the global
fallbacks, and out-of-bounds accesses caught on the CPU.  Results are
checked against the host Gustavson product by tests/emu/emu_main.cpp."""
import os
import subprocess

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


@pytest.fixture(scope="module")
def emu():
    os.makedirs(BUILD, exist_ok=True)
    srcs = [os.path.join(CSRC, f) for f in ("k_rows.hip", "k_sym.hip", "k_symnum.hip", "api.hip")]
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-pthread",
           "-Wno-unused-result", "-I", os.path.join(EMU, "include"), "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "gpuspgemm_amd", "apps"), "-x", "c++", *srcs, "-x", "none",
           os.path.join(EMU, "emu_runtime.cpp"), os.path.join(EMU, "emu_main.cpp"), "-o", EXE]
    subprocess.run(cmd, check=True)
    return EXE


def run_emu(exe, tmp_path, A, B, *extra, tables=None):
    fa, fb = tmp_path / "a.csr", tmp_path / "b.csr"
    A.save_to_filename(str(fa))
    B.save_to_filename(str(fb))
    args = [exe, str(fa), str(fb), *extra]
    for phase, text in (tables or {}).items():
        f = tmp_path / f"table{phase}.txt"
        f.write_text(text)
        args += ["--sym-table" if phase == 0 else "--num-table", str(f)]
    res = subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert res.returncode == 0, res.stdout[-4000:]
    assert "structure identical" in res.stdout
    return res.stdout


def test_uniform_and_stencil(emu, tmp_path):
    A = gen.uniform8(m=512)
    run_emu(emu, tmp_path, A, A)
    S = gen.stencil27(g=6)
    run_emu(emu, tmp_path, S, S)
    run_emu(emu, tmp_path, S.astype(np.float32), S.astype(np.float32), "--f32")


def test_powerlaw_mixed_bins(emu, tmp_path):
    A = gen.powerlaw(m=4000, seed=5)
    run_emu(emu, tmp_path, A, A, "--repeat", "2")


def test_rectangular_and_empty_rows(emu, tmp_path):
    A = gen.random_csr(120, 90, 0.05, seed=1, empty_rows=0.3)
    B = gen.random_csr(90, 70, 0.06, seed=2, empty_rows=0.2)
    run_emu(emu, tmp_path, A, B)


SYM_T = [16, 32, 64, 128, 256, 512, 1024, 4096, 8192]
NUM_P = [8, 16, 32, 64, 128, 128, 512, 1024, 2048, 8192]


@pytest.mark.parametrize("variant", range(len(SYM_T)))
def test_symbolic_variant(emu, tmp_path, variant):
    A = gen.powerlaw(m=1500, seed=7)
    table = f"{variant} {SYM_T[variant] // 2} {INT_MAX}\n9 {INT_MAX} {INT_MAX}\n"
    run_emu(emu, tmp_path, A, A, tables={0: table})


@pytest.mark.parametrize("variant", range(len(NUM_P)))
def test_fused_variant(emu, tmp_path, variant):
    A = gen.powerlaw(m=1500, seed=7)
    table = f"{variant} {INT_MAX} {NUM_P[variant]}\n10 {INT_MAX} {INT_MAX}\n"
    run_emu(emu, tmp_path, A, A, tables={1: table})


def test_global_paths(emu, tmp_path):
    """Every row through the largest-table and global-memory kernels,
    including a row that overflows the largest LDS table."""
    nb = 40_000
    r0 = np.unique(gen.rand_below(3, 1, np.arange(12000), np.full(12000, nb)))
    rows = [r0] + [np.unique(gen.rand_below(3, 3 + i, np.arange(3), np.full(3, nb))) for i in range(20)]
    rpt = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.concatenate(rows).astype(np.int32)
    A = HostCSR(len(rows), nb, rpt, col, gen.rand_unit(3, 9, np.arange(col.size)) * 2 - 1)
    keys = np.unique(np.arange(nb, dtype=np.int64)[:, None] * 200_000
                     + gen.rand_below(4, 1, np.arange(nb * 3), np.full(nb * 3, 200_000)).reshape(nb, 3))
    brpt = np.zeros(nb + 1, dtype=np.int64)
    np.cumsum(np.bincount(keys // 200_000, minlength=nb), out=brpt[1:])
    bcol = (keys % 200_000).astype(np.int32)
    B = HostCSR(nb, 200_000, brpt.astype(np.int32), bcol, gen.rand_unit(4, 2, np.arange(bcol.size)) * 2 - 1)
    run_emu(emu, tmp_path, A, B, tables={0: f"9 {INT_MAX} {INT_MAX}\n", 1: f"10 {INT_MAX} {INT_MAX}\n"})
    run_emu(emu, tmp_path, A, B)
