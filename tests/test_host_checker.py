"""The apps' `validate` checker (gpuspgemm_amd/apps/host_spgemm.hpp, a host
Gustavson standing where the reference's apps call cuSPARSE,
src/apps/common.hpp:108-136) against the oracle: same structure, values
within the FP64 bound.  Built with g++ against tests/emu's HIP stubs; no
GPU needed."""
import os
import subprocess

import numpy as np
import pytest

from gpuspgemm_amd import generators as gen
from gpuspgemm_amd.csr import HostCSR
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")


@pytest.fixture(scope="module")
def host_product():
    build = os.path.join(EMU, "_build")
    os.makedirs(build, exist_ok=True)
    exe = os.path.join(build, f"host_product.{os.getpid()}")
    subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", "-I", os.path.join(EMU, "include"),
                    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "gpuspgemm_amd", "apps"),
                    os.path.join(EMU, "host_product.cpp"), "-o", exe], check=True)
    yield exe
    os.remove(exe)


def _cases():
    yield "powerlaw", gen.powerlaw(m=3000, seed=7), None
    yield "stencil", gen.stencil27(g=6), None
    A = gen.random_csr(300, 200, 0.03, seed=8, empty_rows=0.3)
    yield "rectangular", A, gen.random_csr(200, 150, 0.05, seed=9)
    # exact zero sums: (1, 1) . (1, -1)^T per row
    rpt = np.arange(0, 2 * 50 + 1, 2, dtype=np.int32)
    col = np.tile(np.array([0, 1], dtype=np.int32), 50)
    Z = HostCSR(50, 2, rpt, col, np.ones(100))
    W = HostCSR(2, 3, np.array([0, 1, 2], dtype=np.int32), np.array([2, 2], dtype=np.int32), np.array([1.0, -1.0]))
    yield "zero sums", Z, W


@pytest.mark.parametrize("name,A,B", list(_cases()), ids=[c[0] for c in _cases()])
def test_checker_matches_oracle(host_product, tmp_path, name, A, B):
    B = A if B is None else B
    fa, fb, fc = tmp_path / "a.csr", tmp_path / "b.csr", tmp_path / "c.csr"
    A.save_to_filename(str(fa))
    B.save_to_filename(str(fb))
    subprocess.run([host_product, str(fa), str(fb), str(fc)], check=True, timeout=120)
    C = HostCSR.load_from_filename(str(fc))
    ref = oracle.spgemm(A, B)
    assert (C.m, C.n) == (A.m, B.n)
    np.testing.assert_array_equal(C.rpt.astype(np.int64), ref.rpt)
    np.testing.assert_array_equal(C.col, ref.col)
    scale = np.where(ref.absval > 0, ref.absval, 1.0)
    assert float((np.abs(C.val - ref.val) / scale).max(initial=0.0)) <= 1e-12
