"""The `.csr` binary format (src/libutils/include/utils/csr.cuh:119-191;
writer tasks.py:67-72): little-endian int32 nnz, m, n, int32 rpt[m+1],
int32 col[nnz], float64 val[nnz] for every value type.  Fixtures are bytes
written from the format description, not by the code under test."""
import struct

import numpy as np
import pytest

from gpuspgemm_amd import generators as gen
from gpuspgemm_amd.csr import HostCSR

# 2 x 3 matrix [[1.5, 0, -2], [0, 3.25, 0]]
FIXTURE = struct.pack("<3i", 3, 2, 3) + struct.pack("<3i", 0, 2, 3) + struct.pack("<3i", 0, 2, 1) \
    + struct.pack("<3d", 1.5, -2.0, 3.25)


def test_load_fixture(tmp_path):
    f = tmp_path / "m.csr"
    f.write_bytes(FIXTURE)
    A = HostCSR.load_from_filename(str(f))
    assert (A.nnz, A.m, A.n) == (3, 2, 3)
    assert A.rpt.tolist() == [0, 2, 3] and A.col.tolist() == [0, 2, 1]
    assert A.val.tolist() == [1.5, -2.0, 3.25] and A.val.dtype == np.float64


def test_save_is_byte_identical(tmp_path):
    f = tmp_path / "m.csr"
    f.write_bytes(FIXTURE)
    A = HostCSR.load_from_filename(str(f))
    g = tmp_path / "n.csr"
    A.save_to_filename(str(g))
    assert g.read_bytes() == FIXTURE


def test_float32_narrows_on_load_and_widens_on_save(tmp_path):
    f = tmp_path / "m.csr"
    f.write_bytes(struct.pack("<3i", 1, 1, 1) + struct.pack("<2i", 0, 1) + struct.pack("<i", 0)
                  + struct.pack("<d", 0.1))
    A = HostCSR.load_from_filename(str(f), dtype=np.float32)
    assert A.val.dtype == np.float32 and A.val[0] == np.float32(0.1)
    g = tmp_path / "n.csr"
    A.save_to_filename(str(g))
    assert struct.unpack("<d", g.read_bytes()[-8:])[0] == float(np.float32(0.1))


def test_size_errors(tmp_path):
    f = tmp_path / "m.csr"
    f.write_bytes(FIXTURE[:8])
    with pytest.raises(RuntimeError, match="too small"):
        HostCSR.load_from_filename(str(f))
    f.write_bytes(FIXTURE[:-1])
    with pytest.raises(RuntimeError, match="Failed to read"):
        HostCSR.load_from_filename(str(f))
    f.write_bytes(FIXTURE + b"\0")
    with pytest.raises(RuntimeError, match="too large"):
        HostCSR.load_from_filename(str(f))
    with pytest.raises(RuntimeError, match="Failed to open"):
        HostCSR.load_from_filename(str(tmp_path / "missing.csr"))


def test_round_trip_and_equality_rule(tmp_path):
    A = gen.powerlaw(m=3000, seed=4)
    f = tmp_path / "p.csr"
    A.save_to_filename(str(f))
    B = HostCSR.load_from_filename(str(f))
    assert B == A
    # csr.cuh:232-238: |l-r| <= 1e-8 + 1e-5 max(|l|,|r|)
    C = HostCSR(B.m, B.n, B.rpt, B.col, B.val.copy())
    C.val[0] = C.val[0] * (1 + 5e-6)
    assert C == A
    C.val[0] = A.val[0] + 1e-3 + 1e-5 * abs(A.val[0])
    assert not (C == A)
    D = HostCSR(B.m, B.n, B.rpt, B.col.copy(), B.val)
    D.col[0] += 1
    assert not (D == A)
