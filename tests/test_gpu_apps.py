"""The apps on the GPU: `proposal` and the rocSPARSE comparator validate
against the host product and run under the reference's benchmark protocol
through the harness (tasks.py:96-180)."""
import csv
import os
import subprocess

import pytest

from gpuspgemm_amd import generators as gen
from gpuspgemm_amd import harness

pytestmark = pytest.mark.gpu
BIN = harness.BIN


@pytest.fixture(scope="module")
def data(tmp_path_factory):
    d = tmp_path_factory.mktemp("csr")
    gen.powerlaw(m=20_000, seed=3).save_to_filename(str(d / "powerlaw20k.csr"))
    gen.stencil27(g=16, seed=4).save_to_filename(str(d / "stencil16.csr"))
    return d


@pytest.mark.parametrize("app", ["proposal", "proposal_f32", "rocsparse", "rocsparse_f32"])
def test_validate(gpu_lib, data, app):
    f = str(data / "powerlaw20k.csr")
    r = subprocess.run([os.path.join(BIN, app), f, f, "validate"], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Validation succeeded" in r.stdout, r.stdout[-2000:]


@pytest.mark.parametrize("app", ["proposal", "proposal_f32"])
def test_save_mode(gpu_lib, data, tmp_path, app):
    """`save <output>` end to end (common.hpp:142-152 -> csr.cuh:162-191):
    the file has the .csr layout (int32 nnz, m, n, rpt, col, then float64
    values, FP32 results widened), its structure bytes equal the oracle's
    C and its values are within the tolerance of the app's value type."""
    import numpy as np
    import oracle
    from gpuspgemm_amd.csr import HostCSR
    from parity import TOL
    f = str(data / "powerlaw20k.csr")
    out = str(tmp_path / f"C_{app}.csr")
    r = subprocess.run([os.path.join(BIN, app), f, f, "save", out], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    dtype = np.float32 if app.endswith("f32") else np.float64
    A = HostCSR.load_from_filename(f, dtype=dtype)
    ref = oracle.spgemm(A, A)
    raw = open(out, "rb").read()
    nnz, m, n = np.frombuffer(raw[:12], dtype="<i4")
    assert (nnz, m, n) == (ref.nnz, A.m, A.n)
    assert len(raw) == 12 + 4 * (m + 1) + 12 * nnz
    rpt = np.frombuffer(raw, dtype="<i4", count=m + 1, offset=12)
    col = np.frombuffer(raw, dtype="<i4", count=nnz, offset=12 + 4 * (m + 1))
    val = np.frombuffer(raw, dtype="<f8", count=nnz, offset=12 + 4 * (m + 1 + nnz))
    np.testing.assert_array_equal(rpt, ref.rpt.astype(np.int32))
    np.testing.assert_array_equal(col, ref.col)
    if dtype == np.float32:  # widened from FP32: exactly representable in FP32
        np.testing.assert_array_equal(val, val.astype(np.float32).astype(np.float64))
    scale = np.where(ref.absval > 0, ref.absval, 1.0)
    assert float((np.abs(val - ref.val) / scale).max()) <= TOL[np.dtype(dtype)]


def test_harness_protocol(gpu_lib, data, tmp_path):
    for target in ("proposal", "rocsparse"):
        out = tmp_path / f"{target}.csv"
        rows = harness.benchmark(str(data), target, None, runs=3, warmups=1, pause=0, output=str(out))
        assert len(rows) == 6 and all(r["time_ns"] > 0 and r["flops"] > 0 for r in rows)
        with open(out) as fp:
            assert len(list(csv.DictReader(fp))) == 6
