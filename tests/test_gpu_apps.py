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


def test_harness_protocol(gpu_lib, data, tmp_path):
    for target in ("proposal", "rocsparse"):
        out = tmp_path / f"{target}.csv"
        rows = harness.benchmark(str(data), target, None, runs=3, warmups=1, pause=0, output=str(out))
        assert len(rows) == 6 and all(r["time_ns"] > 0 and r["flops"] > 0 for r in rows)
        with open(out) as fp:
            assert len(list(csv.DictReader(fp))) == 6
