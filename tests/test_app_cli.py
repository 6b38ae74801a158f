"""The C++ apps' command line (src/apps/common.hpp:40-82) without a GPU:
usage, argument errors and missing files exit before any device work."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "gpuspgemm_amd", "bin", "proposal")


@pytest.fixture(scope="module")
def app():
    if not os.path.exists(APP):
        from gpuspgemm_amd import build
        build.build_apps()
    return APP


def run(*args):
    return subprocess.run([*args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=60)


def test_help(app):
    r = run(app, "--help")
    assert r.returncode == 0 and "benchmark" in r.stderr and "validate" in r.stderr


def test_argument_errors(app, tmp_path):
    f = tmp_path / "a.csr"
    f.write_bytes(b"\0" * 16)
    assert run(app).returncode != 0
    assert "File does not exist" in run(app, "nope.csr", str(f), "validate").stderr
    assert run(app, str(f), str(f), "frobnicate").returncode != 0
    assert run(app, str(f), str(f), "save").returncode != 0
    assert run(app, str(f), str(f), "--loglevel=loud", "validate").returncode != 0
    assert run(app, str(f), str(f), "benchmark", "--runs=-1").returncode != 0


def test_dimension_mismatch_is_reported_before_device_work(app, tmp_path):
    from gpuspgemm_amd import generators as gen
    a, b = tmp_path / "a.csr", tmp_path / "b.csr"
    gen.random_csr(5, 7, 0.3, seed=1).save_to_filename(str(a))
    gen.random_csr(6, 5, 0.3, seed=2).save_to_filename(str(b))
    r = run(app, str(a), str(b), "validate")
    assert r.returncode != 0 and "must match" in r.stderr
