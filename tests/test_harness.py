"""Benchmark harness (tasks.py:96-180 protocol) without a GPU: the run-line
pattern against the apps' printf format, table and CSV output through a
stand-in app that prints what the apps print."""
import csv
import os
import stat
import sys

import pytest

from gpuspgemm_amd import harness


def test_pattern_matches_app_format():
    # the apps print "Run %2d: %lldns | %.3f FLOPS" (apps/common.hpp)
    lines = "A: 4 x 4 (8 non-zero elements)\nRun  1: 551234ns | 31391818181.818 FLOPS\n" \
            "Run 10: 549000ns | 31519489981.785 FLOPS\nStats: mean 0.55 ms\n"
    times, flops = harness.parse_runs(lines)
    assert times == [551234.0, 549000.0]
    assert flops == pytest.approx([31391818181.818, 31519489981.785])
    with pytest.raises(ValueError):
        harness.parse_runs("Run 1: garbage\n")


def test_table_and_csv(tmp_path):
    app = tmp_path / "fake_app"
    app.write_text(f"#!{sys.executable}\n"
                   "import sys\n"
                   "runs = int([a for a in sys.argv if a.startswith('--runs=')][0].split('=')[1])\n"
                   "for i in range(runs):\n"
                   "    print(f'Run {i + 1:2d}: {1000 * (i + 1)}ns | {1e9 * (i + 1):.3f} FLOPS')\n")
    app.chmod(app.stat().st_mode | stat.S_IEXEC)
    data = tmp_path / "data"
    data.mkdir()
    for name in ("m1", "m2"):
        (data / f"{name}.csr").write_bytes(b"")
    out = tmp_path / "res.csv"
    with open(tmp_path / "table.txt", "w") as fp:
        rows = harness.benchmark(str(data), "fake", None, runs=3, warmups=0, pause=0, output=str(out),
                                 binary=str(app), out=fp)
    assert len(rows) == 6
    table = (tmp_path / "table.txt").read_text()
    assert "Algorithm: fake" in table and "m1" in table and "Unit: us" in table
    with open(out) as fp:
        got = list(csv.DictReader(fp))
    assert [r["matrix"] for r in got] == ["m1"] * 3 + ["m2"] * 3
    assert float(got[1]["time_ns"]) == 2000.0 and float(got[2]["flops"]) == 3e9
