"""Benchmark harness with the reference's protocol and output.

The reference's task runner (tasks.py:96-180, `invoke benchmark`) runs an
app binary on every matrix of its list as A . A with
`benchmark --runs R --warmups W --pause P`, parses the app's `Run` lines
with one regular expression (tasks.py:105-107), prints a table of mean,
standard deviation and GFLOPS per matrix, and can save every run as CSV
rows `algorithm,matrix,time_ns,flops`.  This module does the same with the
apps built here (`bin/proposal`, `bin/proposal_f32`, and the rocSPARSE
comparator `bin/rocsparse`, `bin/rocsparse_f32`), so the reference's table
can be re-measured on MI355X with the same protocol.  The SuiteSparse
matrices are not downloaded (no network): supply them as `.csr` files in
`--data`, or write the synthetic workloads there with `--generate`.

    python -m gpuspgemm_amd.harness --data DIR [--target proposal] [--runs 10]
        [--warmups 10] [--pause 50] [--matrices a,b] [--output results.csv]
    python -m gpuspgemm_amd.harness --generate DIR
"""
from __future__ import annotations

import argparse
import csv
import os
import re
import subprocess
import sys

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(PKG, "bin")

# tasks.py:24-37 (names only; files are supplied locally)
REFERENCE_MATRICES = [
    "m133-b3", "webbase-1M", "mc2depi", "mario002", "majorbasis", "shipsec1",
    "rma10", "pdb1HYS", "pkustk12", "SiO2", "case39", "gupta3",
]
# the same pattern as tasks.py:105-107
RUN_PATTERN = re.compile(r"Run\s+(?P<id>\d+): (?P<time>.+)ns [|] (?P<flops>\d+[.]\d*) FLOPS")
FORMAT_WIDTH = 7


def parse_runs(stdout: str) -> tuple[list[float], list[float]]:
    """Times (ns) and FLOPS of the `Run` lines; raises on a malformed line."""
    times, flops = [], []
    for line in stdout.splitlines():
        if not line.startswith("Run"):
            continue
        m = RUN_PATTERN.fullmatch(line)
        if not m:
            raise ValueError(f"wrong output format: {line!r}")
        times.append(float(m.group("time")))
        flops.append(float(m.group("flops")))
    return times, flops


def run_matrix(binary: str, path: str, runs: int, warmups: int, pause: int,
               timeout: float = 1800.0) -> tuple[list[float], list[float]]:
    res = subprocess.run([binary, path, path, "benchmark", f"--runs={runs}", f"--warmups={warmups}",
                          f"--pause={pause}"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                         timeout=timeout)
    times, flops = parse_runs(res.stdout)
    if not times:  # the app failed on this matrix: zeros, as tasks.py records them
        sys.stderr.write(res.stdout[-2000:] + res.stderr[-2000:])
        times, flops = [0.0] * runs, [0.0] * runs
    if len(times) != runs:
        raise ValueError(f"{len(times)} runs reported, {runs} expected")
    return times, flops


def benchmark(data: str, target: str = "proposal", matrices: list[str] | None = None, runs: int = 10,
              warmups: int = 10, pause: int = 50, output: str | None = None, binary: str | None = None,
              out=sys.stdout) -> list[dict]:
    binary = binary or os.path.join(BIN, target)
    if matrices is None:
        matrices = sorted(f[:-4] for f in os.listdir(data) if f.endswith(".csr"))
    width = max([len(m) for m in matrices] + [6])
    print(f"Algorithm: {target}", file=out)
    print(f"Runs: {runs}", file=out)
    print(f"Warmups: {warmups}", file=out)
    print(f"Pause: {pause}", file=out)
    print("Unit: us", file=out)
    print(f"{'Matrix':{width}s} | {'Mean':^{FORMAT_WIDTH}s} (+- {'Std':^{FORMAT_WIDTH}s}) | {'GFLOPS':^{FORMAT_WIDTH}s}",
          file=out)
    rows = []
    for name in matrices:
        times, flops = run_matrix(binary, os.path.join(data, f"{name}.csr"), runs, warmups, pause)
        rows += [{"algorithm": target, "matrix": name, "time_ns": t, "flops": f} for t, f in zip(times, flops)]
        us = np.array(times) / 1000.0
        gf = np.array(flops) / 1e9
        print(f"{name:{width}s} | {np.mean(us):{FORMAT_WIDTH}.1f} (+- {np.std(us):{FORMAT_WIDTH}.1f})"
              f" | {np.mean(gf):{FORMAT_WIDTH}.2f}", file=out, flush=True)
    if output:
        with open(output, "w", newline="") as fp:
            w = csv.DictWriter(fp, fieldnames=["algorithm", "matrix", "time_ns", "flops"])
            w.writeheader()
            w.writerows(rows)
    return rows


def generate(data: str) -> list[str]:
    """Write the synthetic workloads as .csr files (values as in the benchmark)."""
    from . import generators as gen
    os.makedirs(data, exist_ok=True)
    mats = {"uniform8": gen.uniform8(), "powerlaw": gen.powerlaw(), "stencil27": gen.stencil27()}
    for name, A in mats.items():
        A.save_to_filename(os.path.join(data, f"{name}.csr"))
    return sorted(mats)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--data", help="directory of <matrix>.csr files")
    ap.add_argument("--generate", metavar="DIR", help="write the synthetic workloads to DIR and exit")
    ap.add_argument("--target", default="proposal", help="app in gpuspgemm_amd/bin")
    ap.add_argument("--matrices", help="comma-separated names (default: every .csr in --data)")
    ap.add_argument("--reference-list", action="store_true", help="use the reference's matrix list")
    ap.add_argument("--runs", type=int, default=10)
    ap.add_argument("--warmups", type=int, default=10)
    ap.add_argument("--pause", type=int, default=50)
    ap.add_argument("--output", help="CSV of every run: algorithm,matrix,time_ns,flops")
    args = ap.parse_args(argv)
    if args.generate:
        print("\n".join(generate(args.generate)))
        return 0
    if not args.data:
        ap.error("--data is required")
    mats = REFERENCE_MATRICES if args.reference_list else (args.matrices.split(",") if args.matrices else None)
    benchmark(args.data, args.target, mats, args.runs, args.warmups, args.pause, args.output)
    return 0


if __name__ == "__main__":
    sys.exit(main())
