"""Autotuner of the bin tables.

The reference fixes, per GPU architecture, which kernel shape (block size,
rows per block, hash-table size) serves which range of row sizes in a
hand-written table (src/libproposal/include/proposal/parameters.cuh:22-413).
Here the shapes are the compiled variants listed by `spg_get_variants`, the
ranges are runtime bin tables (`spg_set_tuning`), and this tuner picks them
by measurement on the GPU it runs on:

  * each bin of the default tables is a slot with caps (NIP, nnz); every
    compiled variant whose table / capacity fits the slot's caps is a
    candidate for it;
  * slots are tuned one at a time, greedily, in two passes: each candidate
    replaces the slot's variant, or moves one of the slot's caps (halved or
    doubled, where the variant still admits it), and the whole multiply is
    timed (median of `reps` calls after warmups) on every tuning workload;
    the score is the sum over workloads of time / time with the starting
    tables; a candidate that wins by more than 1 % is timed again beside
    the current tables and kept only if it wins a second time;
  * the best tables are printed and, with --write, written as a C++ include
    (csrc/tuned_tables.inc) that the library compiles in as its defaults.

    python -m gpuspgemm_amd.tune [--reps 5] [--write PATH] [--json PATH]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import generators as gen
from . import SpgError, device_info, get_tuning, load_library, proposal, proposal_batched, set_tuning, variant_list

INT_MAX = 2**31 - 1


def parse_table(text: str) -> list[tuple[int, int, int]]:
    return [tuple(int(x) for x in line.split()) for line in text.strip().splitlines()]


def table_text(rows: list[tuple[int, int, int]]) -> str:
    return "".join(f"{v} {nip} {nnz}\n" for v, nip, nnz in rows)


def sym_candidates(slot: tuple[int, int, int], variants: list[dict]) -> list[int]:
    """Multi-bin symbolic variants whose table holds the slot's NIP cap at
    most half full and is at most 16x larger."""
    _, nip_cap, _ = slot
    return [v["id"] for v in variants
            if not v["max"] and 2 * nip_cap <= v["table"] <= 16 * nip_cap]


def num_candidates(slot: tuple[int, int, int], variants: list[dict]) -> list[int]:
    """Fused LDS variants whose capacity admits the slot's nnz cap and is at
    most 8x larger; for a slot whose NIP cap is finite, only variants whose
    lanes can stage that many products in registers."""
    _, nip_cap, nnz_cap = slot
    out = []
    for v in variants:
        if v["global"] or not (nnz_cap <= v["capacity"] <= 8 * nnz_cap):
            continue
        if nip_cap != INT_MAX and v["lanes"] * v["slots"] < nip_cap:
            continue
        out.append(v["id"])
    return out


class Workload:
    """A tuning workload: device-resident operands and a call that runs one
    multiply (or one batch) of it."""

    def __init__(self, name: str, pairs: list):
        self.name = name
        self.dev = [(a.to_device(), b.to_device()) for a, b in pairs]

    def run(self) -> None:
        if len(self.dev) == 1:
            proposal(*self.dev[0]).free()
        else:
            for c in proposal_batched([a for a, _ in self.dev], [b for _, b in self.dev]):
                c.free()

    def time(self, reps: int, warmup: int = 2) -> float:
        for _ in range(warmup):
            self.run()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            self.run()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    def free(self) -> None:
        for a, b in self.dev:
            a.free()
            b.free()


def cap_moves(phase: int, slot: tuple[int, int, int], variants: list[dict]) -> list[tuple[int, int, int]]:
    """The slot with one cap halved or doubled, where the variant allows it."""
    v, nip, nnz = slot
    out = []
    info = next(x for x in variants if x["id"] == v)
    if phase == 0:
        for c in (nip // 2, nip * 2):
            if 1 <= c < info["table"] and c != nip:
                out.append((v, c, nnz))
    else:
        if nip != INT_MAX:
            for c in (nip // 2, nip * 2):
                if c >= 1 and c != nip:
                    out.append((v, c, nnz))
        for c in (nnz // 2, nnz * 2):
            if 1 <= c <= info["capacity"] and c != nnz:
                out.append((v, nip, c))
    return out


def default_workloads(quick: bool = False) -> list[Workload]:
    P = gen.powerlaw(m=250_000 if quick else 1_000_000)
    S = gen.stencil27(g=64 if quick else 96)
    out = [Workload("powerlaw", [(P, P)]), Workload("stencil", [(S, S)])]
    # instances drawn like the batched workload's (rows log-uniform in
    # [1e3, 2e5]); fewer of them
    out.append(Workload("batched", [gen.batched_instance(i) for i in range(16 if quick else 64)]))
    return out


def tune(workloads: list[Workload], reps: int = 5, passes: int = 2, log=print) -> dict:
    load_library()
    sv, nv = variant_list(0), variant_list(1)
    set_tuning(0, None)
    set_tuning(1, None)
    tables = {0: parse_table(get_tuning(0)), 1: parse_table(get_tuning(1))}
    for w in workloads:  # first calls pay one-time costs (code loading, pool growth)
        w.time(1, warmup=3)
    base = {w.name: w.time(reps) for w in workloads}
    log(f"start: " + ", ".join(f"{k} {1e3 * v:.3f} ms" for k, v in base.items()))

    def score(t0, t1) -> tuple[float, dict]:
        try:
            set_tuning(0, table_text(t0))
            set_tuning(1, table_text(t1))
        except SpgError:  # a move that breaks the table rules
            return float("inf"), {}
        ts = {w.name: w.time(reps) for w in workloads}
        return sum(ts[k] / base[k] for k in ts), ts

    best, best_ts = score(tables[0], tables[1])
    trials = []
    for _ in range(passes):
        for phase in (1, 0):
            for i in range(len(tables[phase]) - 1):  # the last bin (max / global variant) is fixed
                slot = tables[phase][i]
                vs = sv if phase == 0 else nv
                cands = [(v, slot[1], slot[2]) for v in (sym_candidates if phase == 0 else num_candidates)(slot, vs)
                         if v != slot[0]]
                cands += cap_moves(phase, slot, vs)
                log(f"phase {phase} slot {i} {slot}: {len(cands)} candidates, score {best:.4f}")
                for cand in cands:
                    trial = {p: list(t) for p, t in tables.items()}
                    trial[phase][i] = cand
                    s, ts = score(trial[0], trial[1])
                    trials.append({"phase": phase, "slot": i, "bin": list(cand), "score": s, "ms": {
                        k: round(1e3 * x, 4) for k, x in ts.items()}})
                    if s < best * 0.99:  # a candidate: it wins by more than noise once
                        # confirm: time the current tables and the candidate again,
                        # one after the other; keep the change only if it wins again
                        s_cur, ts_cur = score(tables[0], tables[1])
                        s2, ts2 = score(trial[0], trial[1])
                        if s2 < s_cur * 0.99:
                            log(f"phase {phase} slot {i}: {slot} -> {cand}, score {s_cur:.4f} -> {s2:.4f} "
                                f"(first {best:.4f} -> {s:.4f})")
                            best, best_ts, tables = s2, ts2, trial
                            slot = cand
                        else:
                            best, best_ts = s_cur, ts_cur
    set_tuning(0, table_text(tables[0]))
    set_tuning(1, table_text(tables[1]))
    return {"device": device_info(), "workloads": [w.name for w in workloads],
            "start_ms": {k: round(1e3 * v, 4) for k, v in base.items()},
            "tuned_ms": {k: round(1e3 * v, 4) for k, v in best_ts.items()}, "score": best,
            "sym": tables[0], "num": tables[1], "trials": trials}


def write_include(path: str, result: dict) -> None:
    def rows(t):
        return ", ".join("{%d, %s, %s}" % (v, "INT_MAX" if nip == INT_MAX else nip,
                                           "INT_MAX" if nnz == INT_MAX else nnz) for v, nip, nnz in t)
    dev = result["device"]
    with open(path, "w") as fp:
        fp.write("// Bin tables picked by gpuspgemm_amd/tune.py on %s (%d CUs) from the workloads %s.\n"
                 % (dev["arch"], dev["compute_units"], ", ".join(result["workloads"])))
        fp.write("// Median multiply time before -> after (ms): %s.\n" % ", ".join(
            f"{k} {result['start_ms'][k]} -> {result['tuned_ms'][k]}" for k in result["start_ms"]))
        fp.write("// Regenerate with `python -m gpuspgemm_amd.tune --write <this file>` and rebuild.\n")
        fp.write("// Rows: {variant, nip_cap, nnz_cap}; a row goes to the first bin that admits it.\n")
        fp.write("constexpr TunedBin kTunedSym[] = {%s};\n" % rows(result["sym"]))
        fp.write("constexpr TunedBin kTunedNum[] = {%s};\n" % rows(result["num"]))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--quick", action="store_true", help="smaller workloads")
    ap.add_argument("--write", metavar="PATH", help="write the tuned tables as a C++ include")
    ap.add_argument("--json", metavar="PATH", help="write the full tuning record")
    args = ap.parse_args(argv)
    ws = default_workloads(args.quick)
    try:
        res = tune(ws, reps=args.reps, passes=args.passes, log=lambda s: print(s, flush=True))
    finally:
        for w in ws:
            w.free()
    print("sym:\n" + table_text(res["sym"]) + "num:\n" + table_text(res["num"]), flush=True)
    if args.write:
        write_include(args.write, res)
    if args.json:
        with open(args.json, "w") as fp:
            json.dump(res, fp, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
