"""Per-multiply GPU timeline from a rocprofv3 kernel trace.

    python -m gpuspgemm_amd.timeline <kernel_trace.csv> [--last N] [--skip K]

The trace (`rocprofv3 --kernel-trace --output-format csv`) lists every
dispatch with its hardware queue and start / end timestamps.  A multiply
starts with its `k_nip` launch (or the `k_clear` right before it); for each
of the last N multiplies before the final K this prints the kernels in start order with their
queue, the idle gaps (no kernel of the process running), the busy fraction
(union of kernel intervals over the span from the first start to the last
end) and the time of the two phases: symbolic (first kernel to the end of
the read-back that carries nnz(C)) and fused (first fused launch to the end
of the last read-back), then the means of these over the N multiplies.
The profiler adds its own per-dispatch cost, so spans here are longer than
the unprofiled wall clock.
"""
from __future__ import annotations

import argparse
import csv
import re
import sys


def short_name(name: str) -> str:
    n = name.replace("spg::", "")
    m = re.search(r"k_symnum_bin<(?:double|float), (\d+)>", n)
    if m:
        return "symnum_v" + m.group(1)
    m = re.search(r"k_symnum_multi<(?:double|float), (\d+)>", n)
    if m:
        return "symnum_multi" + m.group(1)
    if "k_sym_multi" in n:
        return "sym_multi_small" if ", 0, 4>" in n else ("sym_multi_large" if ", 7, 9>" in n else "sym_multi")
    m = re.match(r"(?:void )?(k_\w+)", n)
    return m.group(1)[2:] if m else n.split("(")[0][:32]


def load(path: str) -> list[dict]:
    rows = []
    with open(path) as fp:
        for r in csv.DictReader(fp):
            rows.append({"start": int(r["Start_Timestamp"]), "end": int(r["End_Timestamp"]),
                         "queue": int(r["Queue_Id"]), "name": short_name(r["Kernel_Name"])})
    rows.sort(key=lambda k: k["start"])
    return rows


def split_multiplies(ks: list[dict]) -> list[list[dict]]:
    starts = []
    for i, k in enumerate(ks):
        if k["name"] == "nip":
            j = i - 1 if i > 0 and ks[i - 1]["name"] == "clear" else i
            starts.append(j)
    out = []
    for a, b in zip(starts, starts[1:] + [len(ks)]):
        step = [k for k in ks[a:b] if k["name"] not in ("total_nip",) and not k["name"].startswith("__amd")]
        out.append(step)
    return out


def summarize(step: list[dict]) -> dict:
    t0 = step[0]["start"]
    t1 = max(k["end"] for k in step)
    busy, cur, gaps = 0, t0, []
    for k in step:
        if k["start"] > cur:
            gaps.append((cur, k["start"], k["name"]))
            cur = k["start"]
        if k["end"] > cur:
            busy += k["end"] - max(k["start"], cur)
            cur = k["end"]
    readbacks = [k for k in step if k["name"] == "readback"]
    fused = [k for k in step if k["name"].startswith("symnum")]
    sym_end = readbacks[0]["end"] if readbacks else t1
    queues = {}
    for k in step:
        q = queues.setdefault(k["queue"], [0, 0])
        q[0] += k["end"] - k["start"]
        q[1] += 1
    return {"t0": t0, "span_us": (t1 - t0) / 1e3, "busy_us": busy / 1e3, "busy_frac": busy / max(1, t1 - t0),
            "symbolic_us": (sym_end - t0) / 1e3,
            "fused_us": (t1 - fused[0]["start"]) / 1e3 if fused else 0.0,
            "readback_to_fused_us": (fused[0]["start"] - sym_end) / 1e3 if fused else 0.0,
            "gaps": [((a - t0) / 1e3, (b - a) / 1e3, before) for a, b, before in gaps],
            "queues": {q: (v[0] / 1e3, v[1]) for q, v in sorted(queues.items())}}


def report(path: str, last: int, skip: int = 0) -> str:
    steps = split_multiplies(load(path))
    steps = steps[max(0, len(steps) - skip - last):len(steps) - skip]
    lines = []
    means = {k: 0.0 for k in ("span_us", "busy_us", "busy_frac", "symbolic_us", "readback_to_fused_us", "fused_us")}
    for n, st in enumerate(steps):
        for k in means:
            means[k] += summarize(st)[k] / len(steps)
        s = summarize(st)
        lines.append(f"multiply {n + 1} of the last {len(steps)}: span {s['span_us']:.1f} us, busy "
                     f"{s['busy_us']:.1f} us ({100 * s['busy_frac']:.0f} %), symbolic phase {s['symbolic_us']:.1f} us, "
                     f"read-back -> first fused launch {s['readback_to_fused_us']:.1f} us, fused pass "
                     f"{s['fused_us']:.1f} us")
        lines.append("  queue  kernel-us  kernels: " + "; ".join(
            f"q{q} {us:.1f} ({c})" for q, (us, c) in s["queues"].items()))
        lines.append(f"  {'start':>8} {'end':>8} {'us':>7}  queue  kernel")
        for k in st:
            lines.append(f"  {(k['start'] - s['t0']) / 1e3:8.1f} {(k['end'] - s['t0']) / 1e3:8.1f} "
                         f"{(k['end'] - k['start']) / 1e3:7.1f}  q{k['queue']:<4} {k['name']}")
        idle = [g for g in s["gaps"] if g[1] >= 1.0]
        lines.append("  idle gaps >= 1 us (at, length, before): " + (
            ", ".join(f"{a:.1f}+{l:.1f} before {b}" for a, l, b in idle) if idle else "none"))
    if steps:
        lines.append(f"mean of {len(steps)}: span {means['span_us']:.1f} us, busy {means['busy_us']:.1f} us "
                     f"({100 * means['busy_frac']:.0f} %), symbolic phase {means['symbolic_us']:.1f} us, read-back -> "
                     f"first fused launch {means['readback_to_fused_us']:.1f} us, fused pass {means['fused_us']:.1f} us")
    return "\n".join(lines)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("trace")
    ap.add_argument("--last", type=int, default=3)
    ap.add_argument("--skip", type=int, default=0, help="leave out the final K multiplies (e.g. profiled steps)")
    args = ap.parse_args(argv)
    print(report(args.trace, args.last, args.skip))
    return 0


if __name__ == "__main__":
    sys.exit(main())
