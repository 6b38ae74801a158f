#!/usr/bin/env python3
"""Benchmark of C = A.A on one MI355X: GFLOP/s (2 x intermediate products
per second, the reference's metric, src/apps/common.hpp:186-204) and HBM
GB/s of algorithmic bytes.

A step is one full multiply with A and B (two device copies, as the
reference's app makes, src/apps/common.hpp:170-171) resident in HBM: the
wall clock spans the whole call, including the allocation of C and the
final synchronisation, as src/apps/proposal.cu:28-35 times it; freeing C is
outside the timed region, as in the reference.

Default workload: the power-law matrix (1,000,000 rows, Pareto(2) row
lengths), FP64.  The stencil (96^3, 27-point) is measured beside it and
reported under "secondary".

    python bench.py [--gpus 1] [--steps 30] [--warmup 20] [--workload powerlaw]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, published
METRIC = "SpGEMM GFLOP/s (2×intermediate products) + HBM GB/s, FP64; served p50/p99"


def make_workload(name: str):
    from gpuspgemm_amd import generators as gen
    if name == "powerlaw":
        return gen.powerlaw(), {"workload": "powerlaw", "rows": 1_000_000, "row_length": "Pareto(2), mean ~3",
                                "columns": "70% within +-64 of the diagonal, 30% uniform", "seed": 1}
    if name == "stencil":
        return gen.stencil27(), {"workload": "stencil27", "grid": "96^3", "rows": 96 ** 3, "seed": 2}
    if name == "uniform8":
        return gen.uniform8(), {"workload": "uniform8", "rows": 4096, "per_row": 8, "seed": 0}
    raise SystemExit(f"unknown workload {name}")


def csr_bytes(m: int, nnz: int, s: int) -> int:
    return 4 * (m + 1) + (4 + s) * nnz


def row_nip(A, B) -> np.ndarray:
    blen = np.diff(B.rpt.astype(np.int64))
    per_entry = blen[A.col]
    cs = np.concatenate([[0], np.cumsum(per_entry)])
    return cs[A.rpt[1:].astype(np.int64)] - cs[A.rpt[:-1].astype(np.int64)]


def classify(table: str, nip: np.ndarray, nnz: np.ndarray) -> tuple[np.ndarray, list[int]]:
    """Bin of each row under a bin table (first bin whose caps admit it)."""
    bins = [tuple(int(x) for x in line.split()) for line in table.strip().splitlines()]
    out = np.full(nip.shape, len(bins) - 1, dtype=np.int64)
    todo = np.ones(nip.shape, dtype=bool)
    for b, (_, nip_cap, nnz_cap) in enumerate(bins):
        hit = todo & (nip <= nip_cap) & (nnz <= nnz_cap)
        out[hit] = b
        todo &= ~hit
    return out, [v for v, _, _ in bins]


def kernel_bytes(name: str, A, B, C_rpt: np.ndarray, s: int, sym_table: str, num_table: str,
                 merged: dict = None) -> int:
    """Algorithmic (compulsory) bytes of one launch of kernel `name`: the A
    rows it handles (row pointers, columns, values when it reads them),
    every B row those rows reference, read once, and what it writes.
    `merged` maps each merged launch set to its fused variants (the bins
    of launch symnum_multi<set>)."""
    nip = row_nip(A, B)
    nnz_c = np.diff(C_rpt.astype(np.int64))
    arow = np.diff(A.rpt.astype(np.int64))
    if name.startswith("symnum_v") or name.startswith("symnum_multi"):
        b, variants = classify(num_table, nip, nnz_c)
        if name.startswith("symnum_multi"):
            members = (merged or {}).get(int(name[len("symnum_multi"):]), ())
            bins = [i for i, v in enumerate(variants) if v in members]
        else:
            bins = [i for i, v in enumerate(variants) if v == int(name[len("symnum_v"):])]
        sel = (nnz_c > 0) & np.isin(b, bins)
        val_a = val_b = s
        write = (4 + s) * int(nnz_c[sel].sum())
    elif name == "sym_multi":
        b, variants = classify(sym_table, nip, np.zeros_like(nip))
        sel = (nip > 0) & (b < len(variants) - 1)
        val_a = val_b = 0
        write = 4 * int(sel.sum())
    else:
        return 0
    rows = np.nonzero(sel)[0]
    a_bytes = 8 * rows.size + (4 + val_a) * int(arow[rows].sum())
    mask = np.repeat(sel, arow)
    used_b = np.unique(A.col[mask])
    blen = np.diff(B.rpt.astype(np.int64))
    b_bytes = 8 * used_b.size + (4 + val_b) * int(blen[used_b].sum())
    return a_bytes + b_bytes + write


def traffic_record(workload: str, dtype: str, kernel: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fp:
            data = json.load(fp)
        return data.get(f"{workload}/{dtype}/{kernel}")
    except (OSError, ValueError):
        return None


def run_gpu(A, steps: int, warmup: int, pause_ms: float, dtype) -> dict:
    import gpuspgemm_amd as g
    g.load_library()
    Ah = A.astype(dtype)
    dA, dB = Ah.to_device(), Ah.to_device()
    nip = g.get_nip(dA, dB)
    for _ in range(warmup):
        if pause_ms:
            time.sleep(pause_ms / 1e3)
        g.proposal(dA, dB).free()
    # timed steps: no per-kernel events inside the timed region
    times = []
    for _ in range(steps):
        if pause_ms:
            time.sleep(pause_ms / 1e3)
        t0 = time.perf_counter()
        C = g.proposal(dA, dB)
        t1 = time.perf_counter()
        times.append(t1 - t0)
        if _ == steps - 1:
            c_rpt = C.to_host().rpt
            c_nnz = C.nnz
        C.free()
    # the same number of steps again with HIP events around every launch,
    # for the per-kernel times and the roofline of the dominant launch
    kernel_ms: dict[str, list[float]] = {}
    with g.profile() as prof:
        for _ in range(steps):
            g.proposal(dA, dB).free()
            for r in prof.collect():
                kernel_ms.setdefault(r["name"], []).append(r["ms"])
    sym_table, num_table = g.get_tuning(0), g.get_tuning(1)
    merged = {}
    for v in g.variant_list(1):
        if v["merged"] is not None:
            merged.setdefault(v["merged"], []).append(v["id"])
    dA.free()
    dB.free()
    return {"times": np.array(times), "nip": nip, "c_rpt": c_rpt, "c_nnz": c_nnz, "kernel_ms": kernel_ms,
            "sym_table": sym_table, "num_table": num_table, "merged": merged, "info": g.device_info()}


def summarize(A, res: dict, dtype_name: str, workload: str, cfg: dict) -> dict:
    s = 8 if dtype_name == "f64" else 4
    t = res["times"]
    mean = float(t.mean())
    flop = 2.0 * res["nip"]
    b_alg = 2 * csr_bytes(A.m, A.nnz, s) + csr_bytes(A.m, res["c_nnz"], s)
    # dominant kernel by total time over the timed steps
    totals = {k: float(np.sum(v)) for k, v in res["kernel_ms"].items()}
    dom = max(totals, key=totals.get)
    dom_avg_ms = float(np.mean(res["kernel_ms"][dom]))
    dom_bytes = kernel_bytes(dom, A, A, res["c_rpt"], s, res["sym_table"], res["num_table"], res["merged"])
    achieved = dom_bytes / (dom_avg_ms * 1e-3) / 1e9 if dom_bytes else None
    traffic = traffic_record(workload, dtype_name, dom)
    return {
        "value": flop / mean / 1e9,
        "ms_per_step": mean * 1e3,
        "p50_ms": float(np.percentile(t, 50)) * 1e3,
        "p99_ms": float(np.percentile(t, 99)) * 1e3,
        "hbm_gbps": b_alg / mean / 1e9,
        "hbm_frac": b_alg / mean / 1e9 / HBM_PEAK_GBPS,
        "nip": int(res["nip"]),
        "nnz_c": int(res["c_nnz"]),
        "alg_bytes": int(b_alg),
        "kernels_ms": {k: round(float(np.mean(v)), 4) for k, v in sorted(res["kernel_ms"].items())},
        "roofline": {
            "bound": "hbm", "kernel": dom,
            "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4) if achieved else None,
            "traffic": traffic, "alg_bytes_per_launch": int(dom_bytes), "avg_launch_ms": round(dom_avg_ms, 4),
        },
        "config": cfg,
    }


def cpu_baseline(A, nip: int, budget_s: float, warmups: int, runs: int) -> dict:
    """The oracle's row-parallel Gustavson timed on the host: the product
    alone (count pass and scan, then sorted columns and values into fresh
    arrays), all host cores and one thread, `warmups` untimed and up to
    `runs` timed products each (fewer when one thread would pass the
    budget), like the GPU's warmup/run protocol (common.hpp:174-204)."""
    import oracle
    threads = oracle.hw_threads()

    def leg(t: int, budget: float) -> dict:
        one = float(oracle.time_product(A, A, t, 0, 1)[0])  # also pages the operands in
        n = int(max(2, min(runs, budget / max(one, 1e-6))))
        w = min(warmups, 1) if one * (n + warmups) > budget else warmups
        times = oracle.time_product(A, A, t, w, n)
        mean = float(np.mean(times))
        return {"value": round(2.0 * nip / mean / 1e9, 4), "unit": "GFLOP/s", "cores": t, "runs": n,
                "warmups": w, "mean_ms": round(mean * 1e3, 3),
                "p50_ms": round(float(np.percentile(times, 50)) * 1e3, 3),
                "p99_ms": round(float(np.percentile(times, 99)) * 1e3, 3)}

    allc = leg(threads, budget_s / 3)
    single = leg(1, 2 * budget_s / 3)
    out = dict(allc)
    out.update({"kind": "port",
                "sample": f"full product, {allc['runs']} timed runs after {allc['warmups']} warm-ups of the "
                          f"oracle's row-parallel Gustavson on {threads} threads (count pass, scan, sorted "
                          f"columns and values into fresh arrays, FP64), and {single['runs']} on one thread",
                "single_thread": single})
    return out


def _batched_pair(i: int):
    from gpuspgemm_amd import generators as gen
    return gen.batched_instance(i, seed=3)


def _pool_map(fn, items):
    import multiprocessing as mp
    workers = max(1, min(16, os.cpu_count() or 1))
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(fn, items, chunksize=4)


def run_batched(n_inst: int, reps: int) -> dict:
    """Batched form (512 independent products in one call of
    spg_multiply_batched) against the plain per-call loop over the same
    products: intermediate products per second of each."""
    import gpuspgemm_amd as g
    g.load_library()
    pairs = _pool_map(_batched_pair, range(n_inst))
    dev = [(a.to_device(), b.to_device()) for a, b in pairs]
    nip = sum(g.get_nip(da, db) for da, db in dev)
    As, Bs = [a for a, _ in dev], [b for _, b in dev]

    def batched():
        t0 = time.perf_counter()
        cs = g.proposal_batched(As, Bs)
        t = time.perf_counter() - t0
        for c in cs:
            c.free()
        return t

    def loop():
        t0 = time.perf_counter()
        cs = [g.proposal(a, b) for a, b in dev]
        t = time.perf_counter() - t0
        for c in cs:
            c.free()
        return t

    batched(), loop()  # warm the pools
    tb = min(batched() for _ in range(reps))
    tl = min(loop() for _ in range(reps))
    # the same batch with graph replay on: call 1 eager, call 2 captures, then replays
    g.set_graphs(True)
    try:
        batched(), batched()
        tg = min(batched() for _ in range(reps))
        gstats = g.graph_stats()
    finally:
        g.set_graphs(False)
    for a, b in dev:
        a.free()
        b.free()
    rows = int(sum(p[0].m for p in pairs))
    return {"config": {"workload": "batched", "instances": n_inst, "rows": rows, "row_length": "Geometric(1/6)",
                       "rmat_rows": "1/3", "seed": 3}, "dtype": "f64", "nip": int(nip),
            "batched_ms": round(tb * 1e3, 3), "loop_ms": round(tl * 1e3, 3),
            "batched_products_per_s": nip / tb, "loop_products_per_s": nip / tl, "speedup": tl / tb,
            "graph_replay_ms": round(tg * 1e3, 3), "graph_stats": gstats,
            "timing": f"best of {reps}, wall clock around the whole call (outputs allocated inside)"}


def _served_request(k: int) -> dict:
    from gpuspgemm_amd import generators as gen
    return gen.served_request(k)


def served_bound(dev, nips, iso, rows) -> dict:
    """What any schedule of these requests can gain over the serial plain
    path, measured: every request known at once and run in one batched pass
    per value type (best of 3 wall-clock runs), and the share of the serial
    time taken by the requests of at least 2^17 rows, which keep the GPU
    busy on their own."""
    import gpuspgemm_amd as g
    groups = {}
    for i, (a, b) in enumerate(dev):
        groups.setdefault(str(a.dtype), []).append(i)
    best = None
    for _ in range(3):
        t0 = time.perf_counter()
        for idx in groups.values():
            for c in g.proposal_batched([dev[i][0] for i in idx], [dev[i][1] for i in idx]):
                c.free()
        t = time.perf_counter() - t0
        best = t if best is None else min(best, t)
    total, serial = float(sum(nips)), float(np.sum(iso))
    big = np.array(rows) >= (1 << 17)
    # Queueing at the offered load of the serial rate even with the best
    # measured speedup (service times scaled by it): utilisation rho and
    # the Pollaczek-Khinchine mean wait of an M/G/1 queue,
    # rho / (1 - rho) * E[S^2] / (2 E[S]).
    speed = serial / best
    svc = iso / speed
    rho = 1.0 / speed
    mean_wait = rho / (1.0 - rho) * float(np.mean(svc ** 2)) / (2.0 * float(np.mean(svc))) if rho < 1 else None
    return {"batched_all_s": best, "batched_all_products_per_s": total / best,
            "batched_all_vs_serial": speed,
            "serial_share_rows_ge_2^17": float(np.sum(iso[big]) / serial),
            "requests_rows_ge_2^17": int(big.sum()),
            "mg1_rho_at_serial_rate": rho,
            "mg1_mean_wait_ms_at_serial_rate": mean_wait * 1e3 if mean_wait is not None else None,
            "isolated_p99_ms": float(np.percentile(iso, 99) * 1e3)}


def run_served(n_req: int, load: float, budget_gb: float, max_batch: int, bound: bool = False) -> dict:
    """Served form: Poisson arrivals of mixed requests (70 % FP64, 30 % FP32;
    generator drawn from uniform8 / power-law / stencil / batched instance;
    rows log-uniform in [2^10, 2^20]) at an offered load of `load` times the
    rate the serial plain path sustains; operands are device-resident.
    Latency is from submit to completion as the server records it."""
    import threading
    import gpuspgemm_amd as g
    g.load_library()
    reqs = _pool_map(_served_request, range(n_req))
    dev = [(q["A"].to_device(), q["B"].to_device()) for q in reqs]
    nips = [g.get_nip(a, b) for a, b in dev]
    # isolated (serial plain path) latency of every request
    for a, b in dev[: min(8, n_req)]:
        g.proposal(a, b).free()
    iso = []
    t0 = time.perf_counter()
    for a, b in dev:
        t1 = time.perf_counter()
        g.proposal(a, b).free()
        iso.append(time.perf_counter() - t1)
    serial_s = time.perf_counter() - t0
    rate = load * n_req / serial_s  # requests per second offered
    gaps = -np.log(1.0 - gen_unit(5, n_req)) / rate
    arrivals = np.cumsum(gaps)
    lat = [0.0] * n_req
    with g.Server(budget_bytes=int(budget_gb * (1 << 30)), max_batch=max_batch) as srv:
        tickets = [None] * n_req
        done = threading.Event()

        def waiter():
            for i in range(n_req):
                while tickets[i] is None:
                    time.sleep(0.0001)
                c, ms = srv.wait(tickets[i])
                lat[i] = ms * 1e-3
                c.free()
            done.set()

        th = threading.Thread(target=waiter)
        th.start()
        start = time.perf_counter()
        for i in range(n_req):
            now = time.perf_counter() - start
            if arrivals[i] > now:
                time.sleep(arrivals[i] - now)
            tickets[i] = srv.submit(*dev[i])
        th.join()
        makespan = time.perf_counter() - start
        st = srv.stats()
    lat = np.array(lat)
    iso = np.array(iso)
    total = float(sum(nips))
    extra = {"bound": served_bound(dev, nips, iso, [a.m for a, _ in dev])} if bound else {}
    for a, b in dev:
        a.free()
        b.free()
    return {**extra, "config": {"workload": "served", "requests": n_req, "rows": "log-uniform [2^10, 2^20]",
                       "mix": "70% f64 / 30% f32; uniform8, power-law, stencil, batched instance", "seed": 4,
                       "offered_load_req_per_s": round(rate, 2), "load_vs_serial": load,
                       "budget_gb": budget_gb, "max_batch": max_batch},
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "isolated_p50_ms": float(np.percentile(iso, 50) * 1e3),
            "isolated_p99_ms": float(np.percentile(iso, 99) * 1e3),
            "p99_vs_isolated_p99": float(np.percentile(lat, 99) / np.percentile(iso, 99)),
            "products_per_s": total / makespan, "serial_products_per_s": total / serial_s,
            "vs_serial": (total / makespan) / (total / serial_s), "server": st}


def gen_unit(stream: int, n: int) -> np.ndarray:
    from gpuspgemm_amd import generators as gen
    return gen.rand_unit(4, stream, np.arange(n))


def copy_calibration(mbytes: int, reps: int = 5) -> None:
    """Device-to-device copies of a known size, for scaling rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters to bytes (run under --pmc)."""
    import ctypes
    import gpuspgemm_amd as g
    from gpuspgemm_amd import _lib
    lib = g.load_library()
    n = mbytes << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.check(lib.spg_malloc(ctypes.byref(src), n))
    _lib.check(lib.spg_malloc(ctypes.byref(dst), n))
    for _ in range(reps):
        _lib.check(lib.spg_memcpy(dst, src, n))
    _lib.check(lib.spg_device_sync())
    lib.spg_free(src)
    lib.spg_free(dst)
    print(json.dumps({"copy_calibration_bytes": n, "reps": reps}), flush=True)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--pause", type=float, default=0.0, help="ms of sleep before each call (reference: 50)")
    ap.add_argument("--workload", default="powerlaw", choices=["powerlaw", "stencil", "uniform8"])
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32"])
    ap.add_argument("--cpu-budget", type=float, default=10.0, help="seconds of CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run device copies of MB MiB (counter calibration)")
    ap.add_argument("--form", default="all", choices=["all", "plain", "batched", "served"],
                    help="all: the plain line with the batched and served forms attached")
    ap.add_argument("--instances", type=int, default=512, help="batched form: products per batch")
    ap.add_argument("--requests", type=int, default=400, help="served form: requests")
    ap.add_argument("--load", type=float, default=1.0, help="served form: offered load / serial rate")
    ap.add_argument("--sat-load", type=float, default=4.0,
                    help="served form, throughput run: offered load / serial rate")
    ap.add_argument("--mod-load", type=float, default=0.5,
                    help="served form, latency run at a moderate load: offered load / serial rate")
    ap.add_argument("--budget-gb", type=float, default=16.0, help="served form: output memory budget")
    ap.add_argument("--max-batch", type=int, default=32, help="served form: largest batch")
    args = ap.parse_args(argv)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the library runs on one GPU; "
              "a multi-GPU run would shard the rows of A, one process per GPU)", file=sys.stderr)
        return 2
    dtype = np.float64 if args.dtype == "f64" else np.float32
    if args.form == "batched":
        print(json.dumps(run_batched(args.instances, 3)), flush=True)
        return 0
    if args.form == "served":
        print(json.dumps(run_served(args.requests, args.load, args.budget_gb, args.max_batch, bound=True)), flush=True)
        return 0

    A, cfg = make_workload(args.workload)
    res = run_gpu(A, args.steps, args.warmup, args.pause, dtype)
    out = summarize(A, res, args.dtype, args.workload, cfg)
    line = {
        "metric": METRIC, "value": round(out["value"], 3), "unit": "GFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(out["ms_per_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.dtype,
        "data": "synthetic", "config": cfg,
        "hbm_gbps": round(out["hbm_gbps"], 1), "hbm_frac": round(out["hbm_frac"], 4),
        "p50_ms": round(out["p50_ms"], 4), "p99_ms": round(out["p99_ms"], 4),
        "nip": out["nip"], "nnz_c": out["nnz_c"], "alg_bytes": out["alg_bytes"],
        "kernels_ms": out["kernels_ms"], "roofline": out["roofline"],
        "device": res["info"],
    }
    if not args.no_secondary and args.workload == "powerlaw":
        S, scfg = make_workload("stencil")
        sres = run_gpu(S, args.steps, min(args.warmup, 3), args.pause, dtype)
        sout = summarize(S, sres, args.dtype, "stencil", scfg)
        line["secondary"] = {
            "config": scfg, "value": round(sout["value"], 3), "unit": "GFLOP/s",
            "ms_per_step": round(sout["ms_per_step"], 4), "hbm_gbps": round(sout["hbm_gbps"], 1),
            "hbm_frac": round(sout["hbm_frac"], 4), "kernels_ms": sout["kernels_ms"],
            "roofline": sout["roofline"], "nip": sout["nip"], "nnz_c": sout["nnz_c"],
        }
        del S
    if not args.no_secondary and args.workload == "powerlaw" and args.dtype == "f64":
        # the power-law product in FP32 as well (same matrix, values rounded)
        fres = run_gpu(A, args.steps, min(args.warmup, 3), args.pause, np.float32)
        fout = summarize(A, fres, "f32", "powerlaw", cfg)
        line["powerlaw_f32"] = {
            "value": round(fout["value"], 3), "unit": "GFLOP/s", "dtype": "f32",
            "ms_per_step": round(fout["ms_per_step"], 4), "hbm_gbps": round(fout["hbm_gbps"], 1),
            "hbm_frac": round(fout["hbm_frac"], 4), "roofline": fout["roofline"],
        }
    if args.form == "all":
        line["batched"] = run_batched(args.instances, 3)
        line["served"] = run_served(args.requests, args.load, args.budget_gb, args.max_batch, bound=True)
        line["served_moderate"] = run_served(args.requests, args.mod_load, args.budget_gb, args.max_batch)
        line["served_saturated"] = run_served(args.requests, args.sat_load, args.budget_gb, args.max_batch)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(A.astype(dtype), out["nip"], args.cpu_budget, args.warmup, args.steps)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
