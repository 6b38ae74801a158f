"""The per-step timeline tool (gpuspgemm_amd/timeline.py), which makes the
timelines in profiles/, on a hand-made kernel trace whose busy time, idle
gaps and phase times are worked out by hand."""
import csv

from gpuspgemm_amd import timeline

# (name as rocprofv3 writes it, queue, start ns, end ns) of one multiply
STEP = [
    ("spg::k_clear(spg::DevMeta*, int)", 1, 0, 1000),
    ("void spg::k_nip<double>(spg::ProbSet<double>, ...)", 1, 1000, 5000),
    ("void spg::k_sym_multi<double, 0, 4>(...)", 1, 6000, 9000),
    ("void spg::k_sym_multi<double, 4, 7>(...)", 2, 6500, 8000),
    ("spg::k_total_nip(int const*, ...)", 3, 7000, 30000),  # another call's get_nip: left out
    ("void spg::k_scan_count<double>(...)", 1, 10000, 11000),
    ("spg::k_readback(int const*, ...)", 1, 11000, 12000),
    ("void spg::k_symnum_bin<double, 9>(...)", 1, 14000, 20000),
    ("void spg::k_symnum_multi<double, 0>(...)", 2, 15000, 22000),
    ("spg::k_readback(int const*, ...)", 1, 20000, 21000),
    ("spg::k_readback(int const*, ...)", 2, 22000, 23000),
]


def _trace(tmp_path, shifts):
    path = tmp_path / "kernel_trace.csv"
    with open(path, "w", newline="") as fp:
        w = csv.writer(fp)
        w.writerow(["Kernel_Name", "Queue_Id", "Start_Timestamp", "End_Timestamp"])
        for sh in shifts:
            for name, q, a, b in STEP:
                w.writerow([name, q, a + sh, b + sh])
    return str(path)


def test_step_summary(tmp_path):
    steps = timeline.split_multiplies(timeline.load(_trace(tmp_path, [0, 100_000])))
    assert len(steps) == 2
    s = timeline.summarize(steps[0])
    assert [k["name"] for k in steps[0]][:3] == ["clear", "nip", "sym_multi_small"]
    assert all(k["name"] != "total_nip" for k in steps[0])
    assert s["span_us"] == 23.0
    # busy: [0, 5] + [6, 9] + [10, 12] + [14, 23] us
    assert s["busy_us"] == 19.0
    assert [(a, l, b) for a, l, b in s["gaps"]] == [(5.0, 1.0, "sym_multi_small"), (9.0, 1.0, "scan_count"),
                                                    (12.0, 2.0, "symnum_v9")]
    assert s["symbolic_us"] == 12.0 and s["readback_to_fused_us"] == 2.0 and s["fused_us"] == 9.0
    assert s["queues"] == {1: (17.0, 7), 2: (9.5, 3)}


def test_report_skips_the_final_steps(tmp_path):
    out = timeline.report(_trace(tmp_path, [0, 100_000, 200_000]), last=2, skip=1)
    assert out.count("multiply ") == 2 and "mean of 2: span 23.0 us, busy 19.0 us (83 %)" in out
