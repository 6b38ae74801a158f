"""The autotuner on the CPU: its search and its table emission run against
the library's real table checks (spg_set_tuning / spg_get_tuning of the
library sources built for the CPU emulation in tests/emu), with the timing
of each multiply replaced by a deterministic cost of the current tables.
The emitted include must parse back to the tuned tables, and the library
must accept them; tables that break the rules (a NIP cap at or above the
variant's table size, an nnz cap above the capacity, a last bin that is not
the largest-table / global variant with caps INT_MAX) are refused."""
import ctypes
import os
import re
import subprocess

import pytest

import gpuspgemm_amd as g
from gpuspgemm_amd import _lib, tune

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")
CSRC = os.path.join(ROOT, "gpuspgemm_amd", "csrc")
INT_MAX = 2**31 - 1


@pytest.fixture(scope="module")
def emu_lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("emu") / "libspg_emu.so")
    srcs = [os.path.join(CSRC, f) for f in ("k_rows.hip", "k_sym.hip", "k_symnum.hip", "api.hip")]
    cmd = ["g++", "-std=c++17", "-O0", "-shared", "-fPIC", "-pthread", "-w", "-I", os.path.join(EMU, "include"),
           "-I", os.path.join(ROOT, "include"), "-x", "c++", *srcs, "-x", "none",
           os.path.join(EMU, "emu_runtime.cpp"), "-o", out]
    subprocess.run(cmd, check=True)
    lib = ctypes.CDLL(out)
    for name, (res, args) in _lib._SIG.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


@pytest.fixture
def on_emu(emu_lib, monkeypatch):
    monkeypatch.setattr(_lib, "_lib", emu_lib)
    yield emu_lib
    emu_lib.spg_set_tuning(0, None)
    emu_lib.spg_set_tuning(1, None)


class CostWorkload:
    """Stands in for a timed workload: the 'time' of a multiply is a fixed
    function of the current tables that favours some variants and caps."""

    def __init__(self, name, prefer_sym, prefer_num):
        self.name = name
        self.prefer_sym, self.prefer_num = prefer_sym, prefer_num

    def time(self, reps, warmup=2):
        sym = tune.parse_table(g.get_tuning(0))
        num = tune.parse_table(g.get_tuning(1))
        cost = 1.0
        cost += 0.05 * sum(v != self.prefer_sym.get(i, v) for i, (v, _, _) in enumerate(sym))
        cost += 0.05 * sum(v != self.prefer_num.get(i, v) for i, (v, _, _) in enumerate(num))
        return cost


def test_tuner_emits_tables_the_library_accepts(on_emu, tmp_path):
    start_sym = tune.parse_table(g.get_tuning(0))
    start_num = tune.parse_table(g.get_tuning(1))
    sv, nv = g.variant_list(0), g.variant_list(1)
    # prefer, for the first multi bin of each phase, another candidate variant
    want_sym = next(v for v in tune.sym_candidates(start_sym[0], sv) if v != start_sym[0][0])
    want_num = next(v for v in tune.num_candidates(start_num[0], nv) if v != start_num[0][0])
    ws = [CostWorkload("a", {0: want_sym}, {0: want_num}), CostWorkload("b", {0: want_sym}, {})]
    res = tune.tune(ws, reps=1, passes=1, log=lambda s: None)
    assert res["sym"][0][0] == want_sym and res["num"][0][0] == want_num
    assert res["score"] < 2.0
    # the include parses back to the same tables
    inc = tmp_path / "tuned_tables.inc"
    tune.write_include(str(inc), res)
    text = inc.read_text()

    def rows(name):
        body = re.search(name + r"\[\] = \{(.*)\};", text).group(1)
        return [tuple(INT_MAX if x.strip() == "INT_MAX" else int(x) for x in m.split(","))
                for m in re.findall(r"\{([^{}]*)\}", body)]

    assert rows("kTunedSym") == [tuple(r) for r in res["sym"]]
    assert rows("kTunedNum") == [tuple(r) for r in res["num"]]
    # and the library takes them
    g.set_tuning(0, tune.table_text(res["sym"]))
    g.set_tuning(1, tune.table_text(res["num"]))
    assert tune.parse_table(g.get_tuning(0)) == [tuple(r) for r in res["sym"]]
    assert res["sym"][-1][1:] == (INT_MAX, INT_MAX) and res["num"][-1][1:] == (INT_MAX, INT_MAX)


def test_library_refuses_tables_that_break_the_rules(on_emu):
    sv, nv = g.variant_list(0), g.variant_list(1)
    sym = tune.parse_table(g.get_tuning(0))
    num = tune.parse_table(g.get_tuning(1))
    v0 = next(v for v in sv if v["id"] == sym[0][0])
    bad = [
        (0, [(sym[0][0], v0["table"], INT_MAX)] + sym[1:]),             # NIP cap = table size
        (0, sym[:-1] + [(sym[0][0], INT_MAX, INT_MAX)]),                # last bin not the max variant
        (0, sym[:-1] + [(sym[-1][0], 4096, INT_MAX)]),                  # last cap not INT_MAX
        (1, [(num[0][0], num[0][1], next(v for v in nv if v["id"] == num[0][0])["capacity"] + 1)] + num[1:]),
        (1, num[:-1] + [(num[0][0], INT_MAX, INT_MAX)]),                # last bin not the global variant
    ]
    for phase, table in bad:
        with pytest.raises(g.SpgError, match="INVALID_ARGUMENT"):
            g.set_tuning(phase, tune.table_text(table))
    # the defaults are still in place
    assert tune.parse_table(g.get_tuning(0)) == sym
