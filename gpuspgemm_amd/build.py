"""Build the gfx950 shared library and the host apps in-tree.

    python -m gpuspgemm_amd.build            # library + apps
    python -m gpuspgemm_amd.build --lib      # library only

The library is `gpuspgemm_amd/libgpuspgemm.so` (HIP kernels + C ABI).  The
apps (`gpuspgemm_amd/bin/proposal`, `proposal_f32`) are plain C++ compiled
with g++ against include/gpuspgemm.hpp and linked to the library; they do
not include any HIP header.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libgpuspgemm.so")
BIN = os.path.join(PKG, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
    # packed FP32 instructions stay off in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]
SOURCES = ["k_rows.hip", "k_sym.hip", "k_symnum.hip", "api.hip"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines()
                    if "packed-fp32-ops" not in l or "error" in l)
    if res.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}")


def build_lib() -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 4)) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_apps() -> list[str]:
    """bin/proposal(_f32): the library's app (plain C++, no HIP headers);
    bin/rocsparse(_f32): the vendor-library comparator (host code against
    the HIP runtime and rocSPARSE)."""
    os.makedirs(BIN, exist_ok=True)
    outs = []
    cmds = []
    for name, defs in (("proposal", ["-DUSE_DOUBLE_PRECISION"]), ("proposal_f32", [])):
        out = os.path.join(BIN, name)
        cmds.append(["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", *defs,
                     "-I", os.path.join(ROOT, "include"),
                     os.path.join(PKG, "apps", "proposal.cpp"), "-o", out,
                     "-L", PKG, "-lgpuspgemm", "-Wl,-rpath,$ORIGIN/..", "-pthread"])
        outs.append(out)
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(HIPCC)))
    for name, defs in (("rocsparse", ["-DUSE_DOUBLE_PRECISION"]), ("rocsparse_f32", [])):
        out = os.path.join(BIN, name)
        cmds.append([HIPCC, f"--offload-arch={ARCH}", "-std=c++20", "-O2", *defs,
                     "-I", os.path.join(ROOT, "include"),
                     os.path.join(PKG, "apps", "rocsparse.cpp"), "-o", out,
                     "-L", PKG, "-lgpuspgemm", "-L", os.path.join(rocm, "lib"), "-lrocsparse",
                     "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + os.path.join(rocm, "lib"), "-pthread"])
        outs.append(out)
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        list(ex.map(_run, cmds))
    return outs


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    args = ap.parse_args(argv)
    print(build_lib())
    if not args.lib:
        for p in build_apps():
            print(p)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
