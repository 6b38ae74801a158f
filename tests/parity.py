"""Parity helpers: run the GPU path through the C ABI and compare with the
oracle (structure bit-exact, values within the stated tolerance)."""
from __future__ import annotations

import numpy as np

import gpuspgemm_amd as g
import oracle

# Tolerances of the FP64 oracle comparison, applied to the |A||B|-scaled
# error |c - c*| / (|A||B|)_ij: FP64 1e-12, FP32 1e-5.
TOL = {np.dtype(np.float64): 1e-12, np.dtype(np.float32): 1e-5}


def gpu_multiply(A: g.HostCSR, B: g.HostCSR) -> g.HostCSR:
    dA, dB = A.to_device(), B.to_device()
    try:
        dC = g.proposal(dA, dB)
        try:
            return dC.to_host()
        finally:
            dC.free()
    finally:
        dA.free()
        dB.free()


def compare(A: g.HostCSR, B: g.HostCSR, C: g.HostCSR, ref=None) -> dict:
    """Assert parity of C with the oracle product of A and B; return error stats."""
    if ref is None:
        ref = oracle.spgemm(A, B)
    assert (C.m, C.n) == (A.m, B.n)
    assert C.nnz == ref.nnz, f"nnz {C.nnz} != oracle {ref.nnz}"
    np.testing.assert_array_equal(C.rpt.astype(np.int64), ref.rpt)
    np.testing.assert_array_equal(C.col, ref.col)
    c = C.val.astype(np.float64)
    diff = np.abs(c - ref.val)
    scale = np.where(ref.absval > 0, ref.absval, 1.0)
    scaled = float((diff / scale).max()) if diff.size else 0.0
    nz = np.abs(ref.val) > 0
    elem = float((diff[nz] / np.abs(ref.val[nz])).max()) if nz.any() else 0.0
    tol = TOL[np.dtype(C.val.dtype)]
    assert scaled <= tol, f"max |A||B|-scaled error {scaled:.3e} > {tol:g}"
    # the reference's own acceptance rule (csr.cuh:232-238)
    ok = (c == ref.val) | (diff <= 1e-8 + 1e-5 * np.maximum(np.abs(c), np.abs(ref.val)))
    return {"nnz": C.nnz, "max_scaled_err": scaled, "max_elem_rel_err": elem,
            "reference_rule_ok": bool(ok.all())}
