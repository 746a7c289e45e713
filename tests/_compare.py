"""Parity checks between the HIP assembly and the oracle."""

from __future__ import annotations

import numpy as np

# Per-entry tolerance of the FP64 path. An entry e = (r, c) passes when
#   |gpu_e - ref_e| <= RTOL * max(|ref_e|, sqrt(|ref_rr| * |ref_cc|)).
# The floor sqrt(|d_r d_c|) bounds |e| for the positive semi-definite A and B
# (Cauchy-Schwarz), so entries that vanish up to rounding are compared against
# the size their row and column could give them.
RTOL = 1e-12


def entry_errors(rows, cols, got, ref):
    diag = np.zeros(int(max(rows.max(initial=0), cols.max(initial=0))) + 1)
    on = rows == cols
    diag[rows[on]] = np.abs(ref[on])
    floor = np.sqrt(diag[rows] * diag[cols])
    scale = np.maximum(np.abs(ref), floor)
    err = np.abs(got - ref)
    rel = np.where(scale > 0, err / np.where(scale > 0, scale, 1.0), np.where(err > 0, np.inf, 0.0))
    return rel


def assert_gep_matches(gep, ref, rtol: float = RTOL):
    """Pattern bit-exact, values within rtol (definition above)."""
    assert gep.n == ref.n, (gep.n, ref.n)
    np.testing.assert_array_equal(gep.row_ptr, ref.row_ptr)
    np.testing.assert_array_equal(gep.col, ref.col)
    rows = ref.rows()
    worst = {}
    for name, got, exp in (("A", gep.a_val, ref.a_val), ("B", gep.b_val, ref.b_val)):
        rel = entry_errors(rows, ref.col, got, exp)
        k = int(np.argmax(rel)) if rel.size else 0
        worst[name] = float(rel[k]) if rel.size else 0.0
        assert worst[name] <= rtol, (
            f"{name}[{rows[k]},{ref.col[k]}]: got {got[k]!r}, oracle {exp[k]!r}, rel {rel[k]:.3e}")
    return worst


def assert_rows_match(gep, ref, mask, rtol: float = RTOL):
    """The rows selected by mask: pattern bit-exact, values within rtol. ref holds only those
    rows (the oracle's row-mask mode); the scale floor uses the GPU diagonal for columns
    outside the mask (those diagonals are checked wherever their row is selected)."""
    rows = gep.rows()
    sel = mask[rows]
    r, c = rows[sel], gep.col[sel]
    rr = ref.rows()
    np.testing.assert_array_equal(r, rr)
    np.testing.assert_array_equal(c, ref.col)
    worst = {}
    for name, allv, exp in (("A", gep.a_val, ref.a_val), ("B", gep.b_val, ref.b_val)):
        diag = np.zeros(gep.n)
        on = rows == gep.col
        diag[rows[on]] = np.abs(allv[on])
        scale = np.maximum(np.abs(exp), np.sqrt(diag[r] * diag[c]))
        got = allv[sel]
        err = np.abs(got - exp)
        rel = np.where(scale > 0, err / np.where(scale > 0, scale, 1.0), np.where(err > 0, np.inf, 0.0))
        k = int(np.argmax(rel)) if rel.size else 0
        worst[name] = float(rel[k]) if rel.size else 0.0
        assert worst[name] <= rtol, (
            f"{name}[{r[k]},{c[k]}]: got {got[k]!r}, oracle {exp[k]!r}, rel {rel[k]:.3e}")
    return worst
