"""HIP assembly vs the oracle (CPU restatement of galerkin.rs) through the C ABI.

Pattern (row_ptr, col) must match bit for bit; values within the tolerance of
_compare.RTOL (1e-12 relative, with the sqrt(|d_r d_c|) floor)."""

import ctypes as C

import numpy as np
import pytest

import fem_2d_amd as F
import oracle
from _compare import assert_gep_matches
import cases

pytestmark = pytest.mark.gpu

BASIS = {"hierpoly": F.HierPoly, "kol": F.KOLShapeFn}


@pytest.fixture(scope="module")
def ctx():
    return F.Context(0)


@pytest.mark.parametrize("name", list(cases.ALL))
@pytest.mark.parametrize("basis", ["hierpoly", "kol"])
def test_parity(ctx, name, basis):
    print(f"[parity] case={name} basis={basis}", flush=True)
    dom, glq = cases.ALL[name]()
    B = BASIS[basis]
    gep = F.galerkin_sample_gep_hcurl(dom, glq, B, F.CurlCurl, F.L2Inner, ctx=ctx)
    rc, ref = oracle.assemble(dom.view(), basis=B.code, glq=glq, n_threads=8)
    assert rc == 0
    worst = assert_gep_matches(gep, ref)
    print(f"[parity] case={name} basis={basis} nnz={gep.nnz} worst={worst}", flush=True)


def test_swapped_integrals_and_default_glq(ctx):
    """AI = L2Inner, BI = CurlCurl, glq_grid_dim = None (default_ngq, basis.rs:172-177)."""
    dom, _ = cases.mixed_refinements()
    gep = F.galerkin_sample_gep_hcurl(dom, None, F.HierPoly, F.L2Inner, F.CurlCurl, ctx=ctx)
    rc, ref = oracle.assemble(dom.view(), basis=0, a_integral=1, b_integral=0, glq=None)
    assert rc == 0
    assert_gep_matches(gep, ref)


def test_same_integral_twice(ctx):
    dom, glq = cases.slepc_problem()
    gep = F.galerkin_sample_gep_hcurl(dom, glq, F.HierPoly, F.CurlCurl, F.CurlCurl, ctx=ctx)
    np.testing.assert_array_equal(gep.a_val, gep.b_val)


def test_deterministic_and_device_resident(ctx):
    """Two assemblies are bitwise identical; the device-resident API gives the same GEP."""
    dom, glq = cases.small_hp_irregular()
    g1 = F.galerkin_sample_gep_hcurl(dom, glq, F.KOLShapeFn, ctx=ctx)
    dd = F.DeviceDomain(dom, ctx)
    g2 = dd.download(dd.assemble(glq, F.KOLShapeFn))
    g3 = dd.download(dd.assemble(glq, F.KOLShapeFn))
    for g in (g2, g3):
        np.testing.assert_array_equal(g1.row_ptr, g.row_ptr)
        np.testing.assert_array_equal(g1.col, g.col)
        assert np.array_equal(g1.a_val.view(np.uint64), g.a_val.view(np.uint64))
        assert np.array_equal(g1.b_val.view(np.uint64), g.b_val.view(np.uint64))


def test_errors_in_reference_order(ctx):
    """galerkin.rs:41-59: continuity condition, then empty DoF set, then GLQ settings."""
    with pytest.raises(F.WrongContinuityCondition):
        F.galerkin_sample_gep_hcurl(F.Domain.blank(F.ContinuityCondition.HDiv), (3, 3), ctx=ctx)
    with pytest.raises(F.EmptyDOFSet):
        F.galerkin_sample_gep_hcurl(F.Domain.blank(F.ContinuityCondition.HCurl), (3, 3), ctx=ctx)
    dom, _ = cases.nalg_problem()
    with pytest.raises(F.InvalidGLQSettings):
        F.galerkin_sample_gep_hcurl(dom, (3, 8), ctx=ctx)
    with pytest.raises(F.InvalidGLQSettings):
        F.galerkin_sample_gep_hcurl(dom, (8, 3), ctx=ctx)


def test_invalid_view_rejected(ctx):
    """A view whose DoF ids are out of range is rejected by the device checks, not assembled."""
    dom, glq = cases.nalg_problem()
    a = dom.arrays()
    bad_dof = a["bs_dof"].copy()
    bad_dof[3] = dom.num_dofs + 5
    keep = dict(a)
    v = F.Domain.__new__(F.Domain)
    view = type(dom.view())()
    arrs = {}
    for name, ctype in [("parent", C.c_int32), ("loc", C.c_uint8), ("element_box", C.c_double),
                        ("materials", C.c_double), ("orders", C.c_uint8), ("bs_offsets", C.c_int64),
                        ("bs_dir", C.c_uint8), ("bs_i", C.c_uint8), ("bs_j", C.c_uint8)]:
        arrs[name] = np.ascontiguousarray(keep[name].ravel())
        setattr(view, name, arrs[name].ctypes.data_as(C.POINTER(ctype)))
    arrs["bs_dof"] = np.ascontiguousarray(bad_dof)
    view.bs_dof = arrs["bs_dof"].ctypes.data_as(C.POINTER(C.c_int32))
    view.cc = 0
    view.n_elems = a["n_elems"]
    view.n_dofs = a["n_dofs"]
    out = F._native.GepCsr()
    rc = F._native.load().fem2d_galerkin_sample_gep_hcurl(ctx._h, C.byref(view), 0, 0, 1, 8, 8, 0, C.byref(out))
    assert rc == F._native.ERR_INVALID_VIEW
