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


@pytest.mark.parametrize("graphs", [True, False])
def test_block_path_after_smaller_domain(graphs):
    """Regression: a Domain with block-path row groups assembled on a context whose output
    buffers were sized for a smaller Domain. The first attempt finds them too small; its
    block-path fill must not read row pointers that only the streaming fill writes (and that
    fill skips that attempt). The result equals the oracle's, also when the buffers grow twice."""
    print(f"[regress] block path after a smaller Domain, graphs={graphs}", flush=True)
    c = F.Context(0)
    c.set_graphs(graphs)
    small, sglq = cases.nalg_problem()
    F.galerkin_sample_gep_hcurl(small, sglq, ctx=c)
    mid, mglq = cases.slepc_problem()
    F.galerkin_sample_gep_hcurl(mid, mglq, F.KOLShapeFn, ctx=c)
    dom, glq = cases.small_hp_irregular()
    gep = F.galerkin_sample_gep_hcurl(dom, glq, F.KOLShapeFn, ctx=c)
    assert c.last_timing()["n_large"] > 0, "the case must exercise the block path"
    rc, ref = oracle.assemble(dom.view(), basis=1, glq=glq, n_threads=8)
    assert rc == 0
    assert_gep_matches(gep, ref)
    # and back to the small Domain on the grown buffers
    g_small = F.galerkin_sample_gep_hcurl(small, sglq, ctx=c)
    rc, ref_small = oracle.assemble(small.view(), basis=0, glq=sglq)
    assert rc == 0
    assert_gep_matches(g_small, ref_small)
    c.close()


FILL_CHUNKS = [64, 128, 1024, 4096]


@pytest.mark.parametrize("name,basis", [("small_hp_irregular", "kol"), ("waveguide_small", "hierpoly"),
                                        ("deep_corner", "kol")])
def test_fill_chunk_sizes_give_identical_gep(name, basis):
    """Every chunk size of the streaming fill (64 to 4096 entries per work item) writes the same
    GEP, bit for bit, as the default size; one of them is checked against the oracle too."""
    print(f"[fill-chunk] case={name}", flush=True)
    c = F.Context(0)
    c.set_autotune(False)
    dom, glq = cases.ALL[name]()
    B = BASIS[basis]
    dd = F.DeviceDomain(dom, c)
    ref = dd.download(dd.assemble(glq, B))
    for chunk in FILL_CHUNKS:
        c.set_fill_chunk(chunk)
        got = dd.download(dd.assemble(glq, B))
        assert c.last_timing()["fill_chunk"] == chunk
        np.testing.assert_array_equal(got.row_ptr, ref.row_ptr)
        np.testing.assert_array_equal(got.col, ref.col)
        assert np.array_equal(got.a_val.view(np.uint64), ref.a_val.view(np.uint64)), chunk
        assert np.array_equal(got.b_val.view(np.uint64), ref.b_val.view(np.uint64)), chunk
    c.set_fill_chunk(128)
    gep = dd.download(dd.assemble(glq, B))
    rc, oref = oracle.assemble(dom.view(), basis=B.code, glq=glq, n_threads=8)
    assert rc == 0
    assert_gep_matches(gep, oref)
    with pytest.raises(ValueError):  # FEM2D_ERR_INVALID_ARG
        c.set_fill_chunk(96)
    c.close()


def test_stale_device_handle_rejected(ctx):
    """A device GEP handle is valid until the next assembly on its context: afterwards the
    download, the PETSc writer and the eigensolves reject it (the buffers may have been
    reused or freed), while the latest handle works."""
    dom, glq = cases.slepc_problem()
    other, oglq = cases.small_hp_irregular()
    dd = F.DeviceDomain(dom, ctx)
    od = F.DeviceDomain(other, ctx)
    h1 = dd.assemble(glq)
    assert h1.generation != 0
    h2 = od.assemble(oglq, F.KOLShapeFn)  # grows the buffers: h1's memory is gone
    assert h2.generation != h1.generation
    with pytest.raises(ValueError, match="stale"):
        dd.download(h1)
    with pytest.raises(ValueError, match="stale"):
        F.solve_gep_dense(h1, 1.47, ctx=ctx)
    with pytest.raises(ValueError, match="stale"):
        F.nalgebra_solve_gep(h1, 1.47, ctx=ctx)
    with pytest.raises(ValueError, match="stale"):
        dd.print_to_petsc_binary_files(h1, "/nonexistent-dir", "x")
    # a handle from another context is stale on this one
    c2 = F.Context(0)
    h3 = F.DeviceDomain(dom, c2).assemble(glq)
    with pytest.raises(ValueError, match="stale"):
        dd.download(h3)
    c2.close()
    # the latest handle of this context is good
    g = od.download(h2)
    assert g.nnz == h2.nnz


def _same_bits(g, ref):
    np.testing.assert_array_equal(g.row_ptr, ref.row_ptr)
    np.testing.assert_array_equal(g.col, ref.col)
    assert np.array_equal(g.a_val.view(np.uint64), ref.a_val.view(np.uint64))
    assert np.array_equal(g.b_val.view(np.uint64), ref.b_val.view(np.uint64))


def test_graph_replays_match_eager(ctx):
    """Assemblies recorded as a HIP graph and replayed give, on every replay, the bits of the
    kernel-by-kernel launch; eager work on another context and re-recordings for another
    Domain run between the replays."""
    dom, glq = cases.small_hp_irregular()
    other, oglq = cases.mixed_refinements()
    eager = F.Context(0)
    eager.set_graphs(False)
    ref = F.galerkin_sample_gep_hcurl(dom, glq, F.KOLShapeFn, ctx=eager)
    dd = F.DeviceDomain(dom, ctx)
    od = F.DeviceDomain(other, ctx)
    for i in range(4):
        _same_bits(dd.download(dd.assemble(glq, F.KOLShapeFn)), ref)
        F.galerkin_sample_gep_hcurl(other, oglq, ctx=eager)
        eager.synchronize()
        if i == 2:
            od.assemble(oglq)  # a different graph is recorded in between
    _same_bits(dd.download(dd.assemble(glq, F.KOLShapeFn)), ref)
    eager.close()


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


def _view_from_arrays(a):
    """A fem2d_domain_view over (possibly corrupted) copies of Domain.arrays()."""
    view = F._native.DomainView()
    keep = {}
    for name, ctype in [("parent", C.c_int32), ("loc", C.c_uint8), ("element_box", C.c_double),
                        ("materials", C.c_double), ("orders", C.c_uint8), ("bs_offsets", C.c_int64),
                        ("bs_dir", C.c_uint8), ("bs_i", C.c_uint8), ("bs_j", C.c_uint8), ("bs_dof", C.c_int32),
                        ("dof_bs", C.c_int32)]:
        keep[name] = np.ascontiguousarray(a[name].ravel())
        setattr(view, name, keep[name].ctypes.data_as(C.POINTER(ctype)))
    view.cc = 0
    view.n_elems = a["n_elems"]
    view.n_dofs = a["n_dofs"]
    return view, keep


def _corrupt(kind, a):
    a = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in a.items()}
    if kind == "dof_out_of_range":
        a["bs_dof"][3] = a["n_dofs"] + 5
    elif kind == "w_direction":
        a["bs_dir"][5] = 2
    elif kind == "order_beyond_tables":
        a["bs_i"][5] = 30
    elif kind == "three_specs_one_dof":
        a["bs_dof"][1] = a["bs_dof"][0]
        a["bs_dof"][2] = a["bs_dof"][0]
    elif kind == "bad_href_loc":
        child = int(np.flatnonzero(a["parent"] >= 0)[0])
        a["loc"][child] = 9
    elif kind == "parent_cycle":
        a["parent"][0] = 1
        a["parent"][1] = 0
    elif kind == "split_group":
        # swap the DoFs of two specs of one Elem that belong to different row groups
        off = a["bs_offsets"]
        e = int(np.argmax(np.diff(off)))
        lo, hi = int(off[e]), int(off[e + 1]) - 1
        a["bs_dof"][lo], a["bs_dof"][hi] = a["bs_dof"][hi], a["bs_dof"][lo]
    else:
        raise ValueError(kind)
    return a


@pytest.mark.parametrize("kind", ["dof_out_of_range", "w_direction", "order_beyond_tables", "three_specs_one_dof",
                                  "bad_href_loc", "parent_cycle", "split_group"])
def test_invalid_view_rejected(ctx, kind):
    """Views that violate the Domain invariants are rejected with a status code (device checks
    or host checks), never assembled and never allowed to address memory outside the inputs."""
    print(f"[invalid-view] kind={kind}", flush=True)
    dom, glq = cases.mixed_refinements()
    view, keep = _view_from_arrays(_corrupt(kind, dom.arrays()))
    out = F._native.GepCsr()
    rc = F._native.load().fem2d_galerkin_sample_gep_hcurl(ctx._h, C.byref(view), 0, 0, 1, 8, 8, 0, C.byref(out))
    assert rc in (F._native.ERR_INVALID_VIEW, F._native.ERR_UNSUPPORTED), (rc, F._native.last_error())
    # the context stays usable
    gep = F.galerkin_sample_gep_hcurl(dom, glq, F.HierPoly, F.CurlCurl, F.L2Inner, ctx=ctx)
    assert gep.nnz > 0
