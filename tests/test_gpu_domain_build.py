"""Domain::from_mesh on the GPU (fem2d_device_domain_from_mesh) against the host construction.

The device-built Domain must hold the same arrays, bit for bit, as the view of the host
Domain: the sequential restatement of domain.rs:69-159 (from_mesh_serial) on the small cases,
the data-parallel host builder (itself checked against the sequential one on every case by
test_domain_parallel.py) at the full workload sizes. A GEP assembled from it must equal the
one assembled from the uploaded view."""

import numpy as np
import pytest

import fem_2d_amd as F
import cases
from fem_2d_amd import workloads as W

pytestmark = pytest.mark.gpu

KEYS = ("parent", "loc", "element_box", "materials", "bs_offsets", "bs_dir", "bs_i", "bs_j", "bs_dof", "dof_bs")


@pytest.fixture(scope="module")
def ctx():
    return F.Context(0)


def _assert_same_domain(dd, dom):
    ref = dom.arrays()
    assert dd.n_elems == ref["n_elems"]
    assert dd.n_dofs == ref["n_dofs"]
    got = dd.download_view()
    for k in KEYS:
        np.testing.assert_array_equal(got[k], np.asarray(ref[k]).reshape(-1), err_msg=k)


@pytest.mark.parametrize("name", list(cases.ALL))
def test_device_build_matches_sequential_construction(ctx, name):
    print(f"[domain-build] case={name}", flush=True)
    dom, _ = cases.ALL[name]()
    mesh = dom.mesh
    serial = F.Domain.from_mesh_serial(mesh)
    dd = F.DeviceDomain.from_mesh(mesh, ctx=ctx)
    _assert_same_domain(dd, serial)
    print(f"[domain-build] case={name} dofs={dd.n_dofs} specs={dd.n_bs}", flush=True)


@pytest.mark.parametrize("name", ["waveguide_64x32", "hp_irregular", "many_waveguides"])
def test_device_build_full_size(ctx, name):
    print(f"[domain-build] workload={name}", flush=True)
    w = getattr(W, name)()
    dd = F.DeviceDomain.from_mesh(w.domain.mesh, ctx=ctx)
    _assert_same_domain(dd, w.domain)
    print(f"[domain-build] workload={name} dofs={dd.n_dofs} specs={dd.n_bs}", flush=True)


@pytest.mark.parametrize("name,basis", [("small_hp_irregular", F.KOLShapeFn), ("deep_corner", F.HierPoly),
                                        ("small_many_waveguides", F.HierPoly)])
def test_device_built_domain_assembles_the_same_gep(ctx, name, basis):
    print(f"[domain-build] assemble case={name}", flush=True)
    dom, glq = cases.ALL[name]()
    up = F.DeviceDomain(dom, ctx)
    g_up = up.download(up.assemble(glq, basis))
    built = F.DeviceDomain.from_mesh(dom.mesh, ctx=ctx)
    g_built = built.download(built.assemble(glq, basis))
    np.testing.assert_array_equal(g_built.row_ptr, g_up.row_ptr)
    np.testing.assert_array_equal(g_built.col, g_up.col)
    np.testing.assert_array_equal(g_built.a_val, g_up.a_val)
    np.testing.assert_array_equal(g_built.b_val, g_up.b_val)


def test_device_build_of_a_freshly_refined_mesh(ctx):
    """A Mesh whose edge activation was never set by a host Domain construction."""
    m = F.Mesh.unit()
    m.global_h_refinement(F.HRef.T)
    m.h_refine_elems([1, 2], F.HRef.T)
    m.set_global_expansion_orders(F.Orders(3, 4))
    dd = F.DeviceDomain.from_mesh(m, ctx=ctx)
    _assert_same_domain(dd, F.Domain.from_mesh_serial(m))


def test_device_build_errors(ctx):
    # the reference panics on a non-H(curl) condition once the Mesh has Elems
    with pytest.raises(F.errors.MeshPanic, match="only H\\(Curl\\)"):
        F.DeviceDomain.from_mesh(F.Mesh.unit(), F.ContinuityCondition.HDiv, ctx=ctx)
    # a blank Mesh builds an empty Domain, which the assembly rejects like the reference
    dd = F.DeviceDomain.from_mesh(F.Mesh.blank(), ctx=ctx)
    assert dd.n_dofs == 0 and dd.n_elems == 0
    with pytest.raises(F.errors.EmptyDOFSet):
        dd.assemble((8, 8))
