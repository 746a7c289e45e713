"""GPU field evaluation (fem2d_xy_fields / fem2d_device_xy_fields, UniformFieldSpace::xy_fields,
fields.rs:63-127) against the oracle's restatement on the same inputs: every field value within
FIELD_TOL of the oracle's relative to the largest field magnitude (the kernel forms each product
and sum in the reference's order, so most values agree bit for bit); the device-resident variant
equal to the host one; the reference's error variants; and the TE10 field from a GPU assembly and
GPU eigensolve."""

import numpy as np
import pytest

import cases
import fem_2d_amd as F
import oracle
from fem_2d_amd import workloads as W
from test_fields_oracle import _leaf_boxes, tangential_jump

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-13


@pytest.fixture(scope="module")
def ctx():
    return F.Context(0)


def _compare(got, exp):
    scale = max(np.abs(exp).max(initial=0.0), 1e-300)
    err = np.abs(got - exp).max(initial=0.0) / scale
    return err, float(np.mean(got == exp)) if got.size else 1.0


CASES = [("nalg_problem", 0, 6), ("slepc_problem", 0, 6), ("mixed_refinements", 0, 9), ("mixed_refinements", 1, 6),
         ("deep_unit", 0, 7), ("deep_unit", 1, 7), ("waveguide_small", 0, 8), ("small_hp_irregular", 1, 5),
         ("small_many_waveguides", 0, 4)]


@pytest.mark.parametrize("name,basis,d", CASES)
def test_fields_match_oracle(ctx, name, basis, d):
    print(f"[fields] case={name} basis={basis} d={d}", flush=True)
    dom, _ = cases.ALL[name]()
    sol = np.random.default_rng(11).standard_normal(dom.num_dofs)
    rc, ex, ey = oracle.xy_fields(dom.view(), basis, (d, d), sol)
    assert rc == 0
    BS = F.KOLShapeFn if basis else F.HierPoly
    gx, gy = F.xy_fields_array(dom, sol, (d, d), BS, ctx=ctx)
    assert gx.shape == ex.shape
    for got, exp in ((gx, ex), (gy, ey)):
        err, exact = _compare(got, exp)
        print(f"[fields] case={name} max rel err {err:.2e}, bit-exact fraction {exact:.3f}", flush=True)
        assert err <= FIELD_TOL
    worst, _ = tangential_jump(dom, gx, gy, d)
    assert worst <= 1e-10 * max(np.abs(gx).max(), np.abs(gy).max())


def test_device_variant_equals_host(ctx):
    print("[fields] device variant", flush=True)
    dom, _ = cases.mixed_refinements()
    sol = np.random.default_rng(2).standard_normal(dom.num_dofs)
    hx, hy = F.xy_fields_array(dom, sol, (6, 6), ctx=ctx)
    dd = F.DeviceDomain(dom, ctx)
    dx, dy = F.xy_fields_array(dd, sol, (6, 6), ctx=ctx)
    np.testing.assert_array_equal(dx, hx)
    np.testing.assert_array_equal(dy, hy)


def test_full_size_fields(ctx):
    """The 445 380-DoF benchmark Domain (6 000+ leaves, two refinement levels), 8x8 per leaf."""
    print("[fields] waveguide_64x32", flush=True)
    w = W.waveguide_64x32()
    sol = np.random.default_rng(5).standard_normal(w.domain.num_dofs)
    rc, ex, ey = oracle.xy_fields(w.domain.view(), 0, (8, 8), sol)
    assert rc == 0
    dd = F.DeviceDomain(w.domain, ctx)
    gx, gy = F.xy_fields_array(dd, sol, (8, 8), ctx=ctx)
    for got, exp in ((gx, ex), (gy, ey)):
        assert _compare(got, exp)[0] <= FIELD_TOL


def test_field_space_api_and_errors(ctx):
    print("[fields] api and errors", flush=True)
    dom = F.Domain.unit(F.ContinuityCondition.HCurl)
    ufs = F.UniformFieldSpace(dom, [10, 10], ctx=ctx)
    names = ufs.xy_fields("unit_fields", [1.0] * dom.num_dofs)  # fields.rs:46-62
    assert names == ["unit_fields_x", "unit_fields_y"]
    dom, _ = cases.nalg_problem()
    sol = np.ones(dom.num_dofs)
    ufs = F.UniformFieldSpace(dom, [4, 4], ctx=ctx)
    ex, ey = ufs.xy_fields("E", sol)
    ufs.expression_2arg([ex, ey], "E_mag", lambda a, b: (a ** 2 + b ** 2) ** 0.5)
    for e, v in ufs.quantities["E_mag"].items():
        np.testing.assert_allclose(v, np.hypot(ufs.quantities[ex][e], ufs.quantities[ey][e]), rtol=1e-15)
    with pytest.raises(F.MissingQuantity):
        ufs.map_to_quantity("H_x", "H2", lambda a: a * a)
    with pytest.raises(F.MismatchedSolutionSize):
        ufs.xy_fields("E", sol[:-1])
    with pytest.raises(F.MeshPanic):
        F.UniformFieldSpace(dom, [4, 5], ctx=ctx).xy_fields("E", sol)
    gx, gy = F.xy_fields_array(dom, sol, (1, 1), ctx=ctx)  # uniform_range(-1, 1, 1) is [NaN]
    assert np.all(np.isnan(gx)) and np.all(np.isnan(gy))
    gx, _ = F.xy_fields_array(dom, sol, (0, 0), ctx=ctx)
    assert gx.size == 0


def test_te10_field_end_to_end(ctx):
    """GPU assembly -> GPU dense eigensolve -> GPU fields: E_x = 0, E_y proportional to sin(pi x)."""
    print("[fields] TE10 end to end", flush=True)
    w = W.waveguide_small()
    dd = F.DeviceDomain(w.domain, ctx)
    pair = F.solve_gep_dense(dd.assemble(w.glq), 10.0, ctx=ctx)
    assert abs(pair.value - np.pi ** 2) / np.pi ** 2 < 1e-6
    d = 5
    x, y = F.xy_fields_array(dd, pair.vector, (d, d), ctx=ctx)
    _, boxes = _leaf_boxes(w.domain)
    t = np.array([i * (2.0 / (d - 1)) - 1.0 for i in range(d)])
    shape = np.stack([np.outer(np.sin(np.pi * (b[0] + (t + 1) / 2 * (b[2] - b[0]))), np.ones(d)) for b in boxes])
    scale = y.ravel() @ shape.ravel() / (shape.ravel() @ shape.ravel())
    assert np.abs(x).max() < 1e-4 * abs(scale)
    assert np.abs(y - scale * shape).max() < 1e-4 * abs(scale)
