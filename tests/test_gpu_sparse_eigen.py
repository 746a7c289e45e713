"""The GPU sparse shift-invert eigensolve (fem2d_solve_gep_sparse / fem2d_gep_device_solve_sparse,
the stand-in for slepc_solve_gep, slepc_solve.rs:14-68) against

* the reference's SLEPc eigenpair of lib.rs:82-119 (test_eval.dat / test_evec.dat);
* the oracle's sparse shift-invert restatement (ARPACK around the target) on GEPs of both
  shape-function spaces up to 19 296 DoFs, including a disconnected one (independent guides);
* the GPU dense solve on the 4 496-DoF precision-study matrices;
* the analytic TE10 cutoff pi^2 at full size: the 445 380-DoF benchmark GEP, solved from the
  device-resident CSR, its residual checked on the host;
* mixed-precision storage (FP32R) as input.

Eigenvalues within EIG_REL of the oracle's; vectors B-normalised, with relative residual
|A x - lambda B x| / |A x| below RESID on the host matrices, and parallel to the oracle's where
the eigenvalue is simple."""

import struct

import numpy as np
import pytest

import cases
import fem_2d_amd as F
from conftest import golden
from fem_2d_amd import workloads as W
from oracle import eigen as OE

pytestmark = pytest.mark.gpu

EIG_REL = 1e-10
RESID = 1e-9


@pytest.fixture(scope="module")
def ctx():
    return F.Context(0)


def _sparse(gep, target, ctx):
    """F.solve_gep_sparse, printing the solver's statistics (also when it raises)."""
    try:
        return F.solve_gep_sparse(gep, target, ctx=ctx)
    finally:
        st = ctx.last_eigen_stats()
        print(f"[sparse-eigen] stats {st}", flush=True)


def _check_pair(A, B, pair, lam_ref=None):
    x = pair.vector
    assert abs(x @ (B @ x) - 1.0) < 1e-9
    resid = np.linalg.norm(A @ x - pair.value * (B @ x)) / np.linalg.norm(A @ x)
    assert resid < RESID, resid
    if lam_ref is not None:
        assert abs(pair.value - lam_ref) <= EIG_REL * abs(lam_ref), (pair.value, lam_ref)
    return resid


def test_slepc_eigenpair(ctx):
    """lib.rs:82-119: lambda = 1.4745880937 (tolerance 1e-9) and the vector of test_evec.dat."""
    print("[sparse-eigen] slepc_problem", flush=True)
    dom, glq = cases.slepc_problem()
    gep = F.galerkin_sample_gep_hcurl(dom, glq, ctx=ctx)
    pair = _sparse(gep, 1.475, ctx)
    lam = struct.unpack(">d", open(golden("test_eval.dat"), "rb").read())[0]
    raw = open(golden("test_evec.dat"), "rb").read()
    v = np.array(struct.unpack(f">{struct.unpack('>i', raw[4:8])[0]}d", raw[8:]))
    A, B = gep.to_scipy()
    _check_pair(A, B, pair, lam)
    assert abs(pair.value - 1.4745880937) < 1e-9
    x = pair.vector
    cos = abs(x @ (B @ v)) / np.sqrt((x @ (B @ x)) * (v @ (B @ v)))
    assert 1.0 - cos < 1e-9


CASES = [("waveguide_small", F.HierPoly, 10.0), ("waveguide_small", F.HierPoly, 45.0),
         ("mixed_refinements", F.KOLShapeFn, 3.0), ("create_domain", F.HierPoly, 20.0),
         ("small_hp_irregular", F.KOLShapeFn, 5.0), ("small_many_waveguides", F.HierPoly, 9.0)]


@pytest.mark.parametrize("name,BS,target", CASES)
def test_sparse_matches_oracle(ctx, name, BS, target):
    print(f"[sparse-eigen] case={name} target={target}", flush=True)
    dom, glq = cases.ALL[name]()
    dd = F.DeviceDomain(dom, ctx)
    dev = dd.assemble(glq, BS)
    pair = _sparse(dev, target, ctx)
    gep = dd.download(dd.assemble(glq, BS))
    A, B = gep.to_scipy()
    lam, x_ref = OE.sparse_solve_gep(A, B, target)
    resid = _check_pair(A, B, pair, lam)
    # the oracle's vector is the same one up to sign when no other eigenvalue is that close
    w = OE.sparse_eigenvalues_near(A, B, lam, 4)
    if np.sum(np.abs(w - lam) <= 1e-8 * abs(lam)) == 1:
        x = pair.vector
        cos = abs(x @ (B @ x_ref)) / np.sqrt((x @ (B @ x)) * (x_ref @ (B @ x_ref)))
        assert 1.0 - cos < 1e-8
    print(f"[sparse-eigen] case={name} n={gep.n} lambda={pair.value!r} oracle={lam!r} resid={resid:.2e}",
          flush=True)


def test_sparse_matches_dense_gpu_solve(ctx):
    """Precision-study matrices (orders 8, 12x12 Gauss, 4 496 DoFs): the sparse and the dense GPU
    solves find the same TE10 pair."""
    print("[sparse-eigen] precision_study vs dense", flush=True)
    w = W.precision_study()
    dd = F.DeviceDomain(w.domain, ctx)
    dense = F.solve_gep_dense(dd.assemble(w.glq), 10.0, ctx=ctx)
    sparse = _sparse(dd.assemble(w.glq), 10.0, ctx)
    assert abs(sparse.value - dense.value) <= EIG_REL * dense.value
    A, B = dd.download(dd.assemble(w.glq)).to_scipy()
    x, y = sparse.vector, dense.vector
    cos = abs(x @ (B @ y)) / np.sqrt((x @ (B @ x)) * (y @ (B @ y)))
    assert 1.0 - cos < 1e-9


def test_fp32r_input(ctx):
    """A GEP stored as FP32R (FP32 plus a 16-bit residual per value): the same eigenvalue."""
    dom, glq = cases.waveguide_small()
    dd = F.DeviceDomain(dom, ctx)
    p64 = F.solve_gep_sparse(dd.assemble(glq), 10.0, ctx=ctx)
    p32r = F.solve_gep_sparse(dd.assemble(glq, precision=F.FP32R), 10.0, ctx=ctx)
    assert abs(p32r.value - p64.value) <= 1e-10 * p64.value


def test_slepc_solve_gep_dispatch(ctx):
    """slepc_solve_gep picks the dense solve for small GEPs and the sparse one beyond."""
    dom, glq = cases.slepc_problem()
    gep = F.galerkin_sample_gep_hcurl(dom, glq, ctx=ctx)
    assert abs(F.slepc_solve_gep(gep, 1.475, ctx=ctx).value - 1.4745880937) < 1e-9
    dom, glq = cases.small_hp_irregular()
    gep = F.galerkin_sample_gep_hcurl(dom, glq, F.KOLShapeFn, ctx=ctx)
    assert gep.n > F.linalg.SLEPC_DENSE_LIMIT
    assert abs(F.slepc_solve_gep(gep, 5.0, ctx=ctx).value - F.solve_gep_sparse(gep, 5.0, ctx=ctx).value) == 0.0


def test_benchmark_gep_te10(ctx):
    """The 445 380-DoF benchmark GEP (64x32 waveguide, orders 6, two T levels), solved from the
    device-resident CSR: TE10 = pi^2 within 1e-6, B-normalised, residual on the host matrices."""
    print("[sparse-eigen] waveguide_64x32 full size", flush=True)
    w = W.waveguide_64x32()
    dd = F.DeviceDomain(w.domain, ctx)
    pair = _sparse(dd.assemble(w.glq), 10.0, ctx)
    print(f"[sparse-eigen] waveguide_64x32 lambda={pair.value!r}", flush=True)
    assert pair.vector.size == 445380
    assert abs(pair.value - np.pi ** 2) / np.pi ** 2 < 1e-6
    A, B = dd.download(dd.assemble(w.glq)).to_scipy()
    resid = _check_pair(A, B, pair)
    print(f"[sparse-eigen] waveguide_64x32 residual {resid:.2e}", flush=True)
