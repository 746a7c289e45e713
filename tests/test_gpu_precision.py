"""FEM2D_PRECISION_FP32 (mixed precision): the FP64 arithmetic of the plain path, with the A and B
values stored rounded to FP32 (12 bytes per stored entry instead of 20). Checked against

* the FP64 GPU path: the same pattern, and every FP32 value is the FP64 value rounded to float,
  bit for bit (the arithmetic is the same and the rounding happens once, at the store);
* the oracle: maximum relative entry error (the parity definition of _compare, with its diagonal
  floor) at most MAX_REL, per matrix;
* the modes: the Rayleigh quotients of the 10 lowest non-spurious FP64 eigenvectors (eigenvalues
  above the gradient null space) under the FP32 matrices, relative error at most RQ_TOL, on the
  precision-study configuration (config-1 mesh at orders 8, 12x12 Gauss points) and on the
  1096-DoF waveguide.

Storing 4-byte values does not reach a Rayleigh-quotient error of 1e-8 on these meshes: the
uniform meshes repeat identical element values, whose roundings are equal and do not average
out, and B's entries (which do not cancel in v^T B v) carry them into the quotient (2.9e-8 on
the 1096-DoF waveguide). RQ_TOL states the bound the FP32 storage meets."""

import numpy as np
import pytest
import scipy.linalg as sla

import cases
import fem_2d_amd as F
import oracle
from _compare import entry_errors
from fem_2d_amd import workloads as W

pytestmark = pytest.mark.gpu

MAX_REL = 1e-6  # largest relative entry error allowed for FP32 storage
RQ_TOL = 1e-7  # largest relative Rayleigh-quotient error of the 10 lowest modes


@pytest.fixture(scope="module")
def ctx():
    return F.Context(0)


def _precision_study():
    w = W.precision_study()
    return w.domain, w.glq


CASES = {
    "nalg_problem": cases.nalg_problem,
    "slepc_problem": cases.slepc_problem,
    "mixed_refinements": cases.mixed_refinements,
    "waveguide_small": cases.waveguide_small,
    "precision_study": _precision_study,
}


@pytest.mark.parametrize("name", list(CASES))
def test_fp32_values_are_fp64_values_rounded(ctx, name):
    print(f"[fp32] rounding case={name}", flush=True)
    dom, glq = CASES[name]()
    g64 = F.galerkin_sample_gep_hcurl(dom, glq, F.HierPoly, F.CurlCurl, F.L2Inner, ctx=ctx)
    g32 = F.galerkin_sample_gep_hcurl(dom, glq, F.HierPoly, F.CurlCurl, F.L2Inner, ctx=ctx, precision=F.FP32)
    np.testing.assert_array_equal(g32.row_ptr, g64.row_ptr)
    np.testing.assert_array_equal(g32.col, g64.col)
    for got, exp in ((g32.a_val, g64.a_val), (g32.b_val, g64.b_val)):
        np.testing.assert_array_equal(got, exp.astype(np.float32).astype(np.float64))
    # the device-resident path stores the same FP32 values
    dd = F.DeviceDomain(dom, ctx)
    dev = dd.assemble(glq, F.HierPoly, F.CurlCurl, F.L2Inner, precision=F.FP32)
    assert dev.precision == F.FP32 and not dev.a_val and dev.a_val32
    g32d = dd.download(dev)
    np.testing.assert_array_equal(g32d.a_val, g32.a_val)
    np.testing.assert_array_equal(g32d.b_val, g32.b_val)


@pytest.mark.parametrize("name", list(CASES))
def test_fp32_entry_error_against_oracle(ctx, name):
    print(f"[fp32] entry error case={name}", flush=True)
    dom, glq = CASES[name]()
    g32 = F.galerkin_sample_gep_hcurl(dom, glq, F.HierPoly, F.CurlCurl, F.L2Inner, ctx=ctx, precision=F.FP32)
    rc, ref = oracle.assemble(dom.view(), basis=0, glq=glq, n_threads=16)
    assert rc == 0
    np.testing.assert_array_equal(g32.row_ptr, ref.row_ptr)
    np.testing.assert_array_equal(g32.col, ref.col)
    rows = ref.rows()
    worst = {}
    for m, got, exp in (("A", g32.a_val, ref.a_val), ("B", g32.b_val, ref.b_val)):
        worst[m] = float(entry_errors(rows, ref.col, got, exp).max(initial=0.0))
        assert worst[m] <= MAX_REL, (m, worst[m])
    print(f"[fp32] case={name} nnz={g32.nnz} max relative entry error A {worst['A']:.2e} B {worst['B']:.2e}",
          flush=True)


def _lowest_modes(A, B, k, lo, hi):
    """The k lowest generalized eigenpairs (A, B) with eigenvalue in (lo, hi]; B-orthonormal."""
    lam, vec = sla.eigh(A, B, subset_by_value=(lo, hi), driver="gvx")
    assert lam.size >= k, lam
    return lam[:k], vec[:, :k]


@pytest.mark.parametrize("name", ["waveguide_small", "precision_study"])
def test_fp32_rayleigh_quotients_of_lowest_modes(ctx, name):
    """With a = 1, b = 0.5 the air-filled guide's modes are (m pi)^2 + (2 n pi)^2 (9.87, 39.5,
    39.5, 49.3, ...); the gradient null space sits at 0, far below 1e-3 pi^2. The window (1, 300]
    holds the 10 lowest non-spurious modes."""
    print(f"[fp32] rayleigh case={name}", flush=True)
    dom, glq = CASES[name]()
    g64 = F.galerkin_sample_gep_hcurl(dom, glq, F.HierPoly, F.CurlCurl, F.L2Inner, ctx=ctx)
    g32 = F.galerkin_sample_gep_hcurl(dom, glq, F.HierPoly, F.CurlCurl, F.L2Inner, ctx=ctx, precision=F.FP32)
    A64, B64 = g64.to_dense()
    A32, B32 = g32.to_dense()
    lam, vec = _lowest_modes(A64, B64, 10, 1.0, 300.0)
    assert abs(lam[0] - np.pi ** 2) / np.pi ** 2 < 1e-5, lam[0]  # TE10 cutoff of the unit-width guide
    worst = 0.0
    for i in range(lam.size):
        v = vec[:, i]
        rq64 = (v @ A64 @ v) / (v @ B64 @ v)
        rq32 = (v @ A32 @ v) / (v @ B32 @ v)
        assert abs(rq64 - lam[i]) / lam[i] < 1e-10  # the FP64 matrices reproduce their own modes
        worst = max(worst, abs(rq32 - lam[i]) / lam[i])
    print(f"[fp32] case={name} n={g32.n} 10 lowest modes {np.round(lam, 4).tolist()} "
          f"max Rayleigh-quotient error {worst:.2e}", flush=True)
    assert worst <= RQ_TOL, worst


def test_fp32_full_size_download_is_rounded_fp64(ctx):
    """The 25.7 M-entry benchmark GEP: the download streams it in 32 MiB chunks and widens FP32 on
    the way; every value equals the FP64 value rounded to float, across all chunks."""
    print("[fp32] full-size download", flush=True)
    w = W.waveguide_64x32()
    dd = F.DeviceDomain(w.domain, ctx)
    g64 = dd.download(dd.assemble(w.glq))
    g32 = dd.download(dd.assemble(w.glq, precision=F.FP32))
    assert g32.nnz == g64.nnz == 25704606
    np.testing.assert_array_equal(g32.row_ptr, g64.row_ptr)
    np.testing.assert_array_equal(g32.col, g64.col)
    for got, exp in ((g32.a_val, g64.a_val), (g32.b_val, g64.b_val)):
        np.testing.assert_array_equal(got, exp.astype(np.float32).astype(np.float64))
