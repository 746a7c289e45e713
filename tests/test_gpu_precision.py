"""The mixed-precision storage modes: the FP64 arithmetic of the plain path, with values rounded
once, at the store.

* FEM2D_PRECISION_FP32: A and B stored as FP32 (12 bytes per stored entry instead of 20);
* FEM2D_PRECISION_FP32R: A and B stored as FP32 plus a 16-bit residual each (16 bytes per stored
  entry): x = hi + q 2^(e - 38), hi = fl32(x), e the binary exponent of hi, q in int16.

Checked against

* the FP64 GPU path: the same pattern, and every stored value the FP64 value rounded to the
  mode's format bit for bit (FP32: to float; FP32R: the float plus its residual, restated below
  in numpy), and within the format's error bound (2^-24, 2^-38 relative) - the arithmetic is the
  same and the rounding happens once - on small cases of both shape-function spaces and over
  the whole matrices of the full benchmark, hp-irregular (KOL, 1.78 M DoFs) and 256-waveguide
  workloads;
* the oracle: maximum relative entry error (the parity definition of _compare, with its diagonal
  floor) at most MAX_REL per matrix, on small cases and on sampled rows at full size;
* the modes: the Rayleigh quotients of the 10 lowest non-spurious FP64 eigenvectors (eigenvalues
  above the gradient null space) under the stored matrices, relative error at most the mode's
  bound, on the precision-study configuration (config-1 mesh at orders 8, 12x12 Gauss points)
  and the 1096-DoF waveguide; and the TE10 eigenvector of every instance of the 256-guide batch.

RQ_TOL = 1e-8 is the target bound; FP32R meets it. Four-byte values do not: on the uniform meshes
identical element values round alike and do not average out, and B's entries (which do not
cancel in v^T B v) carry them into the quotient (2.9e-8 on the 1096-DoF waveguide); A's rounding
alone reaches 5.4e-8 on the TE10 modes of the 256-guide batch, where curl-curl cancellation
amplifies it. FP32_RQ_TOL states the bound FP32 storage meets."""

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse.linalg as spl

import cases
import fem_2d_amd as F
import oracle
from _compare import entry_errors
from fem_2d_amd import workloads as W

pytestmark = pytest.mark.gpu

MAX_REL = 1e-6  # largest relative entry error allowed for the mixed modes
RQ_TOL = 1e-8  # largest relative Rayleigh-quotient error of the 10 lowest modes (FP32R)
FP32_RQ_TOL = 1e-7  # the bound FP32 storage of both matrices meets


def fp32_round(x):
    return x.astype(np.float32).astype(np.float64)


def fp32r_round(x):
    """The FP32R format restated: hi = fl32(x), e the binary exponent of hi (of the smallest
    normal for subnormals and zero), q = rint((x - hi) 2^(38 - e)) clamped to int16."""
    hi = x.astype(np.float32)
    eb = ((hi.view(np.uint32) >> 23) & 0xFF).astype(np.int64)
    e = np.maximum(eb, 1) - 127
    q = np.clip(np.rint(np.ldexp(x - hi.astype(np.float64), 38 - e)), -32768, 32767)
    return hi.astype(np.float64) + np.ldexp(q, e - 38)


# mode -> (precision code, Rayleigh-quotient bound, rounding of a value, relative error bound)
MODES = {"fp32": (F.FP32, FP32_RQ_TOL, fp32_round, 2.0 ** -24),
         "fp32r": (F.FP32R, RQ_TOL, fp32r_round, 2.0 ** -38)}


@pytest.fixture(scope="module")
def ctx():
    return F.Context(0)


def _precision_study():
    w = W.precision_study()
    return w.domain, w.glq


# case -> (domain factory, shape-function space)
CASES = {
    "nalg_problem": (cases.nalg_problem, F.HierPoly),
    "slepc_problem": (cases.slepc_problem, F.HierPoly),
    "mixed_refinements": (cases.mixed_refinements, F.HierPoly),
    "waveguide_small": (cases.waveguide_small, F.HierPoly),
    "precision_study": (_precision_study, F.HierPoly),
    "mixed_refinements_kol": (cases.mixed_refinements, F.KOLShapeFn),
    "small_hp_irregular_kol": (cases.small_hp_irregular, F.KOLShapeFn),
    "deep_unit_kol": (cases.deep_unit, F.KOLShapeFn),
}


def _rounded(g, g64, mode):
    """g's values are g64's rounded to the mode's format, bit for bit, and within its bound."""
    _, _, rnd, bound = MODES[mode]
    np.testing.assert_array_equal(g.row_ptr, g64.row_ptr)
    np.testing.assert_array_equal(g.col, g64.col)
    for got, exp in ((g.a_val, g64.a_val), (g.b_val, g64.b_val)):
        assert np.array_equal(got.view(np.uint64), rnd(exp).view(np.uint64))
        assert np.all(np.abs(got - exp) <= bound * np.abs(exp))


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name", list(CASES))
def test_values_are_fp64_values_rounded(ctx, name, mode):
    print(f"[mixed] rounding case={name} mode={mode}", flush=True)
    prec = MODES[mode][0]
    make, BS = CASES[name]
    dom, glq = make()
    g64 = F.galerkin_sample_gep_hcurl(dom, glq, BS, F.CurlCurl, F.L2Inner, ctx=ctx)
    g = F.galerkin_sample_gep_hcurl(dom, glq, BS, F.CurlCurl, F.L2Inner, ctx=ctx, precision=prec)
    _rounded(g, g64, mode)
    # the device-resident path stores the same values, in FP32 arrays (and the residuals)
    dd = F.DeviceDomain(dom, ctx)
    dev = dd.assemble(glq, BS, F.CurlCurl, F.L2Inner, precision=prec)
    assert dev.precision == prec
    assert dev.a_val32 and dev.b_val32 and not dev.a_val and not dev.b_val
    assert bool(dev.res16) == (prec == F.FP32R)
    gd = dd.download(dev)
    _rounded(gd, g64, mode)


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name", list(CASES))
def test_entry_error_against_oracle(ctx, name, mode):
    print(f"[mixed] entry error case={name} mode={mode}", flush=True)
    prec = MODES[mode][0]
    make, BS = CASES[name]
    dom, glq = make()
    g = F.galerkin_sample_gep_hcurl(dom, glq, BS, F.CurlCurl, F.L2Inner, ctx=ctx, precision=prec)
    rc, ref = oracle.assemble(dom.view(), basis=BS.code, glq=glq, n_threads=16)
    assert rc == 0
    np.testing.assert_array_equal(g.row_ptr, ref.row_ptr)
    np.testing.assert_array_equal(g.col, ref.col)
    rows = ref.rows()
    worst = {}
    for m, got, exp in (("A", g.a_val, ref.a_val), ("B", g.b_val, ref.b_val)):
        worst[m] = float(entry_errors(rows, ref.col, got, exp).max(initial=0.0))
        assert worst[m] <= MAX_REL, (m, worst[m])
    print(f"[mixed] case={name} mode={mode} nnz={g.nnz} max relative entry error A {worst['A']:.2e} "
          f"B {worst['B']:.2e}", flush=True)


def _lowest_modes(A, B, k, lo, hi):
    """The k lowest generalized eigenpairs (A, B) with eigenvalue in (lo, hi]; B-orthonormal."""
    lam, vec = sla.eigh(A, B, subset_by_value=(lo, hi), driver="gvx")
    assert lam.size >= k, lam
    return lam[:k], vec[:, :k]


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name", ["waveguide_small", "precision_study"])
def test_rayleigh_quotients_of_lowest_modes(ctx, name, mode):
    """With a = 1, b = 0.5 the air-filled guide's modes are (m pi)^2 + (2 n pi)^2 (9.87, 39.5,
    39.5, 49.3, ...); the gradient null space sits at 0, far below 1e-3 pi^2. The window (1, 300]
    holds the 10 lowest non-spurious modes."""
    print(f"[mixed] rayleigh case={name} mode={mode}", flush=True)
    prec, bound, _, _ = MODES[mode]
    make, BS = CASES[name]
    dom, glq = make()
    g64 = F.galerkin_sample_gep_hcurl(dom, glq, BS, F.CurlCurl, F.L2Inner, ctx=ctx)
    g = F.galerkin_sample_gep_hcurl(dom, glq, BS, F.CurlCurl, F.L2Inner, ctx=ctx, precision=prec)
    A64, B64 = g64.to_dense()
    Am, Bm = g.to_dense()
    lam, vec = _lowest_modes(A64, B64, 10, 1.0, 300.0)
    assert abs(lam[0] - np.pi ** 2) / np.pi ** 2 < 1e-5, lam[0]  # TE10 cutoff of the unit-width guide
    worst = 0.0
    for i in range(lam.size):
        v = vec[:, i]
        rq64 = (v @ A64 @ v) / (v @ B64 @ v)
        rqm = (v @ Am @ v) / (v @ Bm @ v)
        assert abs(rq64 - lam[i]) / lam[i] < 1e-10  # the FP64 matrices reproduce their own modes
        worst = max(worst, abs(rqm - lam[i]) / lam[i])
    print(f"[mixed] case={name} mode={mode} n={g.n} 10 lowest modes {np.round(lam, 4).tolist()} "
          f"max Rayleigh-quotient error {worst:.2e}", flush=True)
    assert worst <= bound, worst


def test_fp32_full_size_download_is_rounded_fp64(ctx):
    """The 25.7 M-entry benchmark GEP: the download streams it in 32 MiB chunks and widens FP32 on
    the way; every value equals the FP64 value rounded to float, across all chunks."""
    print("[fp32] full-size download", flush=True)
    w = W.waveguide_64x32()
    dd = F.DeviceDomain(w.domain, ctx)
    g64 = dd.download(dd.assemble(w.glq))
    g32 = dd.download(dd.assemble(w.glq, precision=F.FP32))
    assert g32.nnz == g64.nnz == 25704606
    _rounded(g32, g64, "fp32")
    gm = dd.download(dd.assemble(w.glq, precision=F.FP32R))
    _rounded(gm, g64, "fp32r")


_full = {}


def _full_size(name, ctx):
    """(workload, FP64 GEP, {mode: GEP}) at full size, cached for the tests of one workload."""
    if name not in _full:
        _full.clear()
        w = W.hp_irregular() if name == "hp_irregular" else W.many_waveguides()
        dd = F.DeviceDomain(w.domain, ctx)
        g64 = dd.download(dd.assemble(w.glq, w.basis, F.CurlCurl, F.L2Inner))
        gm = {m: dd.download(dd.assemble(w.glq, w.basis, F.CurlCurl, F.L2Inner, precision=MODES[m][0]))
              for m in MODES}
        _full[name] = (w, g64, gm)
    return _full[name]


@pytest.mark.parametrize("name", ["hp_irregular", "many_waveguides"])
def test_full_size_mixed_values_and_sampled_rows(ctx, name):
    """The KOL hp-irregular workload (1.78 M DoFs, 10x10 Gauss, block-path row groups) and the
    256-guide batch: over the whole matrices every stored value is the FP64 one rounded to its
    storage type; on 3 000 sampled rows the relative entry error against the oracle is at most
    MAX_REL."""
    print(f"[mixed] full size workload={name}", flush=True)
    w, g64, gm = _full_size(name, ctx)
    for m, g in gm.items():
        _rounded(g, g64, m)
    rng = np.random.default_rng(3)
    mask = np.zeros(g64.n, bool)
    mask[rng.choice(g64.n, min(g64.n, 3000), replace=False)] = True
    rc, ref = oracle.assemble(w.domain.view(), basis=w.basis.code, glq=w.glq, n_threads=16, row_mask=mask)
    assert rc == 0
    rows = g64.rows()
    sel = mask[rows]
    diag = {}
    for m, g in gm.items():
        for name_v, allv, exp in (("A", g.a_val, ref.a_val), ("B", g.b_val, ref.b_val)):
            ref64 = g64.a_val if name_v == "A" else g64.b_val
            d = np.zeros(g.n)
            on = rows == g.col
            d[rows[on]] = np.abs(ref64[on])
            r, c = rows[sel], g.col[sel]
            scale = np.maximum(np.abs(exp), np.sqrt(d[r] * d[c]))
            err = np.abs(allv[sel] - exp)
            rel = np.where(scale > 0, err / np.where(scale > 0, scale, 1.0), np.where(err > 0, np.inf, 0.0))
            diag[(m, name_v)] = float(rel.max(initial=0.0))
            assert diag[(m, name_v)] <= MAX_REL, (m, name_v, diag[(m, name_v)])
    print(f"[mixed] full size workload={name} n={g64.n} nnz={g64.nnz} sampled-row entry errors "
          f"{ {f'{k[0]}:{k[1]}': f'{v:.2e}' for k, v in diag.items()} }", flush=True)


def test_many_waveguides_te10_rayleigh_quotients(ctx):
    """Every instance of the 256-guide batch: the TE10 eigenvector of the instance's FP64 matrices
    (shift-invert inverse iteration at 0.9 of the cutoff), and its Rayleigh quotient under each
    mode's matrices; relative error at most the mode's bound."""
    print("[mixed] many_waveguides TE10 Rayleigh quotients", flush=True)
    w, g64, gm = _full_size("many_waveguides", ctx)
    A, B = (M.tocsr() for M in g64.to_scipy())
    mats = {m: tuple(M.tocsr() for M in g.to_scipy()) for m, g in gm.items()}
    a = w.domain.arrays()
    parent = a["parent"]
    base = np.arange(len(parent))
    while np.any(parent[base] >= 0):
        up = parent[base] >= 0
        base[up] = parent[base[up]]
    bs_elem = np.repeat(np.arange(len(parent)), np.diff(a["bs_offsets"]))
    worst = {m: 0.0 for m in MODES}
    for inst in w.meta["instances"]:
        e0, e1 = inst["elems"]
        dofs = np.unique(a["bs_dof"][(base[bs_elem] >= e0) & (base[bs_elem] < e1)])
        Ai, Bi = A[dofs][:, dofs].tocsc(), B[dofs][:, dofs].tocsc()
        sigma = 0.9 * W.te10_cutoff(inst["a"])
        lu = spl.splu((Ai - sigma * Bi).tocsc())
        x = np.random.default_rng(0).standard_normal(Ai.shape[0])
        for _ in range(40):
            y = lu.solve(Bi @ x)
            x = y / np.sqrt(y @ (Bi @ y))
        rq64 = (x @ (Ai @ x)) / (x @ (Bi @ x))
        for m, (Am, Bm) in mats.items():
            Ami, Bmi = Am[dofs][:, dofs], Bm[dofs][:, dofs]
            rqm = (x @ (Ami @ x)) / (x @ (Bmi @ x))
            worst[m] = max(worst[m], abs(rqm - rq64) / rq64)
    print(f"[mixed] many_waveguides: {len(w.meta['instances'])} TE10 Rayleigh-quotient errors "
          f"{ {m: f'{v:.2e}' for m, v in worst.items()} }", flush=True)
    for m, v in worst.items():
        assert v <= MODES[m][1], (m, v)
