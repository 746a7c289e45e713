"""The HIP assembly at the full sizes of the five workload configurations.

* Sampled rows against the oracle: the oracle's row-mask mode integrates only the pairs that
  land in the selected rows, so those rows come out complete and exact in seconds at any size.
  Pattern bit-exact, values within _compare.RTOL.
* Whole-matrix structure: CSR well-formed, columns strictly ascending from the diagonal, B's
  diagonal positive.
* Physics: the TE10 cutoff (pi / a)^2 of the air-filled guides from the GPU matrices, for the
  single guide and for every instance of the 256-guide batch."""

import numpy as np
import pytest
import scipy.sparse.linalg as spl

import fem_2d_amd as F
import oracle
from _compare import assert_rows_match
from fem_2d_amd import workloads as W

pytestmark = pytest.mark.gpu

FULL = {
    "waveguide_small": W.waveguide_small,
    "waveguide_64x32": W.waveguide_64x32,
    "hp_irregular": W.hp_irregular,
    "precision_study": W.precision_study,
    "many_waveguides": W.many_waveguides,
}


@pytest.fixture(scope="module")
def ctx():
    return F.Context(0)


_cache = {}


def _workload(name):
    if name not in _cache:
        _cache.clear()
        _cache[name] = FULL[name]()
    return _cache[name]


def _assemble(w, ctx):
    dd = F.DeviceDomain(w.domain, ctx)
    return dd.download(dd.assemble(w.glq, w.basis, F.CurlCurl, F.L2Inner))


def _structure(g):
    rp = g.row_ptr
    assert rp[0] == 0 and rp[-1] == g.nnz and np.all(np.diff(rp) >= 1)
    rows = g.rows()
    first = rp[:-1]
    np.testing.assert_array_equal(g.col[first], np.arange(g.n))  # the diagonal leads every row
    inner = np.ones(g.nnz, bool)
    inner[first] = False
    assert np.all(np.diff(g.col.astype(np.int64))[inner[1:]] > 0)  # strictly ascending within rows
    assert np.all(g.col >= rows)
    assert np.all(g.b_val[first] > 0)


@pytest.mark.parametrize("name", list(FULL))
def test_full_size_sampled_rows_match_oracle(ctx, name):
    print(f"[full-size] workload={name}", flush=True)
    w = _workload(name)
    g = _assemble(w, ctx)
    _structure(g)
    rng = np.random.default_rng(7)
    mask = np.zeros(g.n, bool)
    mask[rng.choice(g.n, min(g.n, 3000), replace=False)] = True
    mask[:64] = True
    mask[-64:] = True
    rc, ref = oracle.assemble(w.domain.view(), basis=w.basis.code, glq=w.glq, n_threads=16, row_mask=mask)
    assert rc == 0
    worst = assert_rows_match(g, ref, mask)
    print(f"[full-size] workload={name} n={g.n} nnz={g.nnz} rows_checked={int(mask.sum())} worst={worst}",
          flush=True)


def _eigenvalue_near(A, B, sigma, iters=40):
    """Shift-invert inverse iteration: the generalized eigenvalue closest to sigma. With sigma
    at 0.9 of the TE10 cutoff the gradient null space (eigenvalue 0) and the next modes are
    at least 9x farther away, so every iteration contracts the error by 10 or more."""
    lu = spl.splu((A - sigma * B).tocsc())
    x = np.random.default_rng(0).standard_normal(A.shape[0])
    for _ in range(iters):
        y = lu.solve(B @ x)
        x = y / np.sqrt(y @ (B @ y))
    return float(x @ (A @ x)) / float(x @ (B @ x))


def test_waveguide_te10_from_gpu_matrices(ctx):
    w = _workload("waveguide_small")
    A, B = _assemble(w, ctx).to_scipy()
    te10 = W.te10_cutoff(w.meta["a"])
    lam = _eigenvalue_near(A.tocsc(), B.tocsc(), 0.9 * te10)
    assert abs(lam - te10) / te10 < 1e-6, lam


def test_many_waveguides_te10_per_instance(ctx):
    """Every instance of the 256-guide batch: the GEP decouples by instance and its lowest
    non-gradient eigenvalue is the instance's TE10 cutoff (pi / a)^2."""
    w = _workload("many_waveguides")
    g = _assemble(w, ctx)
    A, B = g.to_scipy()
    A, B = A.tocsr(), B.tocsr()
    a = w.domain.arrays()
    parent = a["parent"]
    base = np.arange(len(parent))
    while np.any(parent[base] >= 0):  # base Elem of every Elem
        up = parent[base] >= 0
        base[up] = parent[base[up]]
    bs_elem = np.repeat(np.arange(len(parent)), np.diff(a["bs_offsets"]))
    worst = 0.0
    for inst in w.meta["instances"]:
        e0, e1 = inst["elems"]
        dofs = np.unique(a["bs_dof"][(base[bs_elem] >= e0) & (base[bs_elem] < e1)])
        Ai = A[dofs][:, dofs].tocsc()
        Bi = B[dofs][:, dofs].tocsc()
        te10 = W.te10_cutoff(inst["a"])
        lam = _eigenvalue_near(Ai, Bi, 0.9 * te10)
        rel = abs(lam - te10) / te10
        worst = max(worst, rel)
        # p = 2 on 4 Elems across is the coarsest instance; its discretisation error is ~1e-4
        assert rel < 2e-3, (inst, lam, te10)
    print(f"[full-size] many_waveguides: {len(w.meta['instances'])} TE10 cutoffs, worst relative error {worst:.2e}",
          flush=True)
