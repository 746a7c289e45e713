"""Data-parallel h-refinement batches against the sequential loop (mesh.rs:844-914).

Batches of at least 256 Elems run on host threads with ids planned from per-refinement counts;
the Mesh must come out identical to the one the sequential loop (n_threads=1) builds: every Elem,
Edge and Node (ids, connectivity, splits, geometry, activation), and the Domain built from it.
Batches that would fail part-way (an Edge split below the minimum length) must fail the same way
and leave the same Mesh."""

import numpy as np
import pytest

import fem_2d_amd as F
from fem_2d_amd.workloads import SplitMix64

PICKS = {
    "T": lambda u: F.HRef.T if u < 0.5 else None,
    "TUV": lambda u: F.HRef.T if u < 0.2 else F.HRef.u() if u < 0.35 else F.HRef.v() if u < 0.5 else None,
    "ext": lambda u: (F.HRef.T if u < 0.15 else F.HRef.U(0) if u < 0.3 else F.HRef.V(1) if u < 0.45
                      else F.HRef.u() if u < 0.55 else None),
}


def _batch(mesh, rng, pick):
    refs = []
    for e in range(mesh.num_elems):
        if mesh.elem_is_h_refineable(e):
            r = pick(rng.uniform())
            if r is not None:
                refs.append((e, r))
    return refs


def _assert_same_mesh(a, b):
    assert (a.num_elems, a.num_edges, a.num_nodes) == (b.num_elems, b.num_edges, b.num_nodes)
    assert a.elems == b.elems
    assert a.edges == b.edges
    assert a.nodes == b.nodes
    va = F.Domain.from_mesh_serial(a).arrays()
    vb = F.Domain.from_mesh_serial(b).arrays()
    for k in ("bs_offsets", "bs_dir", "bs_i", "bs_j", "bs_dof", "dof_bs", "parent", "loc"):
        np.testing.assert_array_equal(va[k], vb[k], err_msg=k)


@pytest.mark.parametrize("pick", list(PICKS))
@pytest.mark.parametrize("threads", [0, 3])
def test_parallel_batches_match_the_sequential_loop(pick, threads):
    rng_a, rng_b = SplitMix64(11), SplitMix64(11)
    seq = F.Mesh.rect_grid(32, 20, 0.0, 2.0, 0.0, 1.0)
    par = F.Mesh.rect_grid(32, 20, 0.0, 2.0, 0.0, 1.0)
    for level in range(3):
        ra, rb = _batch(seq, rng_a, PICKS[pick]), _batch(par, rng_b, PICKS[pick])
        assert ra == rb
        if level == 0:
            assert len(ra) >= 256  # the data-parallel path runs
        seq.execute_h_refinements(ra, n_threads=1)
        par.execute_h_refinements(rb, n_threads=threads)
        _assert_same_mesh(seq, par)


def test_parallel_batch_on_irregular_mesh_with_shared_splits():
    """Neighbours refined in one batch share their split Edges; some Edges were split before."""
    seq = F.Mesh.rect_grid(40, 20, 0.0, 1.0, 0.0, 0.5)
    seq.set_global_expansion_orders(F.Orders(3, 3))
    seq.h_refine_elems(list(range(0, 800, 7)), F.HRef.T)
    par = seq.clone()
    refs = [(e, F.HRef.T if e % 3 else F.HRef.u()) for e in range(seq.num_elems) if seq.elem_is_h_refineable(e)]
    assert len(refs) >= 256
    seq.execute_h_refinements(refs, n_threads=1)
    par.execute_h_refinements(refs, n_threads=0)
    _assert_same_mesh(seq, par)


def test_parallel_batch_error_matches_the_sequential_loop():
    """A batch whose splits would go below MIN_EDGE_LENGTH fails as the sequential loop does,
    leaving the same partially refined Mesh."""
    def deep():
        m = F.Mesh.rect_grid(16, 16, 0.0, 1.0, 0.0, 1.0)
        for _ in range(10):  # down to 2^-14 edges, whose split would go below MIN_EDGE_LENGTH
            m.h_refine_elems([m.num_elems - 1], F.HRef.T)
        return m

    results = []
    for threads in (1, 0):
        m = deep()
        refs = [(e, F.HRef.T) for e in range(m.num_elems) if m.elem_is_h_refineable(e)]
        assert len(refs) >= 256
        try:
            m.execute_h_refinements(refs, n_threads=threads)
            err = None
        except F.Fem2dError as exc:
            err = (type(exc), str(exc))
        results.append((err, m))
    assert results[0][0] is not None and results[0][0][0] is F.HRefError
    assert results[0][0] == results[1][0]
    _assert_same_mesh(results[0][1], results[1][1])
