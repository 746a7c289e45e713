"""Domain construction (basis specs, edge matching, DoF numbering) pinned by the reference's
doc-tests (src/fem_domain/domain.rs) and by the DoF counts of its known-answer problems."""

import numpy as np
import pytest

import fem_2d_amd as F
from fem_2d_amd import workloads as W

import cases


def test_unit_local_basis_specs():
    """domain.rs:240-251: 2 elem-type basis specs in each direction."""
    m = F.Mesh.unit()
    m.set_global_expansion_orders(F.Orders(2, 2))
    d = F.Domain.from_mesh(m, F.ContinuityCondition.HCurl)
    assert len(d.local_basis_specs(0)) == 4


def test_descendant_basis_specs():
    """domain.rs:270-290: 4 descendants with 4 elem-type and 4 edge-type specs each."""
    m = F.Mesh.unit()
    m.set_global_expansion_orders(F.Orders(2, 2))
    m.global_h_refinement(F.HRef.T)
    d = F.Domain.from_mesh(m, F.ContinuityCondition.HCurl)
    desc = d.descendant_basis_specs(0)
    assert len(desc) == 4
    assert all(len(bs) == 8 for _, bs in desc)


def test_ancestor_basis_specs():
    """domain.rs:315-338"""
    m = F.Mesh.unit()
    m.set_global_expansion_orders(F.Orders(2, 2))
    m.global_h_refinement(F.HRef.T)
    m.h_refine_elems([1], F.HRef.T)
    d = F.Domain.from_mesh(m, F.ContinuityCondition.HCurl)
    anc = dict(d.ancestor_basis_specs(5))
    assert len(anc) == 2
    assert len(anc[0]) == 0
    assert len(anc[1]) == 4


@pytest.mark.parametrize("name,expected", [
    ("nalg_problem", 60),       # lib.rs:45-80
    ("readme_example", 600),    # README.md example
    ("slepc_problem", 624),     # test_evec.dat holds 624 entries
    ("waveguide_small", 1096),  # 36 leaves x 24 + 58 active edges x 4
])
def test_dof_counts(name, expected):
    d, _ = cases.ALL[name]()
    assert d.num_dofs == expected


def test_precision_study_dofs():
    w = W.precision_study()
    assert w.domain.num_dofs == 36 * 112 + 58 * 8


def test_dof_numbering_rules():
    """Elem-type DoFs of leaves come first in elem order, then edge DoFs; every DoF has one
    (elem-type) or two (edge-type, on the edge's active pair) basis specs (domain.rs:83-149)."""
    d, _ = cases.mixed_refinements()
    mesh = d.mesh
    a = d.arrays()
    n_elem_dofs = 0
    for e in range(mesh.num_elems):
        el = mesh.elem(e)
        if el.has_children():
            continue
        ni, nj = el.poly_orders.ni, el.poly_orders.nj
        n_elem_dofs += ni * (nj - 1) + (ni - 1) * nj
    for dof in range(d.num_dofs):
        specs = d.dof_basis_specs(dof)
        if dof < n_elem_dofs:
            assert len(specs) == 1
            assert not mesh.elem(specs[0][0]).has_children()
        else:
            assert len(specs) == 2
            e0, e1 = specs[0][0], specs[1][0]
            # the two specs sit on the active elem pair of one edge
            shared = set(mesh.elem(e0).edges) & set(mesh.elem(e1).edges)
            assert any(mesh.edge(k).active_elems in ((e0, e1), (e1, e0)) for k in shared)
    # each elem's basis specs are ordered by DoF id
    for e in range(a["n_elems"]):
        dofs = a["bs_dof"][a["bs_offsets"][e]:a["bs_offsets"][e + 1]]
        assert np.all(np.diff(dofs) > 0)


def test_blank_domains():
    assert F.Domain.blank(F.ContinuityCondition.HCurl).num_dofs == 0
    d = F.Domain.blank(F.ContinuityCondition.HDiv)
    assert d.view().cc == 1


def test_non_hcurl_domain_panics():
    """BasisSpec::new is only implemented for H(Curl) (basis_spec.rs:132-145)."""
    with pytest.raises(F.MeshPanic):
        F.Domain.from_mesh(F.Mesh.unit(), F.ContinuityCondition.HDiv)
