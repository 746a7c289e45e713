"""Mesh construction and hp-refinement, pinned by the reference's own doc-tests and unit
tests (src/fem_domain/domain/mesh.rs, mesh/h_refinement.rs, mesh/p_refinement.rs)."""

import math

import pytest

import fem_2d_amd as F
from conftest import golden


def test_unit_geometry():
    """mesh.rs:363-428"""
    m = F.Mesh.unit()
    p = m.elem_points(0)
    assert abs((p[2][1] - p[0][1]) - 2.0) < 1e-14
    assert abs((p[1][0] - p[0][0]) - 2.0) < 1e-14
    assert abs(math.dist(p[0], p[3]) - math.sqrt(8.0)) < 1e-14
    e0 = m.edge(0)
    assert abs(e0.length - 2.0) < 1e-14


def test_descendants_and_ancestors():
    """mesh.rs:448-468, 505-518"""
    m = F.Mesh.unit()
    m.h_refine_elems([0], F.HRef.T)
    m.h_refine_elems([1], F.HRef.T)
    m.h_refine_elems([5], F.HRef.U(None))
    assert set(m.descendant_elems(1, True)) == {1, 5, 6, 7, 8, 9, 10}
    assert set(m.descendant_elems(1, False)) == {5, 6, 7, 8, 9, 10}
    assert m.ancestor_elems(10, True) == [10, 5, 1, 0]
    assert m.ancestor_elems(10, False) == [5, 1, 0]


def test_descendant_edges_share_nodes():
    """mesh.rs:552-575"""
    m = F.Mesh.unit()
    m.h_refine_elems([0], F.HRef.U(None))
    m.h_refine_elems([1, 2], F.HRef.U(None))
    e0 = m.edge(0)
    relevant = {e0.child_node, e0.nodes[0], e0.nodes[1]}
    for ce in m.descendant_edges(0, True):
        assert relevant & set(m.edge(ce).nodes)


def test_max_expansion_orders():
    """mesh.rs:608-625"""
    m = F.Mesh.unit()
    m.h_refine_elems([0], F.HRef.T)
    m.set_expansion_orders([(1, F.Orders(1, 1)), (2, F.Orders(3, 3)), (3, F.Orders(2, 7)), (4, F.Orders(4, 1))])
    assert m.max_expansion_orders() == [4, 7]


def test_global_h_refinement_counts():
    """mesh.rs:700-712"""
    m = F.Mesh.unit()
    assert m.num_elems == 1
    m.global_h_refinement(F.HRef.T)
    assert m.num_elems == 5
    m.global_h_refinement(F.HRef.V(None))
    assert m.num_elems == 13


def test_h_refine_elems_counts_and_errors():
    """mesh.rs:737-752"""
    m = F.Mesh.unit()
    m.h_refine_elems([0], F.HRef.T)
    assert m.num_elems == 5
    m.h_refine_elems([2, 3, 4], F.HRef.V(None))
    assert m.num_elems == 11
    with pytest.raises(F.HRefError) as e:
        m.h_refine_elems([15], F.HRef.T)
    assert e.value.kind == "ElemDoesNotExist"
    with pytest.raises(F.HRefError) as e:
        m.h_refine_elems([0], F.HRef.T)
    assert e.value.kind == "ElemNotRefineable"
    with pytest.raises(F.HRefError) as e:
        m.h_refine_elems([1, 1], F.HRef.T)
    assert e.value.kind == "DuplicateElemIds"
    with pytest.raises(F.HRefError) as e:  # a repeat that is not adjacent
        m.h_refine_elems([2, 1, 3, 1], F.HRef.T)
    assert e.value.kind == "DuplicateElemIds"
    assert m.num_elems == 11


def test_h_refine_with_filter():
    """mesh.rs:772-788"""
    m = F.Mesh.unit()
    m.h_refine_elems([0], F.HRef.T)
    m.p_refine_elems([1, 2], F.PRef(3, 3))
    m.h_refine_with_filter(lambda e: F.HRef.U(None) if e.poly_orders.ni >= 3 and e.poly_orders.nj >= 3 else None)
    assert m.num_elems == 9


def test_execute_h_refinements_combines():
    """mesh.rs:821-843: U + V on one Elem is a T refinement; errors leave the mesh untouched."""
    m = F.Mesh.unit()
    m.execute_h_refinements([(0, F.HRef.T)])
    assert m.num_elems == 5
    m.execute_h_refinements([(1, F.HRef.U(None)), (1, F.HRef.V(None))])
    assert m.num_elems == 9
    assert len(m.elem(1).child_ids()) == 4
    with pytest.raises(F.HRefError):
        m.execute_h_refinements([(2, F.HRef.T), (0, F.HRef.T)])
    assert m.num_elems == 9


def test_global_p_refinement():
    """mesh.rs:1228-1264"""
    m = F.Mesh.unit()
    m.global_h_refinement(F.HRef.T)
    m.global_h_refinement(F.HRef.T)
    m.global_p_refinement(F.PRef(3, 5))
    assert all((e.poly_orders.ni, e.poly_orders.nj) == (4, 6) for e in m.elems)

    m = F.Mesh.unit()
    m.global_h_refinement(F.HRef.T)
    m.set_expansion_on_elems([1], F.Orders.try_new(F.MAX_POLYNOMIAL_ORDER - 2, 2))
    m.set_expansion_on_elems([2], F.Orders.try_new(3, 5))
    m.global_p_refinement(F.PRef(3, -3))
    assert (m.elem(1).poly_orders.ni, m.elem(1).poly_orders.nj) == (F.MAX_POLYNOMIAL_ORDER, 1)
    assert (m.elem(2).poly_orders.ni, m.elem(2).poly_orders.nj) == (6, 2)


def test_p_refine_elems():
    """mesh.rs:1290-1308"""
    m = F.Mesh.unit()
    assert (m.elem(0).poly_orders.ni, m.elem(0).poly_orders.nj) == (1, 1)
    m.p_refine_elems([0], F.PRef(5, 2))
    assert (m.elem(0).poly_orders.ni, m.elem(0).poly_orders.nj) == (6, 3)
    with pytest.raises(F.PRefError):
        m.p_refine_elems([15], F.PRef(1, 1))
    with pytest.raises(F.PRefError):
        m.p_refine_elems([0, 0], F.PRef(1, 1))
    with pytest.raises(F.PRefError) as e:  # a repeat that is not adjacent
        m.p_refine_elems([0, 1, 0], F.PRef(1, 1))
    assert e.value.kind == "DuplicateElemIds"


def _iso_aniso_mesh():
    m = F.Mesh.unit()
    m.global_h_refinement(F.HRef.T)
    m.execute_h_refinements([(1, F.HRef.T), (2, F.HRef.T), (3, F.HRef.U(None)), (4, F.HRef.V(None))])
    return m


def test_p_refine_with_filter():
    """mesh.rs:1330-1372"""
    m = _iso_aniso_mesh()
    m.p_refine_with_filter(lambda e: F.PRef(3, 1) if e.h_levels.u >= 2 and e.h_levels.v >= 2 else None)
    for c in m.descendant_elems(1, False) + m.descendant_elems(2, False):
        assert (m.elem(c).poly_orders.ni, m.elem(c).poly_orders.nj) == (4, 2)
    for c in m.descendant_elems(3, False) + m.descendant_elems(4, False):
        assert (m.elem(c).poly_orders.ni, m.elem(c).poly_orders.nj) == (1, 1)


def test_p_refine_with_filter_bounded():
    """mesh.rs:1408-1430"""
    m = _iso_aniso_mesh()
    m.p_refine_with_filter_bounded(
        lambda e, b: F.PRef(b[0][1] - 5, b[1][1] - 5) if e.h_levels.u >= 2 and e.h_levels.v >= 2
        else F.PRef(b[0][0] + 1, b[1][0] + 1))


def test_execute_p_refinements():
    """mesh.rs:1460-1485"""
    m = F.Mesh.unit()
    m.execute_p_refinements([(0, F.PRef(1, 1))])
    assert (m.elem(0).poly_orders.ni, m.elem(0).poly_orders.nj) == (2, 2)
    m.execute_p_refinements([(0, F.PRef(1, 0)), (0, F.PRef(0, 1))])
    assert (m.elem(0).poly_orders.ni, m.elem(0).poly_orders.nj) == (3, 3)
    with pytest.raises(F.PRefError):
        m.execute_p_refinements([(0, F.PRef(1, 1)), (7, F.PRef(1, 1))])
    assert (m.elem(0).poly_orders.ni, m.elem(0).poly_orders.nj) == (3, 3)


def test_expansion_order_setters():
    """mesh.rs:1520-1644"""
    m = F.Mesh.unit()
    m.global_h_refinement(F.HRef.T)
    m.global_h_refinement(F.HRef.T)
    m.set_global_expansion_orders(F.Orders(4, 5))
    assert all((e.poly_orders.ni, e.poly_orders.nj) == (4, 5) for e in m.elems)

    m = F.Mesh.unit()
    m.set_expansion_on_elems([0], F.Orders(6, 3))
    assert (m.elem(0).poly_orders.ni, m.elem(0).poly_orders.nj) == (6, 3)

    m = F.Mesh.unit()
    m.global_h_refinement(F.HRef.T)
    m.h_refine_elems([2, 4], F.HRef.V(None))
    m.set_expansions_with_filter(lambda e: F.Orders(4, 3) if e.has_children() else None)
    assert (m.elem(2).poly_orders.ni, m.elem(2).poly_orders.nj) == (4, 3)
    assert (m.elem(4).poly_orders.ni, m.elem(4).poly_orders.nj) == (4, 3)

    m = F.Mesh.unit()
    m.set_expansion_orders([(0, F.Orders(2, 2))])
    assert (m.elem(0).poly_orders.ni, m.elem(0).poly_orders.nj) == (2, 2)
    with pytest.raises(F.PRefError):
        m.set_expansion_orders([(0, F.Orders(2, 1)), (0, F.Orders(1, 2))])
    with pytest.raises(F.PRefError):
        m.set_expansion_orders([(0, F.Orders(2, 2)), (7, F.Orders(2, 2))])


MESH_A_X = [[0.0, 1.0, 0.0, 1.0], [1.0, 2.0, 1.0, 2.0], [0.0, 1.0, 0.0, 1.0], [1.0, 2.0, 1.0, 2.0]]
MESH_A_Y = [[0.0, 0.0, 0.5, 0.5], [0.0, 0.0, 0.5, 0.5], [0.5, 0.5, 1.0, 1.0], [0.5, 0.5, 1.0, 1.0]]
MESH_A_NEIGHBORS = [[None, 2, None, 1], [None, 3, 0, None], [0, None, None, 3], [1, None, 2, None]]


def test_mesh_from_file():
    """mesh.rs:1849-1892: coordinates, materials and neighbours across active edges."""
    m = F.Mesh.from_file(golden("test_mesh_a.json"))
    for el, eps in zip(range(4), [1.0, 1.0, 2.0, 2.0]):
        _, mats = m.element(el)
        assert abs(mats[0] - eps) < 1e-14
    for eid in range(4):
        e = m.elem(eid)
        for k in range(4):
            nd = m.node(e.nodes[k])
            assert abs(nd.x - MESH_A_X[eid][k]) < 1e-14
            assert abs(nd.y - MESH_A_Y[eid][k]) < 1e-14
            if MESH_A_NEIGHBORS[eid][k] is not None:
                assert m.edge(e.edges[k]).other_active_elem_id(eid) == MESH_A_NEIGHBORS[eid][k]
    for el in range(4):
        pts, _ = m.element(el)
        assert pts[3][0] > pts[0][0] and pts[3][0] >= pts[1][0] and pts[3][0] > pts[2][0]
        assert pts[3][1] > pts[0][1] and pts[3][1] > pts[1][1] and pts[3][1] >= pts[2][1]
        assert pts[0][0] < pts[1][0] and pts[0][1] < pts[2][1]


def test_basic_refinements():
    """mesh.rs:1894-1915"""
    mc = F.Mesh.from_file(golden("test_mesh_c.json"))
    mc.h_refine_elems([0], F.HRef.T)
    mc.h_refine_elems([1, 2], F.HRef.U(None))
    mc.h_refine_elems([3, 4], F.HRef.V(None))
    mb = F.Mesh.from_file(golden("test_mesh_b.json"))
    mb.p_refine_elems([0], F.PRef(2, 2))
    mb.p_refine_elems([1], F.PRef(2, 1))
    mb.p_refine_elems([2], F.PRef(1, 2))
    assert [(mb.elem(k).poly_orders.ni, mb.elem(k).poly_orders.nj) for k in range(3)] == [(3, 3), (3, 2), (2, 3)]


def test_proper_edge_order():
    """mesh.rs:1917-1948: every Elem keeps SW/SE/NW/NE node order after mixed refinements."""
    m = F.Mesh.from_file(golden("test_mesh_b.json"))
    m.global_h_refinement(F.HRef.T)
    boundary = [e.boundary for e in m.edges]
    m.h_refine_with_filter(lambda e: F.HRef.u() if any(boundary[k] for k in e.edges) else None)
    m.h_refine_with_filter(lambda e: F.HRef.T if 4 in e.nodes else F.HRef.V(None))
    m.global_h_refinement(F.HRef.U(1))
    for e in m.elems:
        p = [m.node(n) for n in e.nodes]
        assert p[0].x < p[3].x and p[0].y < p[3].y
        assert abs(p[0].y - p[1].y) < 1e-12 and abs(p[0].x - p[2].x) < 1e-12
        assert abs(p[3].y - p[2].y) < 1e-12 and abs(p[3].x - p[1].x) < 1e-12


@pytest.mark.parametrize("case", ["h_non_existent", "h_with_children", "double_h", "bad_extended", "min_edge",
                                  "p_non_existent", "double_p", "neg_p_i", "neg_p_j", "over_max_i", "over_max_j"])
def test_refinement_errors(case):
    """mesh.rs:1983-2082: the reference's should_panic tests raise errors here."""
    m = F.Mesh.from_file(golden("test_mesh_c.json"))
    with pytest.raises(F.Fem2dError):
        if case == "h_non_existent":
            m.h_refine_elems([0, 1], F.HRef.T)
        elif case == "h_with_children":
            m.h_refine_elems([0], F.HRef.T)
            m.h_refine_elems([0], F.HRef.T)
        elif case == "double_h":
            m.execute_h_refinements([(0, F.HRef.T), (1, F.HRef.T), (0, F.HRef.u())])
        elif case == "bad_extended":
            m.h_refine_elems([0], F.HRef.u_extened(2))
        elif case == "min_edge":
            for _ in range(18):
                m.h_refine_with_filter(lambda e: F.HRef.T if (not e.has_children() and e.nodes[0] == 0) else None)
        elif case == "p_non_existent":
            m.p_refine_elems([0, 1], F.PRef(1, 1))
        elif case == "double_p":
            m.p_refine_elems([0, 0], F.PRef(1, 1))
        elif case == "neg_p_i":
            m.set_global_expansion_orders(F.Orders(3, 3))
            m.p_refine_elems([0], F.PRef(-3, 1))
        elif case == "neg_p_j":
            m.set_global_expansion_orders(F.Orders(3, 3))
            m.p_refine_elems([0], F.PRef(1, -3))
        elif case == "over_max_i":
            m.p_refine_elems([0], F.PRef(F.MAX_POLYNOMIAL_ORDER, 0))
        elif case == "over_max_j":
            m.p_refine_elems([0], F.PRef(0, F.MAX_POLYNOMIAL_ORDER))


def test_min_edge_length_error_kind():
    m = F.Mesh.from_file(golden("test_mesh_c.json"))
    with pytest.raises(F.HRefError) as e:
        for _ in range(18):
            m.h_refine_with_filter(lambda el: F.HRef.T if (not el.has_children() and el.nodes[0] == 0) else None)
    assert e.value.kind == "MinEdgeLength"


ANISO = [[[-1.0, 1.0], [-1.0, 1.0]], [[-1.0, 1.0], [0.0, 1.0]], [[0.0, 1.0], [0.0, 1.0]], [[0.0, 1.0], [0.0, 0.5]],
         [[0.0, 0.5], [0.0, 0.5]]]
ISO = [[[-1.0, 1.0], [-1.0, 1.0]], [[-1.0, 0.0], [-1.0, 0.0]], [[-0.5, 0.0], [-1.0, -0.5]],
       [[-0.5, -0.25], [-0.75, -0.5]], [[-0.375, -0.25], [-0.625, -0.5]]]


@pytest.mark.parametrize("steps,table", [
    ([("V", 1), ("U", 1), ("V", 0), ("U", 0)], ANISO),   # N, E, S, W
    ([("T", 0), ("T", 1), ("T", 2), ("T", 3)], ISO),     # SW, SE, NW, NE
])
def test_h_ref_loc_ranges(steps, table):
    """h_refinement.rs:354-417: parametric ranges along a stack of refinements."""
    m = F.Mesh.unit()
    e = 0
    assert m.parametric_range(e) == table[0]
    for k, (kind, idx) in enumerate(steps):
        r = {"T": F.HRef.T, "U": F.HRef.u(), "V": F.HRef.v()}[kind]
        m.h_refine_elems([e], r)
        e = m.elem(e).children[idx]
        got = m.parametric_range(e)
        for uv in range(2):
            for mm in range(2):
                assert abs(got[uv][mm] - table[k + 1][uv][mm]) < 1e-14
    # relative range from the first ancestor equals the absolute one
    assert m.parametric_range(e, 0) == m.parametric_range(e)


def test_pref_constraints():
    """p_refinement.rs:356-378"""
    assert F.PRef(4, 6).constrained_to([[0, 1], [-2, 5]]).as_array() == [1, 5]
    assert F.PRef(-3, -2).constrained_to([[-2, 4], [0, 3]]).as_array() == [-2, 0]
    pr = F.PRef(0, 10)
    assert pr.falls_within([[0, 1], [-2, 10]])
    assert not pr.falls_within([[-2, -1], [3, 10]])
    assert not pr.falls_within([[-1, 1], [11, 14]])


def test_json_errors():
    with pytest.raises(F.MeshPanic):
        F.Mesh.from_json('{"Elements": [{"materials": [1,0,1,0], "node_ids": [0,1,2,2]}], "Nodes": [[0,0],[1,0],[0,1]]}')
    with pytest.raises(F.MeshPanic):
        F.Mesh.from_json('{"Elements": [], "Nodes": [[0,0],[0,0]]}')
    with pytest.raises(FileNotFoundError):
        F.Mesh.from_file("/nonexistent/mesh.json")
