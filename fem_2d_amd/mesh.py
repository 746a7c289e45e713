"""Mesh, h-/p-refinement types: the reference's input-building API
(src/fem_domain/domain/mesh.rs and mesh/*.rs) over the C ABI."""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Callable, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N
from .errors import MeshPanic, PRefError, check

MIN_EDGE_LENGTH = 3.0518e-5  # mesh.rs:36
MAX_POLYNOMIAL_ORDER = 20  # mesh.rs:42

_LOC_NAMES = ["SW", "SE", "NW", "NE", "W", "E", "S", "N"]


class HRef:
    """h-refinement variant (h_refinement.rs:60-157).

    ``HRef.T``, ``HRef.U(None)``, ``HRef.V(None)``; ``HRef.U(0|1)`` / ``HRef.V(0|1)``
    refine the 0th/1st child again in the other direction."""

    _T = 0
    _U = 1
    _V = 2

    def __init__(self, kind: int, ext: Optional[int] = None):
        self.kind = kind
        self.ext = -1 if ext is None else int(ext)

    @classmethod
    def t(cls) -> "HRef":
        return cls(cls._T)

    @classmethod
    def u(cls) -> "HRef":
        return cls(cls._U)

    @classmethod
    def v(cls) -> "HRef":
        return cls(cls._V)

    @classmethod
    def U(cls, ext: Optional[int] = None) -> "HRef":
        return cls(cls._U, ext)

    @classmethod
    def V(cls, ext: Optional[int] = None) -> "HRef":
        return cls(cls._V, ext)

    @classmethod
    def u_extened(cls, child_idx: int) -> "HRef":
        from .errors import HRefError
        if child_idx not in (0, 1):
            raise HRefError("Extended refinement index must be 0 or 1!", 6)
        return cls(cls._U, child_idx)

    @classmethod
    def v_extened(cls, child_idx: int) -> "HRef":
        from .errors import HRefError
        if child_idx not in (0, 1):
            raise HRefError("Extended refinement index must be 0 or 1!", 6)
        return cls(cls._V, child_idx)

    def __eq__(self, other) -> bool:
        return isinstance(other, HRef) and (self.kind, self.ext) == (other.kind, other.ext)

    def __repr__(self) -> str:
        name = "TUV"[self.kind]
        return f"HRef.{name}" if self.ext < 0 else f"HRef.{name}({self.ext})"


HRef.T = HRef(HRef._T)


class PRef:
    """p-refinement deltas (p_refinement.rs:196-260)."""

    def __init__(self, di: int, dj: int):
        self.di = int(di)
        self.dj = int(dj)

    @classmethod
    def from_(cls, di: int, dj: int) -> "PRef":
        return cls(di, dj)

    def as_array(self) -> List[int]:
        return [self.di, self.dj]

    def constrained_to(self, bounds) -> "PRef":
        (u0, u1), (v0, v1) = bounds
        return PRef(min(max(self.di, u0), u1), min(max(self.dj, v0), v1))

    def falls_within(self, bounds) -> bool:
        (u0, u1), (v0, v1) = bounds
        return u0 <= self.di <= u1 and v0 <= self.dj <= v1

    def __repr__(self) -> str:
        return f"PRef({self.di:+d}, {self.dj:+d})"


class Orders:
    """Expansion orders (p_refinement.rs:266-297); ``Orders(i, j)`` panics like Orders::new."""

    def __init__(self, i: int, j: int):
        if not (1 <= i <= MAX_POLYNOMIAL_ORDER and 1 <= j <= MAX_POLYNOMIAL_ORDER):
            raise MeshPanic(f"Expansion orders outside valid range ({i}, {j})")
        self.i = int(i)
        self.j = int(j)

    @classmethod
    def new(cls, i: int, j: int) -> "Orders":
        return cls(i, j)

    @classmethod
    def try_new(cls, i: int, j: int) -> "Orders":
        if not (1 <= i <= MAX_POLYNOMIAL_ORDER and 1 <= j <= MAX_POLYNOMIAL_ORDER):
            raise PRefError(f"Invalid expansion orders: ({i}, {j})", 6)
        return cls(i, j)

    def to_array(self) -> List[int]:
        return [self.i, self.j]


@dataclass(frozen=True)
class PolyOrders:
    ni: int
    nj: int


@dataclass(frozen=True)
class HLevels:
    u: int
    v: int


@dataclass(frozen=True)
class Elem:
    """Snapshot of an Elem (elem.rs:100-240)."""

    id: int
    nodes: Tuple[int, int, int, int]
    edges: Tuple[int, int, int, int]
    element: int
    parent: Optional[int]
    loc: Optional[str]
    h_levels: HLevels
    poly_orders: PolyOrders
    children: Tuple[int, ...]
    depth: int

    def has_children(self) -> bool:
        return len(self.children) > 0

    def child_ids(self) -> Optional[List[int]]:
        return list(self.children) if self.children else None

    def parent_id(self) -> Optional[int]:
        return self.parent


@dataclass(frozen=True)
class Edge:
    """Snapshot of an Edge (edge.rs:64-286)."""

    id: int
    nodes: Tuple[int, int]
    boundary: bool
    dir: str
    length: float
    parent: Optional[int]
    children: Optional[Tuple[int, int]]
    child_node: Optional[int]
    active_elems: Optional[Tuple[int, int]]

    def active_elem_pair(self):
        return self.active_elems

    def other_active_elem_id(self, elem_id: int) -> Optional[int]:
        if self.active_elems is None:
            return None
        a, b = self.active_elems
        if elem_id == a:
            return b
        if elem_id == b:
            return a
        raise MeshPanic(f"{elem_id} isn't an active Elem on Edge {self.id}; Cannot retrieve other active Elem!")


@dataclass(frozen=True)
class Node:
    id: int
    x: float
    y: float
    boundary: bool


def _i32(seq: Sequence[int]):
    arr = (C.c_int32 * max(1, len(seq)))(*[int(s) for s in seq])
    return arr


class Mesh:
    """Mesh with refinement-by-superposition hp-refinement (mesh.rs:44-1666)."""

    def __init__(self, handle, owner=None):
        self._h = handle
        self._owner = owner  # keeps a Domain alive for borrowed meshes

    def __del__(self):
        try:
            if self._owner is None and self._h:
                N.load().fem2d_mesh_free(self._h)
        except Exception:
            pass

    # -- construction -------------------------------------------------------
    @classmethod
    def _make(cls, fn, *args) -> "Mesh":
        h = C.c_void_p()
        check(fn(*args, C.byref(h)))
        return cls(h)

    @classmethod
    def from_file(cls, path: str) -> "Mesh":
        return cls._make(N.load().fem2d_mesh_from_file, str(path).encode())

    @classmethod
    def from_json(cls, text: str) -> "Mesh":
        return cls._make(N.load().fem2d_mesh_from_json, text.encode())

    @classmethod
    def unit(cls) -> "Mesh":
        return cls._make(N.load().fem2d_mesh_unit)

    @classmethod
    def blank(cls) -> "Mesh":
        return cls._make(N.load().fem2d_mesh_blank)

    @classmethod
    def from_parts(cls, points, node_ids, materials) -> "Mesh":
        """Nodes ``[[x, y], ...]``, elements ``[[n0, n1, n2, n3], ...]`` (SW, SE, NW, NE) and
        materials ``[[eps_re, eps_im, mu_re, mu_im], ...]``; same result as the JSON file."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 2)
        ids = np.ascontiguousarray(node_ids, dtype=np.int32).reshape(-1, 4)
        mats = np.ascontiguousarray(materials, dtype=np.float64).reshape(-1, 4)
        return cls._make(N.load().fem2d_mesh_from_parts, len(pts), pts.ctypes.data_as(C.POINTER(C.c_double)),
                         len(ids), ids.ctypes.data_as(C.POINTER(C.c_int32)),
                         mats.ctypes.data_as(C.POINTER(C.c_double)))

    @classmethod
    def rect_grid(cls, nx: int, ny: int, x0: float, x1: float, y0: float, y1: float,
                  materials=(1.0, 0.0, 1.0, 0.0)) -> "Mesh":
        """Structured nx-by-ny quad mesh of [x0, x1] x [y0, y1]; node (i, j) has id j*(nx+1)+i."""
        xs = np.linspace(x0, x1, nx + 1)
        ys = np.linspace(y0, y1, ny + 1)
        pts = np.array([[xs[i], ys[j]] for j in range(ny + 1) for i in range(nx + 1)], dtype=np.float64)
        ids = []
        for j in range(ny):
            for i in range(nx):
                n0 = j * (nx + 1) + i
                ids.append([n0, n0 + 1, n0 + nx + 1, n0 + nx + 2])
        mats = np.tile(np.asarray(materials, dtype=np.float64), (nx * ny, 1))
        return cls.from_parts(pts, ids, mats)

    def clone(self) -> "Mesh":
        return Mesh._make(N.load().fem2d_mesh_clone, self._h)

    # -- queries ------------------------------------------------------------
    def _counts(self):
        ne, nn, nd, nel = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        check(N.load().fem2d_mesh_counts(self._h, C.byref(ne), C.byref(nn), C.byref(nd), C.byref(nel)))
        return ne.value, nn.value, nd.value, nel.value

    @property
    def num_elems(self) -> int:
        return self._counts()[0]

    @property
    def num_nodes(self) -> int:
        return self._counts()[1]

    @property
    def num_edges(self) -> int:
        return self._counts()[2]

    @property
    def num_elements(self) -> int:
        return self._counts()[3]

    def elem(self, elem_id: int) -> Elem:
        info = N.ElemInfo()
        check(N.load().fem2d_mesh_elem_info(self._h, int(elem_id), C.byref(info)))
        return Elem(
            id=info.id,
            nodes=tuple(info.nodes),
            edges=tuple(info.edges),
            element=info.element,
            parent=None if info.parent < 0 else info.parent,
            loc=None if info.loc < 0 else _LOC_NAMES[info.loc],
            h_levels=HLevels(info.h_levels[0], info.h_levels[1]),
            poly_orders=PolyOrders(info.orders[0], info.orders[1]),
            children=tuple(info.children[k] for k in range(info.n_children)),
            depth=info.depth,
        )

    @property
    def elems(self) -> List[Elem]:
        return [self.elem(k) for k in range(self.num_elems)]

    def edge(self, edge_id: int) -> Edge:
        info = N.EdgeInfo()
        check(N.load().fem2d_mesh_edge_info(self._h, int(edge_id), C.byref(info)))
        return Edge(
            id=info.id,
            nodes=(info.nodes[0], info.nodes[1]),
            boundary=bool(info.boundary),
            dir="U" if info.dir == 0 else "V",
            length=info.length,
            parent=None if info.parent < 0 else info.parent,
            children=None if info.children[0] < 0 else (info.children[0], info.children[1]),
            child_node=None if info.child_node < 0 else info.child_node,
            active_elems=None if info.active[0] < 0 else (info.active[0], info.active[1]),
        )

    @property
    def edges(self) -> List[Edge]:
        return [self.edge(k) for k in range(self.num_edges)]

    def node(self, node_id: int) -> Node:
        x, y, b = C.c_double(), C.c_double(), C.c_int32()
        check(N.load().fem2d_mesh_node_info(self._h, int(node_id), C.byref(x), C.byref(y), C.byref(b)))
        return Node(int(node_id), x.value, y.value, bool(b.value))

    @property
    def nodes(self) -> List[Node]:
        return [self.node(k) for k in range(self.num_nodes)]

    def element(self, element_id: int):
        pts = (C.c_double * 8)()
        mats = (C.c_double * 4)()
        check(N.load().fem2d_mesh_element_info(self._h, int(element_id), pts, mats))
        return [[pts[2 * k], pts[2 * k + 1]] for k in range(4)], list(mats)

    def elem_points(self, elem_id: int):
        e = self.elem(elem_id)
        return [(self.node(n).x, self.node(n).y) for n in e.nodes]

    def _ids(self, fn, ident: int, include: bool) -> List[int]:
        n = C.c_int64()
        check(fn(self._h, int(ident), int(include), None, 0, C.byref(n)))
        buf = (C.c_int32 * max(1, n.value))()
        check(fn(self._h, int(ident), int(include), buf, n.value, C.byref(n)))
        return list(buf[: n.value])

    def descendant_elems(self, elem_id: int, include_starting_elem: bool) -> List[int]:
        return self._ids(N.load().fem2d_mesh_descendant_elems, elem_id, include_starting_elem)

    def ancestor_elems(self, elem_id: int, include_starting_elem: bool) -> List[int]:
        return self._ids(N.load().fem2d_mesh_ancestor_elems, elem_id, include_starting_elem)

    def descendant_edges(self, edge_id: int, include_starting_edge: bool) -> List[int]:
        return self._ids(N.load().fem2d_mesh_descendant_edges, edge_id, include_starting_edge)

    def max_expansion_orders(self) -> List[int]:
        out = (C.c_int32 * 2)()
        check(N.load().fem2d_mesh_max_expansion_orders(self._h, out))
        return [out[0], out[1]]

    def elem_is_h_refineable(self, elem_id: int) -> bool:
        out = C.c_int32()
        check(N.load().fem2d_mesh_elem_is_h_refineable(self._h, int(elem_id), C.byref(out)))
        return bool(out.value)

    def elem_p_refinement_window(self, elem_id: int):
        out = (C.c_int32 * 4)()
        check(N.load().fem2d_mesh_elem_p_refinement_window(self._h, int(elem_id), out))
        return [[out[0], out[1]], [out[2], out[3]]]

    def parametric_range(self, elem_id: int, from_ancestor: Optional[int] = None):
        """Elem::parametric_range (or relative_parametric_range from an ancestor)."""
        out = (C.c_double * 4)()
        check(N.load().fem2d_mesh_elem_parametric_range(
            self._h, int(elem_id), -1 if from_ancestor is None else int(from_ancestor), out))
        return [[out[0], out[1]], [out[2], out[3]]]

    # -- h-refinement -------------------------------------------------------
    def global_h_refinement(self, refinement: HRef) -> None:
        check(N.load().fem2d_mesh_global_h_refinement(self._h, refinement.kind, refinement.ext))

    def h_refine_elems(self, elems: Sequence[int], refinement: HRef) -> None:
        ids = _i32(elems)
        check(N.load().fem2d_mesh_h_refine_elems(self._h, ids, len(elems), refinement.kind, refinement.ext))

    def execute_h_refinements(self, refinements: Iterable[Tuple[int, HRef]], n_threads: int = 0) -> None:
        """Mesh::execute_h_refinements (mesh.rs:844-914). Batches of at least 256 Elems run on
        ``n_threads`` host threads (0: the machine's cores, at most 16; 1: sequentially), with the
        ids of the sequential loop."""
        refs = list(refinements)
        ids = _i32([r[0] for r in refs])
        kinds = _i32([r[1].kind for r in refs])
        exts = _i32([r[1].ext for r in refs])
        check(N.load().fem2d_mesh_execute_h_refinements_threads(self._h, ids, kinds, exts, len(refs), int(n_threads)))

    def h_refine_with_filter(self, filt: Callable[[Elem], Optional[HRef]]) -> None:
        """mesh.rs:789-808: the filter sees every h-refineable Elem in id order."""
        refs = []
        for e in self.elems:
            if self.elem_is_h_refineable(e.id):
                r = filt(e)
                if r is not None:
                    refs.append((e.id, r))
        self.execute_h_refinements(refs)

    # -- p-refinement / expansion orders ------------------------------------
    def global_p_refinement(self, refinement: PRef) -> None:
        check(N.load().fem2d_mesh_global_p_refinement(self._h, refinement.di, refinement.dj))

    def p_refine_elems(self, elems: Sequence[int], refinement: PRef) -> None:
        ids = _i32(elems)
        check(N.load().fem2d_mesh_p_refine_elems(self._h, ids, len(elems), refinement.di, refinement.dj))

    def execute_p_refinements(self, refinements: Iterable[Tuple[int, PRef]]) -> None:
        refs = list(refinements)
        check(N.load().fem2d_mesh_execute_p_refinements(
            self._h, _i32([r[0] for r in refs]), _i32([r[1].di for r in refs]), _i32([r[1].dj for r in refs]),
            len(refs)))

    def p_refine_with_filter(self, filt: Callable[[Elem], Optional[PRef]]) -> None:
        """mesh.rs:1373-1398: refinements are clamped to each Elem's valid window."""
        refs = []
        for e in self.elems:
            r = filt(e)
            if r is not None:
                refs.append((e.id, r.constrained_to(self.elem_p_refinement_window(e.id))))
        self.execute_p_refinements(refs)

    def p_refine_with_filter_bounded(self, filt) -> None:
        """mesh.rs:1431-1448: the closure receives the valid window and must respect it."""
        refs = []
        for e in self.elems:
            r = filt(e, self.elem_p_refinement_window(e.id))
            if r is not None:
                refs.append((e.id, r))
        self.execute_p_refinements(refs)

    def set_global_expansion_orders(self, orders: Orders) -> None:
        check(N.load().fem2d_mesh_set_global_expansion_orders(self._h, orders.i, orders.j))

    def set_expansion_on_elems(self, elems: Sequence[int], orders: Orders) -> None:
        self.set_expansion_orders([(e, orders) for e in elems])

    def set_expansion_orders(self, poly_orders: Iterable[Tuple[int, Orders]]) -> None:
        items = list(poly_orders)
        check(N.load().fem2d_mesh_set_expansion_orders(
            self._h, _i32([k for k, _ in items]), _i32([o.i for _, o in items]), _i32([o.j for _, o in items]),
            len(items)))

    def set_expansions_with_filter(self, filt: Callable[[Elem], Optional[Orders]]) -> None:
        items = []
        for e in self.elems:
            o = filt(e)
            if o is not None:
                items.append((e.id, o))
        self.set_expansion_orders(items)
