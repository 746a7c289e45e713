"""Domain: basis specs and DoF numbering (reference src/fem_domain/domain.rs) over the C ABI."""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from enum import IntEnum
from typing import List, Tuple

import numpy as np

from . import _native as N
from .errors import check
from .mesh import Mesh


class ContinuityCondition(IntEnum):
    """domain.rs:18-23"""

    HCurl = 0
    HDiv = 1
    Discontinuous = 2


class BasisDir(IntEnum):
    """basis_spec.rs:259-268"""

    U = 0
    V = 1
    W = 2


@dataclass(frozen=True)
class BasisSpec:
    """The integration data of a BasisSpec (basis_spec.rs:221-232) plus its owner."""

    i: int
    j: int
    dir: BasisDir
    elem_id: int
    elem_idx: int
    dof_id: int

    def integration_data(self):
        return ([self.i, self.j], self.dir, self.dof_id)


class Domain:
    """Domain (domain.rs:42-159): a Mesh with its basis specs and DoFs."""

    def __init__(self, handle):
        self._h = handle
        self._view = None
        self._arrays = None

    def __del__(self):
        try:
            if self._h:
                N.load().fem2d_domain_free(self._h)
        except Exception:
            pass

    @classmethod
    def from_mesh(cls, mesh: Mesh, cc: ContinuityCondition = ContinuityCondition.HCurl,
                  n_threads: int = 0) -> "Domain":
        """Domain::from_mesh (domain.rs:69-159), built on ``n_threads`` host threads (0: the
        machine's cores, at most 16)."""
        h = C.c_void_p()
        check(N.load().fem2d_domain_from_mesh_threads(mesh._h, int(cc), int(n_threads), C.byref(h)))
        return cls(h)

    @classmethod
    def from_mesh_serial(cls, mesh: Mesh, cc: ContinuityCondition = ContinuityCondition.HCurl) -> "Domain":
        """Domain::from_mesh in the reference's sequential order (the same Domain as from_mesh)."""
        h = C.c_void_p()
        check(N.load().fem2d_domain_from_mesh_serial(mesh._h, int(cc), C.byref(h)))
        return cls(h)

    @classmethod
    def blank(cls, cc: ContinuityCondition) -> "Domain":
        h = C.c_void_p()
        check(N.load().fem2d_domain_blank(int(cc), C.byref(h)))
        return cls(h)

    @classmethod
    def unit(cls, cc: ContinuityCondition = ContinuityCondition.HCurl) -> "Domain":
        return cls.from_mesh(Mesh.unit(), cc)

    @property
    def num_dofs(self) -> int:
        n = C.c_int64()
        check(N.load().fem2d_domain_num_dofs(self._h, C.byref(n)))
        return n.value

    @property
    def mesh(self) -> Mesh:
        h = C.c_void_p()
        check(N.load().fem2d_domain_mesh(self._h, C.byref(h)))
        return Mesh(h, owner=self)

    def view(self) -> N.DomainView:
        """Flattened view (fem2d_domain_view); valid while this Domain lives."""
        if self._view is None:
            v = N.DomainView()
            check(N.load().fem2d_domain_get_view(self._h, C.byref(v)))
            self._view = v
        return self._view

    def arrays(self) -> dict:
        """Numpy copies of the flattened view."""
        if self._arrays is None:
            v = self.view()
            ne = v.n_elems

            def arr(ptr, n, dt):
                if n == 0:
                    return np.zeros(0, dtype=dt)
                return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)

            off = arr(v.bs_offsets, ne + 1, np.int64) if ne > 0 else np.zeros(1, np.int64)
            nbs = int(off[-1])
            self._arrays = dict(
                cc=int(v.cc), n_elems=ne, n_dofs=int(v.n_dofs),
                parent=arr(v.parent, ne, np.int32), loc=arr(v.loc, ne, np.uint8),
                element_box=arr(v.element_box, 4 * ne, np.float64).reshape(-1, 4),
                materials=arr(v.materials, 2 * ne, np.float64).reshape(-1, 2),
                orders=arr(v.orders, 2 * ne, np.uint8).reshape(-1, 2),
                bs_offsets=off, bs_dir=arr(v.bs_dir, nbs, np.uint8), bs_i=arr(v.bs_i, nbs, np.uint8),
                bs_j=arr(v.bs_j, nbs, np.uint8), bs_dof=arr(v.bs_dof, nbs, np.int32),
                dof_bs=arr(v.dof_bs, 2 * int(v.n_dofs), np.int32))
        return self._arrays

    def local_basis_specs(self, elem_id: int) -> List[BasisSpec]:
        """domain.rs:253-259"""
        a = self.arrays()
        b0, b1 = int(a["bs_offsets"][elem_id]), int(a["bs_offsets"][elem_id + 1])
        return [BasisSpec(int(a["bs_i"][t]), int(a["bs_j"][t]), BasisDir(int(a["bs_dir"][t])), elem_id, t - b0,
                          int(a["bs_dof"][t])) for t in range(b0, b1)]

    def descendant_basis_specs(self, elem_id: int) -> List[Tuple[int, List[BasisSpec]]]:
        """domain.rs:291-304"""
        return [(d, self.local_basis_specs(d)) for d in self.mesh.descendant_elems(elem_id, False)]

    def ancestor_basis_specs(self, elem_id: int) -> List[Tuple[int, List[BasisSpec]]]:
        """domain.rs:339-352"""
        return [(a, self.local_basis_specs(a)) for a in self.mesh.ancestor_elems(elem_id, False)]

    def dof_basis_specs(self, dof: int) -> List[Tuple[int, int]]:
        """(elem id, index in that elem's basis specs) of a DoF's 1 or 2 BasisSpecs (dof.rs:38-52)."""
        n = C.c_int32()
        eids = (C.c_int32 * 4)()
        idx = (C.c_int32 * 4)()
        check(N.load().fem2d_domain_dof_info(self._h, int(dof), C.byref(n), eids, idx))
        return [(eids[k], idx[k]) for k in range(n.value)]
