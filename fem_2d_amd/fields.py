"""Field solutions over a Domain: the mirror of UniformFieldSpace (reference
src/fem_domain/domain/fields.rs:17-400).

``xy_fields`` evaluates the x and y fields of a solution vector on every leaf Elem with the
library's HIP kernel (fem2d_xy_fields / fem2d_device_xy_fields); the element-wise operations on
stored quantities and the VTK writer are host code, as in the reference."""

from __future__ import annotations

import ctypes as C
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from . import _native as N
from .domain import Domain
from .errors import MissingQuantity, check
from .galerkin import Context, DeviceDomain, HierPoly, default_context


def uniform_range(lo: float, hi: float, n: int) -> np.ndarray:
    """fields.rs:402-405: n points from lo with step (hi - lo) / (n - 1)."""
    step = (hi - lo) / (n - 1) if n != 1 else float("inf") * (hi - lo)
    return np.array([i * step + lo for i in range(n)], dtype=np.float64)


def _leaves(domain: Domain) -> np.ndarray:
    parent = np.asarray(domain.arrays()["parent"])
    has_child = np.zeros(parent.size, dtype=bool)
    has_child[parent[parent >= 0]] = True
    return np.flatnonzero(~has_child)


def xy_fields_array(domain, solution, density: Sequence[int], BSpace=HierPoly, ctx: Optional[Context] = None):
    """(x, y) field arrays of shape [leaves, density_u, density_v] (leaves in ascending Elem id).

    ``domain`` is a :class:`Domain` (the view is uploaded for the call) or a
    :class:`DeviceDomain`; the evaluation runs on the GPU."""
    ctx = ctx or default_context()
    du, dv = int(density[0]), int(density[1])
    lib = N.load()
    sol = np.ascontiguousarray(solution, dtype=np.float64)
    if isinstance(domain, DeviceDomain):
        n_leaves = C.c_int64()
        check(lib.fem2d_device_domain_num_leaves(domain._h, C.byref(n_leaves)))
        total = n_leaves.value * max(du, 0) * max(dv, 0)
        bufs = [C.c_void_p() for _ in range(3)]
        try:
            for b, nbytes in zip(bufs, (8 * sol.size, 8 * total, 8 * total)):
                check(lib.fem2d_device_malloc(ctx._h, nbytes, C.byref(b)))
            check(lib.fem2d_memcpy_to_device(ctx._h, bufs[0], sol.ctypes.data, 8 * sol.size))
            check(lib.fem2d_device_xy_fields(ctx._h, domain._h, BSpace.code, du, dv, bufs[0], sol.size, bufs[1],
                                             bufs[2]))
            x = np.empty(total)
            y = np.empty(total)
            check(lib.fem2d_memcpy_to_host(ctx._h, x.ctypes.data, bufs[1], 8 * total))
            check(lib.fem2d_memcpy_to_host(ctx._h, y.ctypes.data, bufs[2], 8 * total))
        finally:
            for b in bufs:
                if b.value:
                    lib.fem2d_device_free(ctx._h, b)
        return x.reshape(n_leaves.value, max(du, 0), max(dv, 0)), y.reshape(n_leaves.value, max(du, 0), max(dv, 0))
    n_leaves = _leaves(domain).size
    total = n_leaves * max(du, 0) * max(dv, 0)
    x = np.zeros(max(total, 1))
    y = np.zeros(max(total, 1))
    P = C.POINTER(C.c_double)
    check(lib.fem2d_xy_fields(ctx._h, C.byref(domain.view()), BSpace.code, du, dv, sol.ctypes.data_as(P), sol.size,
                              x.ctypes.data_as(P), y.ctypes.data_as(P)))
    return x[:total].reshape(n_leaves, max(du, 0), max(dv, 0)), y[:total].reshape(n_leaves, max(du, 0), max(dv, 0))


class UniformFieldSpace:
    """UniformFieldSpace (fields.rs:17-38): named field quantities on a uniform point grid of
    ``densities`` points over every leaf Elem."""

    def __init__(self, domain: Domain, densities: Sequence[int], ctx: Optional[Context] = None):
        self.domain = domain
        self.densities = [int(densities[0]), int(densities[1])]
        self.parametric_points = [uniform_range(-1.0, 1.0, self.densities[0]),
                                  uniform_range(-1.0, 1.0, self.densities[1])]
        self.quantities: Dict[str, Dict[int, np.ndarray]] = {}
        self.ctx = ctx

    def xy_fields(self, vector_name: str, solution, BSpace=HierPoly) -> List[str]:
        """fields.rs:63-127: stores {vector_name}_x and {vector_name}_y and returns their names."""
        x, y = xy_fields_array(self.domain, solution, self.densities, BSpace, self.ctx)
        leaves = _leaves(self.domain)
        xn, yn = f"{vector_name}_x", f"{vector_name}_y"
        self.quantities[xn] = {int(e): x[k] for k, e in enumerate(leaves)}
        self.quantities[yn] = {int(e): y[k] for k, e in enumerate(leaves)}
        return [xn, yn]

    def _get(self, name: str) -> Dict[int, np.ndarray]:
        if name not in self.quantities:
            raise MissingQuantity(f"Missing quantity '{name}', Cannot apply operation!")
        return self.quantities[name]

    def map_to_quantity(self, name: str, result_name: str, operator: Callable[[float], float]) -> None:
        """fields.rs:262-286: result = operator applied to every value of ``name``."""
        q = self._get(name)
        self.quantities[result_name] = {e: np.vectorize(operator, otypes=[np.float64])(v) for e, v in q.items()}

    def expression_2arg(self, operand_names: Sequence[str], result_name: str,
                        expression: Callable[[float, float], float]) -> None:
        """fields.rs:292-331: result = expression(a, b) point by point (the leaves of the first operand)."""
        qa, qb = self._get(operand_names[0]), self._get(operand_names[1])
        f = np.vectorize(expression, otypes=[np.float64])
        self.quantities[result_name] = {ea: f(va, vb) for (ea, va), vb in zip(sorted(qa.items()),
                                                                               [qb[e] for e in sorted(qb)])}

    def print_all_to_vtk(self, path: str) -> None:
        """fields.rs:135-139"""
        self.print_quantities_to_vkt(path, list(self.quantities))

    def print_quantities_to_vkt(self, path: str, quantity_names: Sequence[str]) -> None:
        """fields.rs:145-245: legacy-VTK unstructured grid of the leaf grids with the quantities as
        point scalars (the header's time stamp is the only content that differs run to run)."""
        import time

        for qn in quantity_names:
            self._get(qn)
        nx, ny = self.densities
        leaves = _leaves(self.domain)
        mesh = self.domain.mesh
        n_points = nx * ny * leaves.size
        out = ["# vtk DataFile Version 3.0\n", f"# File generated by fem_2d on: {time.time():.9f}s\n\n", "ASCII\n",
               "DATASET UNSTRUCTURED_GRID\n", f"\nPOINTS {n_points} double\n"]
        for e in leaves:
            pts = mesh.elem_points(int(e))
            (x0, y0), (x1, y1) = pts[0], pts[3]
            for x in uniform_range(x0, x1, nx):
                for y in uniform_range(y0, y1, ny):
                    out.append(f"{x:.10f} {y:.10f} 0.0\n")
        n_cells = (nx - 1) * (ny - 1) * leaves.size
        out.append(f"\nCELLS {n_cells} {5 * n_cells}\n")
        for k in range(leaves.size):
            for i in range(nx - 1):
                for j in range(ny - 1):
                    p0 = nx * i + j + (nx * ny) * k
                    out.append(f"4\t{p0}\t{p0 + 1}\t{p0 + nx + 1}\t{p0 + nx}\n")
        out.append(f"\nCELL_TYPES {n_cells}\n")
        out.append(" 9" * n_cells + "\n")
        out.append(f"POINT_DATA {n_points}\n")
        for qn in quantity_names:
            # the values follow each other with a trailing space and no line break, as in
            # FieldQuantity::write_vtk_quantity (fields.rs:367-381)
            out.append(f"SCALARS {qn} double 1 \nLOOKUP_TABLE default\n")
            q = self.quantities[qn]
            out.append("".join(f"{v:.15f} " for e in sorted(q) for v in q[e].ravel()))
        out.append("\n")
        with open(path, "w") as f:
            f.write("".join(out))
