"""ORACLE - test infrastructure only.

ctypes wrapper of oracle/_build/liboracle.so, the CPU restatement of the
reference's Galerkin assembly (see oracle.cpp). Only the tests, the smoke check
and the benchmark's CPU-baseline leg import this module.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


class OracleGep(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("nnz", C.c_int64),
        ("pairs", C.c_int64),
        ("row_ptr", C.POINTER(C.c_int64)),
        ("col", C.POINTER(C.c_int32)),
        ("a_val", C.POINTER(C.c_double)),
        ("b_val", C.POINTER(C.c_double)),
    ]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `python -m fem_2d_amd.build_native`")
        lib = C.CDLL(LIB_PATH)
        lib.oracle_gauss_quadrature_points.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double),
                                                       C.POINTER(C.c_double)]
        lib.oracle_shape_tables.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)] + \
            [C.POINTER(C.c_double)] * 4
        lib.oracle_default_ngq.argtypes = [C.c_int]
        lib.oracle_xy_fields.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_int64,
                                         C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.oracle_default_ngq.restype = C.c_int
        lib.oracle_assemble.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int64, C.POINTER(C.c_uint8), C.POINTER(OracleGep)]
        lib.oracle_assemble.restype = C.c_int
        lib.oracle_gep_free.argtypes = [C.POINTER(OracleGep)]
        _lib = lib
    return _lib


def gauss_quadrature_points(n: int, include_endpoints: bool = False):
    """glq.rs:179-222 (Golub-Welsch)."""
    m = n + (2 if include_endpoints else 0)
    p = (C.c_double * m)()
    w = (C.c_double * m)()
    load().oracle_gauss_quadrature_points(n, int(include_endpoints), p, w)
    return np.array(p[:]), np.array(w[:])


def shape_tables(basis: int, n_max: int, points):
    """N, N', T, T' tables [order][point] of HierPoly (basis 0) or the Legendre space (basis 1)."""
    x = np.ascontiguousarray(points, dtype=np.float64)
    np_ = x.size
    outs = [np.zeros((n_max + 1) * np_) for _ in range(4)]
    ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    load().oracle_shape_tables(basis, n_max, np_, ptr(x), *[ptr(o) for o in outs])
    return [o.reshape(n_max + 1, np_) for o in outs]


def default_ngq(max_order: int) -> int:
    return load().oracle_default_ngq(max_order)


class OracleResult:
    def __init__(self, n, row_ptr, col, a_val, b_val, pairs):
        self.n = n
        self.row_ptr = row_ptr
        self.col = col
        self.a_val = a_val
        self.b_val = b_val
        self.pairs = pairs

    @property
    def nnz(self):
        return int(self.col.size)

    def rows(self):
        return np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.row_ptr))

    def to_dense(self):
        A = np.zeros((self.n, self.n))
        B = np.zeros((self.n, self.n))
        r = self.rows()
        A[r, self.col] = self.a_val
        A[self.col, r] = self.a_val
        B[r, self.col] = self.b_val
        B[self.col, r] = self.b_val
        return A, B


def assemble(view, basis: int = 0, a_integral: int = 0, b_integral: int = 1, glq=(8, 8), n_threads: int = 1,
             elem_limit: int = -1, row_mask=None):
    """Reference-algorithm assembly of a fem2d_domain_view (ctypes struct). Returns
    (status, OracleResult or None); status follows galerkin.rs's error order.
    row_mask (bool array over DoFs, optional): compute only the rows it selects (complete
    and exact), which keeps full-size domains checkable in seconds."""
    g = OracleGep()
    gu, gv = (0, 0) if glq is None else glq
    mask = None
    if row_mask is not None:
        mask = np.ascontiguousarray(np.asarray(row_mask, dtype=np.uint8))
        assert mask.size == view.n_dofs
    rc = load().oracle_assemble(C.addressof(view), basis, a_integral, b_integral, gu, gv, n_threads, elem_limit,
                                None if mask is None else mask.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(g))
    if rc != 0:
        return rc, None
    try:
        n, nnz = g.n, g.nnz

        def take(ptr, count, dt):
            if count == 0:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dt, copy=True)

        res = OracleResult(n, take(g.row_ptr, n + 1, np.int64), take(g.col, nnz, np.int32),
                           take(g.a_val, nnz, np.float64), take(g.b_val, nnz, np.float64), g.pairs)
    finally:
        load().oracle_gep_free(C.byref(g))
    return 0, res


def xy_fields(view, basis: int, density, solution):
    """UniformFieldSpace::xy_fields (fields.rs:63-127) on the CPU: (rc, x, y) with x, y of shape
    [leaves, density_u, density_v]; rc 1 = MismatchedSolutionSize, 2 = the reference's index panic."""
    import numpy as np

    parent = np.ctypeslib.as_array(view.parent, shape=(view.n_elems,)) if view.n_elems else np.zeros(0, np.int32)
    has_child = np.zeros(view.n_elems, dtype=bool)
    has_child[parent[parent >= 0]] = True
    n_leaves = int(np.count_nonzero(~has_child))
    du, dv = int(density[0]), int(density[1])
    total = n_leaves * max(du, 0) * max(dv, 0)
    x = np.zeros(max(total, 1))
    y = np.zeros(max(total, 1))
    sol = np.ascontiguousarray(solution, dtype=np.float64)
    P = C.POINTER(C.c_double)
    rc = load().oracle_xy_fields(C.addressof(view), int(basis), du, dv, sol.ctypes.data_as(P),
                                 sol.size, x.ctypes.data_as(P), y.ctypes.data_as(P))
    return rc, x[:total].reshape(n_leaves, max(du, 0), max(dv, 0)), y[:total].reshape(n_leaves, max(du, 0),
                                                                                      max(dv, 0))
