"""Generalized eigensolves of an assembled GEP on the GPU: the drop-ins for
``nalgebra_solve_gep`` (reference src/fem_problem/linalg/nalgebra_solve.rs:14-50) and
``slepc_solve_gep`` (linalg/slepc_solve.rs:14-68): dense up to the dense solver's size limit,
sparse (block-tridiagonal level-set ordering, dense Schur blocks) beyond.

All run in libfem2d_amd.so (rocSOLVER / rocBLAS on the context's stream after the library's
own kernels); there is no CPU fallback. Each takes either a host :class:`GEP` or a
device-resident GEP as :meth:`DeviceDomain.assemble` returns it, which is solved in HBM without
a download."""

from __future__ import annotations

import ctypes as C
from typing import List, Optional, Union

import numpy as np

from . import _native as N
from .errors import check
from .galerkin import GEP, Context, default_context

MAX_DENSE_SIZE = N.NALGEBRA_MAX_DENSE_SIZE  # nalgebra_solve.rs:6
DENSE_GEP_MAX_SIZE = N.DENSE_GEP_MAX_SIZE


class EigenPair:
    """EigenPair (linalg.rs:84-97)."""

    def __init__(self, value: float, vector: np.ndarray):
        self.value = float(value)
        self.vector = vector

    def normalized_eigenvector(self) -> List[float]:
        """The vector scaled to unit 2-norm (linalg.rs:92-96)."""
        return list(self.vector / np.sqrt(np.sum(self.vector ** 2)))


def _solve(host_fn: str, device_fn: str, gep: Union[GEP, N.GepDevice], target: float,
           ctx: Optional[Context]) -> EigenPair:
    ctx = ctx or default_context()
    lib = N.load()
    out = N.EigenPair()
    if isinstance(gep, N.GepDevice):
        check(getattr(lib, device_fn)(ctx._h, C.byref(gep), float(target), C.byref(out)))
    else:
        csr = gep._as_csr()
        check(getattr(lib, host_fn)(ctx._h, C.byref(csr), float(target), C.byref(out)))
    try:
        vec = np.ctypeslib.as_array(out.vector, shape=(out.n,)).copy() if out.n else np.zeros(0)
        return EigenPair(out.value, vec)
    finally:
        lib.fem2d_eigenpair_free(C.byref(out))


def nalgebra_solve_gep(gep: Union[GEP, N.GepDevice], target_eigenvalue: float,
                       ctx: Optional[Context] = None) -> EigenPair:
    """nalgebra_solve_gep (nalgebra_solve.rs:14-50) with its arithmetic: at most 1000 DoFs
    (ProblemTooLarge), B = L L^T (FailedToInvertB), C = B^-1 A, the symmetric eigenproblem of C's
    lower triangle, SpuriouslyConverged if every eigenvalue is below 1e-12 in magnitude, else the
    eigenpair nearest ``target_eigenvalue`` with a unit-2-norm vector."""
    return _solve("fem2d_nalgebra_solve_gep", "fem2d_gep_device_nalgebra_solve", gep, target_eigenvalue, ctx)


def solve_gep_dense(gep: Union[GEP, N.GepDevice], target_eigenvalue: float,
                    ctx: Optional[Context] = None) -> EigenPair:
    """The symmetric-definite problem A x = lambda B x, the one slepc_solve_gep (slepc_solve.rs:14-68)
    hands to SLEPc, for up to DENSE_GEP_MAX_SIZE DoFs: the eigenpair nearest ``target_eigenvalue``
    (shift-invert inverse iteration; the whole spectrum when that is inconclusive), its vector
    B-normalised. Raises ProblemTooLarge, FailedToInvertB, SpuriouslyConverged or FailedToConverge."""
    return _solve("fem2d_solve_gep_dense", "fem2d_gep_device_solve_dense", gep, target_eigenvalue, ctx)


def solve_gep_sparse(gep: Union[GEP, N.GepDevice], target_eigenvalue: float,
                     ctx: Optional[Context] = None) -> EigenPair:
    """The eigenpair of A x = lambda B x nearest ``target_eigenvalue`` for GEPs of any size whose
    factor fits the device memory, standing in for slepc_solve_gep (slepc_solve.rs:14-68):
    shift-invert inverse iteration on a block-tridiagonal (breadth-first level set) ordering of
    A - target B, its dense Schur-complement blocks inverted on the GPU; the vector B-normalised.
    Raises ProblemTooLarge or FailedToConverge."""
    return _solve("fem2d_solve_gep_sparse", "fem2d_gep_device_solve_sparse", gep, target_eigenvalue, ctx)


# the dense solve is used up to this many DoFs by slepc_solve_gep, the sparse one beyond (at 4496
# DoFs the dense solve takes 106-154 ms, its LU panels and substitutions latency-bound, the sparse
# one 36 ms; at 1096 both take ~15 ms)
SLEPC_DENSE_LIMIT = 2048


def slepc_solve_gep(gep: Union[GEP, N.GepDevice], target_eigenvalue: float,
                    ctx: Optional[Context] = None) -> EigenPair:
    """slepc_solve_gep (slepc_solve.rs:14-68) without the SLEPc process: the eigenpair nearest
    ``target_eigenvalue``, from the dense solve up to SLEPC_DENSE_LIMIT DoFs and from the sparse
    one beyond."""
    if gep.n <= SLEPC_DENSE_LIMIT:
        return solve_gep_dense(gep, target_eigenvalue, ctx)
    return solve_gep_sparse(gep, target_eigenvalue, ctx)
