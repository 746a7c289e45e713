"""ORACLE - test infrastructure only (tests may use it; the product never does).

CPU restatements of the two eigensolves that follow the assembly, on dense numpy matrices:

* nalgebra_solve_gep (nalgebra_solve.rs:14-50, nalgebra 0.30.1): more than MAX_DENSE_SIZE = 1000
  DoFs -> ProblemTooLarge (15-17); B's Cholesky factorisation failing -> FailedToInvertB (19, 47-48);
  C = B^-1 A from the factor's inverse (20-21); SymmetricEigen::new(C) (22), which reads only the
  lower triangle of its argument (C itself is not symmetric); every eigenvalue below 1e-12 in
  magnitude -> SpuriouslyConverged (24-26); else the first eigenvalue with the smallest distance to
  the target and its eigenvector (28-46; nalgebra's vectors have unit 2-norm).
* the symmetric-definite problem A x = lambda B x that slepc_solve_gep (slepc_solve.rs:14-68)
  hands to SLEPc: LAPACK dsygvd through scipy, the eigenpair nearest the target, B-normalised;
  and, for sparse matrices of any size, ARPACK's shift-invert Lanczos around the target (scipy
  eigsh), SLEPc's own method for a target eigenvalue.

The sparse restatement is pinned against the dense one on small GEPs, the SLEPc eigenpair below
and the TE10 cutoff pi^2 of the 1096-DoF waveguide (tests/test_oracle_sparse_eigen.py).

Pinned by the reference's known answers, which the tests check against these functions: the
nalgebra eigenvalue 2.6479657 of lib.rs:45-80 (tolerance 1e-6) - which only the lower-triangle
reading reproduces; the true generalized eigenvalue nearest 2.64 there is 3.618 - and the SLEPc
eigenpair of lib.rs:82-119 (1.4745880937, tolerance 1e-9; the vector in test_evec.dat)."""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla

MAX_DENSE_SIZE = 1000  # nalgebra_solve.rs:6
SPURIOUS = 1e-12  # nalgebra_solve.rs:24


class ProblemTooLarge(Exception):
    pass


class FailedToInvertB(Exception):
    pass


class SpuriouslyConverged(Exception):
    pass


def _nearest(w: np.ndarray, target: float) -> int:
    """Index of the first eigenvalue at the smallest distance from target (nalgebra_solve.rs:28-36)."""
    best, delta_min = 0, np.inf
    for k, e in enumerate(w):
        d = abs(e - target)
        if d < delta_min:
            best, delta_min = k, d
    return best


def nalgebra_solve_gep(A: np.ndarray, B: np.ndarray, target: float):
    """(value, unit vector) as nalgebra_solve_gep computes them; eigenvalues in ascending order."""
    n = A.shape[0]
    if n > MAX_DENSE_SIZE:
        raise ProblemTooLarge(n)
    try:
        L = np.linalg.cholesky(B)
    except np.linalg.LinAlgError as e:
        raise FailedToInvertB(str(e))
    Linv = np.linalg.inv(L)
    C = (Linv.T @ Linv) @ A  # B^-1 A
    lower = np.tril(C) + np.tril(C, -1).T
    w, V = np.linalg.eigh(lower)
    if np.all(np.abs(w) < SPURIOUS):
        raise SpuriouslyConverged()
    k = _nearest(w, target)
    return float(w[k]), V[:, k]


def dense_solve_gep(A: np.ndarray, B: np.ndarray, target: float):
    """(value, B-normalised vector) of A x = lambda B x nearest target; eigenvalues ascending."""
    try:
        w, V = sla.eigh(A, B, driver="gvd")
    except np.linalg.LinAlgError as e:
        raise FailedToInvertB(str(e))
    if np.all(np.abs(w) < SPURIOUS):
        raise SpuriouslyConverged()
    k = _nearest(w, target)
    return float(w[k]), V[:, k]


def sparse_solve_gep(A, B, target: float):
    """(value, B-normalised vector) of A x = lambda B x nearest target for sparse symmetric A, B
    (scipy.sparse matrices): ARPACK's shift-invert Lanczos (scipy eigsh with sigma = target, a
    sparse LU of A - target B), the way slepc_solve_gep's SLEPc run is driven with a target."""
    import scipy.sparse.linalg as spl

    w, V = spl.eigsh(A.tocsc(), k=1, M=B.tocsc(), sigma=target, which="LM", tol=1e-13,
                     v0=np.ones(A.shape[0]))
    x = V[:, 0]
    x = x / np.sqrt(x @ (B @ x))
    return float(w[0]), x


def sparse_eigenvalues_near(A, B, sigma: float, k: int):
    """The k eigenvalues of A x = lambda B x nearest sigma (ARPACK shift-invert), ascending."""
    import scipy.sparse.linalg as spl

    w = spl.eigsh(A.tocsc(), k=k, M=B.tocsc(), sigma=sigma, which="LM", tol=1e-13,
                  v0=np.ones(A.shape[0]), return_eigenvectors=False)
    return np.sort(w)
