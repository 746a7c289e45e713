"""Pins of the oracle's sparse shift-invert restatement (oracle.eigen.sparse_solve_gep, ARPACK
around the target), which checks the GPU's sparse eigensolve on GEPs beyond dense sizes:

* the dense LAPACK restatement (dense_solve_gep) on small GEPs of both shape-function spaces
  and several targets: the same eigenvalue and, up to sign, the same B-normalised vector;
* the SLEPc eigenpair of lib.rs:82-119 (test_eval.dat / test_evec.dat), which slepc_solve_gep
  returns for that problem;
* the analytic TE10 cutoff pi^2 of the air-filled 1 x 0.5 guide (1096-DoF waveguide);
* a disconnected GEP (two independent problems), whose eigenpairs belong to one of them."""

import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp

import cases
import oracle
from oracle import eigen as OE
from test_oracle_known_answers import _golden_eigenpair


def _sparse(res):
    rows = res.rows()
    off = rows != res.col
    r = np.concatenate([rows, res.col[off]])
    c = np.concatenate([res.col, rows[off]])
    A = sp.csr_matrix((np.concatenate([res.a_val, res.a_val[off]]), (r, c)), shape=(res.n, res.n))
    B = sp.csr_matrix((np.concatenate([res.b_val, res.b_val[off]]), (r, c)), shape=(res.n, res.n))
    return A, B


def _assemble(name, basis=0):
    dom, glq = cases.ALL[name]()
    rc, res = oracle.assemble(dom.view(), basis=basis, glq=glq, n_threads=8)
    assert rc == 0
    return res


@pytest.mark.parametrize("name,basis,target", [("slepc_problem", 0, 1.475), ("slepc_problem", 0, 7.0),
                                               ("mixed_refinements", 1, 3.0), ("waveguide_small", 0, 45.0)])
def test_sparse_restatement_matches_dense(name, basis, target):
    res = _assemble(name, basis)
    A, B = _sparse(res)
    lam_s, x_s = OE.sparse_solve_gep(A, B, target)
    lam_d, x_d = OE.dense_solve_gep(A.toarray(), B.toarray(), target)
    assert abs(lam_s - lam_d) <= 1e-10 * abs(lam_d)
    assert abs(x_s @ (B @ x_s) - 1.0) < 1e-10
    assert np.linalg.norm(A @ x_s - lam_s * (B @ x_s)) <= 1e-9 * np.linalg.norm(A @ x_s)
    w = sl.eigh(A.toarray(), B.toarray(), eigvals_only=True)
    if np.sum(np.abs(w - lam_d) <= 1e-8 * abs(lam_d)) == 1:  # a simple eigenvalue: one vector up to sign
        cos = abs(x_s @ (B @ x_d)) / np.sqrt((x_s @ (B @ x_s)) * (x_d @ (B @ x_d)))
        assert 1.0 - cos < 1e-9


def test_sparse_restatement_reproduces_slepc_eigenpair():
    res = _assemble("slepc_problem")
    A, B = _sparse(res)
    lam_ref, v = _golden_eigenpair()
    lam, x = OE.sparse_solve_gep(A, B, 1.475)
    assert abs(lam - 1.4745880937) < 1e-9  # lib.rs:104
    assert abs(lam - lam_ref) < 1e-10
    cos = abs(x @ (B @ v)) / np.sqrt((x @ (B @ x)) * (v @ (B @ v)))
    assert 1.0 - cos < 1e-9


def test_sparse_restatement_te10():
    res = _assemble("waveguide_small")
    A, B = _sparse(res)
    lam, _ = OE.sparse_solve_gep(A, B, 10.0)
    assert abs(lam - np.pi ** 2) / np.pi ** 2 < 1e-6


def test_sparse_restatement_on_disconnected_instances():
    """Two independent problems in one GEP (block diagonal): the pair nearest the target is the
    1096-DoF guide's TE10, supported on that block only."""
    A1, B1 = _sparse(_assemble("slepc_problem"))
    A2, B2 = _sparse(_assemble("waveguide_small"))
    A, B = sp.block_diag((A1, A2)).tocsr(), sp.block_diag((B1, B2)).tocsr()
    lam, x = OE.sparse_solve_gep(A, B, 9.5)
    lam_d, _ = OE.dense_solve_gep(A.toarray(), B.toarray(), 9.5)
    assert abs(lam - lam_d) <= 1e-10 * lam_d
    assert abs(lam - np.pi ** 2) / np.pi ** 2 < 1e-6
    assert np.abs(x[:A1.shape[0]]).max() < 1e-8 * np.abs(x).max()
