"""The oracle's PETSc AIJ binary writer (oracle/petsc.py) pinned by a known answer.

The matrix is the one of the reference's own writer test (sparse_matrix.rs:270-287): inserts
[0,0]=1 twice, [9,9]=10, [4,3]=0.25 and [0,8], [8,0]=0.125 into a 10x10 SparseMatrix, i.e. the
upper map {[0,0]: 2, [0,8]: 0.25, [3,4]: 0.25, [9,9]: 10}. The expected bytes are written out
by hand from the format (sparse_matrix.rs:225-265): header 00 12 7B 50 (the literal b"\\0\\x12{P"),
rows, columns, nnz, then row counts, columns and values of the full symmetric matrix in row order,
all big-endian."""

import struct

import numpy as np

from oracle import petsc


def test_reference_writer_matrix_known_bytes():
    rows = np.array([0, 0, 3, 9])
    cols = np.array([0, 8, 4, 9])
    vals = np.array([2.0, 0.25, 0.25, 10.0])
    got = petsc.aij_binary_bytes(10, rows, cols, vals)
    expected = (bytes([0x00, 0x12, 0x7B, 0x50]) + struct.pack(">III", 10, 10, 6)
                + struct.pack(">10I", 2, 0, 0, 1, 1, 0, 0, 0, 1, 1)
                + struct.pack(">6I", 0, 8, 4, 3, 0, 9)
                + struct.pack(">6d", 2.0, 0.25, 0.25, 0.25, 0.25, 10.0))
    assert got == expected


def test_read_back_is_the_full_symmetric_matrix():
    rng = np.random.default_rng(3)
    n = 40
    dense = np.triu(rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.2))
    np.fill_diagonal(dense, 1.0 + np.arange(n))
    r, c = np.nonzero(dense)
    data = petsc.aij_binary_bytes(n, r, c, dense[r, c])
    m, row_ptr, cols, vals = petsc.read_aij_binary(data)
    full = np.zeros((n, n))
    for i in range(n):
        full[i, cols[row_ptr[i]:row_ptr[i + 1]]] = vals[row_ptr[i]:row_ptr[i + 1]]
        assert np.all(np.diff(cols[row_ptr[i]:row_ptr[i + 1]]) > 0)
    np.testing.assert_array_equal(full, dense + np.triu(dense, 1).T)
