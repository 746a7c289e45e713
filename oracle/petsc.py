"""ORACLE - test infrastructure only (tests and the smoke check may use it; the product never does).

Bytes of a PETSc AIJ binary matrix file as the reference writes them: SparseMatrix ->
AIJMatrixBinary (sparse_matrix.rs:184-223) takes the upper-triangular map together with its
transpose, keyed and ordered by [row, col], counts entries per row (a diagonal entry once, an
off-diagonal entry in both its rows); print_to_petsc_binary_file (sparse_matrix.rs:225-265)
writes the header bytes 00 12 7B 50 (1211216, PETSc's MAT_FILE_CLASSID), rows, columns and nnz
as big-endian u32, then the row counts (u32), column indices (u32) and values (f64), all
big-endian. GEP::print_to_petsc_binary_files (linalg.rs:44-52) writes A and B this way to
{dir}/tmp/{prefix}_a.dat and {dir}/tmp/{prefix}_b.dat."""

from __future__ import annotations

import struct

import numpy as np

MAT_FILE_CLASSID = 1211216


def aij_binary_bytes(n: int, rows, cols, vals) -> bytes:
    """File bytes of the symmetric matrix whose upper triangle is (rows, cols, vals), r <= c."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    vals = np.asarray(vals, dtype=np.float64)
    off = rows != cols
    R = np.concatenate([rows, cols[off]])
    C = np.concatenate([cols, rows[off]])
    V = np.concatenate([vals, vals[off]])
    order = np.lexsort((C, R))
    counts = np.bincount(R, minlength=n)
    head = struct.pack(">IIII", MAT_FILE_CLASSID, n, n, R.size)
    return (head + counts.astype(">u4").tobytes() + C[order].astype(">u4").tobytes()
            + V[order].astype(">f8").tobytes())


def read_aij_binary(data: bytes):
    """(n, row_ptr, cols, vals) of a PETSc AIJ binary file (full matrix, CSR)."""
    classid, m, n, nz = struct.unpack(">IIII", data[:16])
    assert classid == MAT_FILE_CLASSID and m == n
    o = 16
    counts = np.frombuffer(data, dtype=">u4", count=m, offset=o).astype(np.int64)
    o += 4 * m
    cols = np.frombuffer(data, dtype=">u4", count=nz, offset=o).astype(np.int64)
    o += 4 * nz
    vals = np.frombuffer(data, dtype=">f8", count=nz, offset=o).astype(np.float64)
    assert o + 8 * nz == len(data)
    row_ptr = np.concatenate([[0], np.cumsum(counts)])
    return n, row_ptr, cols, vals
