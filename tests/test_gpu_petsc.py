"""PETSc AIJ binary files written by the library (transpose and byte order on the GPU) against
the oracle's restatement of the reference writer (oracle/petsc.py), byte for byte."""

import ctypes as C
import os
import struct

import numpy as np
import pytest

import fem_2d_amd as F
import cases
from conftest import golden
from oracle import petsc
from fem_2d_amd import _native as N
from fem_2d_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return F.Context(0)


def _expected(g, which):
    vals = g.a_val if which == "a" else g.b_val
    return petsc.aij_binary_bytes(g.n, g.rows(), g.col, vals)


def _outdir(tmp_path):
    os.makedirs(tmp_path / "tmp", exist_ok=True)  # the reference writes into {dir}/tmp/
    return str(tmp_path)


@pytest.mark.parametrize("name", ["nalg_problem", "slepc_problem", "mixed_refinements", "waveguide_small"])
def test_host_gep_files_match_oracle_bytes(ctx, tmp_path, name):
    print(f"[petsc] case={name}", flush=True)
    dom, glq = cases.ALL[name]()
    g = F.galerkin_sample_gep_hcurl(dom, glq, ctx=ctx)
    d = _outdir(tmp_path)
    g.print_to_petsc_binary_files(d, name, ctx=ctx)
    for which in "ab":
        got = open(os.path.join(d, "tmp", f"{name}_{which}.dat"), "rb").read()
        assert got == _expected(g, which)


def test_reference_writer_matrix(ctx, tmp_path):
    """The 10x10 matrix of sparse_matrix.rs:270-287 through the host-CSR entry point."""
    g = F.GEP(10, np.array([0, 2, 2, 2, 3, 3, 3, 3, 3, 3, 4]), np.array([0, 8, 4, 9], dtype=np.int32),
              np.array([2.0, 0.25, 0.25, 10.0]), np.array([2.0, 0.25, 0.25, 10.0]))
    d = _outdir(tmp_path)
    g.print_to_petsc_binary_files(d, "unit", ctx=ctx)
    got = open(os.path.join(d, "tmp", "unit_a.dat"), "rb").read()
    assert got[:4] == bytes([0x00, 0x12, 0x7B, 0x50])
    assert got == petsc.aij_binary_bytes(10, np.array([0, 0, 3, 9]), np.array([0, 8, 4, 9]),
                                         np.array([2.0, 0.25, 0.25, 10.0]))


def test_slepc_eigenpair_from_written_files(ctx, tmp_path):
    """The files carry the GEP of lib.rs:82-119: the reference's SLEPc eigenpair
    (test_eval.dat / test_evec.dat) satisfies A v = lambda B v read back from them."""
    dom, glq = cases.slepc_problem()
    g = F.galerkin_sample_gep_hcurl(dom, glq, ctx=ctx)
    d = _outdir(tmp_path)
    g.print_to_petsc_binary_files(d, "slepc", ctx=ctx)
    mats = []
    for which in "ab":
        n, rp, cols, vals = petsc.read_aij_binary(open(os.path.join(d, "tmp", f"slepc_{which}.dat"), "rb").read())
        M = np.zeros((n, n))
        for i in range(n):
            M[i, cols[rp[i]:rp[i + 1]]] = vals[rp[i]:rp[i + 1]]
        mats.append(M)
    A, B = mats
    lam = struct.unpack(">d", open(golden("test_eval.dat"), "rb").read())[0]
    raw = open(golden("test_evec.dat"), "rb").read()
    nv = struct.unpack(">i", raw[4:8])[0]
    v = np.array(struct.unpack(f">{nv}d", raw[8:8 + 8 * nv]))
    assert np.linalg.norm(A @ v - lam * (B @ v)) / np.linalg.norm(A @ v) < 1e-10


def test_device_gep_files_full_size(ctx, tmp_path):
    """The 64x32 waveguide (25.7 M stored entries, 51 M in the files) from the device-resident GEP."""
    w = W.waveguide_64x32()
    dd = F.DeviceDomain(w.domain, ctx)
    dev = dd.assemble(w.glq, w.basis)
    d = _outdir(tmp_path)
    dd.print_to_petsc_binary_files(dev, d, "wg")
    g = dd.download(dev)
    for which in "ab":
        got = open(os.path.join(d, "tmp", f"wg_{which}.dat"), "rb").read()
        assert got == _expected(g, which), which


def test_missing_directory_is_an_io_error(ctx, tmp_path):
    dom, glq = cases.nalg_problem()
    g = F.galerkin_sample_gep_hcurl(dom, glq, ctx=ctx)
    with pytest.raises(OSError):
        g.print_to_petsc_binary_files(str(tmp_path / "absent"), "x", ctx=ctx)


@pytest.mark.parametrize("precision", [F.FP32, F.FP32R])
def test_device_gep_files_of_mixed_storage(ctx, tmp_path, precision):
    """A GEP stored in a mixed mode: the files hold the stored values widened to f64 (FP32R: the
    FP32 part plus its residual), the same values the download returns."""
    print(f"[petsc] mixed storage precision={precision}", flush=True)
    dom, glq = cases.small_hp_irregular()
    dd = F.DeviceDomain(dom, ctx)
    dev = dd.assemble(glq, F.KOLShapeFn, precision=precision)
    d = _outdir(tmp_path)
    dd.print_to_petsc_binary_files(dev, d, "mixed")
    g = dd.download(dev)
    for which in "ab":
        got = open(os.path.join(d, "tmp", f"mixed_{which}.dat"), "rb").read()
        assert got == _expected(g, which), which
