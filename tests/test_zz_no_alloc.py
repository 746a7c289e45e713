"""fem2d_assemble_device_noalloc (the assembly with caller-provided outputs and the library's
static per-device scratch array, no device allocation in the call) against fem2d_assemble_device
through the C ABI: the same pattern and the same bits, on two calls in a row, the second with a
smaller Domain; the FP32 and FP32R storage modes and the block path included; the undersized-output and
oversized-Domain errors."""

import ctypes as C

import numpy as np
import pytest

import cases
import fem_2d_amd as F
from fem_2d_amd import _native as N
from fem_2d_amd import workloads as W

pytestmark = pytest.mark.gpu


class DeviceArrays:
    """Caller-owned device buffers (allocated before the no-allocation call, freed after)."""

    def __init__(self, ctx, n_rows, capacity, value_bytes):
        self.ctx, self.lib = ctx, N.load()
        self.ptrs = []
        self.row_ptr = self._alloc(8 * (n_rows + 1))
        self.col = self._alloc(4 * max(capacity, 1))
        self.a = self._alloc(value_bytes * max(capacity, 1))
        self.b = self._alloc(value_bytes * max(capacity, 1))
        self.res = self._alloc(4 * max(capacity, 1))
        self.capacity = capacity

    def _alloc(self, nbytes):
        p = C.c_void_p()
        F.errors.check(self.lib.fem2d_device_malloc(self.ctx._h, int(nbytes), C.byref(p)))
        self.ptrs.append(p)
        return p

    def fetch(self, p, count, dtype):
        out = np.empty(count, dtype=dtype)
        if count:
            F.errors.check(self.lib.fem2d_memcpy_to_host(self.ctx._h, out.ctypes.data_as(C.c_void_p), p,
                                                         int(out.nbytes)))
        return out

    def free(self):
        for p in self.ptrs:
            self.lib.fem2d_device_free(self.ctx._h, p)
        self.ptrs = []


def noalloc(ctx, dd, glq, basis, precision, arrays, capacity=None):
    nnz = C.c_int64(-7)
    cap = arrays.capacity if capacity is None else capacity
    rc = N.load().fem2d_assemble_device_noalloc(ctx._h, dd._h, basis.code, F.CurlCurl.code, F.L2Inner.code,
                                               glq[0], glq[1], int(precision), int(cap), arrays.row_ptr,
                                               arrays.col, arrays.a, arrays.b,
                                               arrays.res if precision == F.FP32R else None, C.byref(nnz))
    return rc, nnz.value


def check_same(ctx, dom, glq, basis, precision, arrays=None):
    """Assemble with both entry points; returns the arrays used (grown when too small)."""
    print(f"[noalloc] n_dofs={dom.num_dofs} basis={basis.__name__} precision={precision}", flush=True)
    dd = F.DeviceDomain(dom, ctx)
    ref = dd.download(dd.assemble(glq, basis, F.CurlCurl, F.L2Inner, precision=precision))
    if arrays is None:
        probe = DeviceArrays(ctx, dom.num_dofs, 0, 8)
        rc, nnz = noalloc(ctx, dd, glq, basis, precision, probe)
        assert rc == N.ERR_INVALID_ARG and nnz == ref.nnz, (rc, nnz, N.last_error())
        assert "capacity" in N.last_error()
        probe.free()
        arrays = DeviceArrays(ctx, dom.num_dofs, nnz, 8)
    rc, nnz = noalloc(ctx, dd, glq, basis, precision, arrays)
    assert rc == N.OK, N.last_error()
    assert nnz == ref.nnz
    row_ptr = arrays.fetch(arrays.row_ptr, dom.num_dofs + 1, np.int64)
    col = arrays.fetch(arrays.col, nnz, np.int32)
    vdt = np.float64 if precision == F.FP64 else np.float32
    a = arrays.fetch(arrays.a, nnz, vdt).astype(np.float64)
    b = arrays.fetch(arrays.b, nnz, vdt).astype(np.float64)
    if precision == F.FP32R:  # the FP32 parts plus their residuals (low half A, high half B)
        r = arrays.fetch(arrays.res, nnz, np.uint32)
        for v, q in ((a, (r & 0xFFFF).astype(np.uint16).view(np.int16)),
                     (b, (r >> 16).astype(np.uint16).view(np.int16))):
            eb = ((v.astype(np.float32).view(np.uint32) >> 23) & 0xFF).astype(np.int64)
            v += np.ldexp(q.astype(np.float64), np.maximum(eb, 1) - 127 - 38)
    np.testing.assert_array_equal(row_ptr, ref.row_ptr)
    np.testing.assert_array_equal(col, ref.col)
    assert np.array_equal(a.view(np.uint64), ref.a_val.view(np.uint64))
    assert np.array_equal(b.view(np.uint64), ref.b_val.view(np.uint64))
    return arrays


@pytest.mark.parametrize("precision", [F.FP64, F.FP32, F.FP32R])
def test_noalloc_matches_assemble_device_twice(precision):
    """First the KOL hp-irregular Domain (block-path row groups included), then, on the same
    buffers and the same scratch array, the smaller 624-DoF problem."""
    ctx = F.Context(0)
    big, bglq = cases.small_hp_irregular()
    arrays = check_same(ctx, big, bglq, F.KOLShapeFn, precision)
    assert ctx.last_timing()["n_large"] > 0
    small, sglq = cases.slepc_problem()
    check_same(ctx, small, sglq, F.HierPoly, precision, arrays)
    arrays.free()
    ctx.close()


def test_noalloc_rejects_domain_beyond_scratch():
    """The 445 380-DoF benchmark Domain needs more work memory than the static array holds: a
    clear error, no device work, and the entry point stays usable afterwards."""
    ctx = F.Context(0)
    w = W.waveguide_64x32()
    dd = F.DeviceDomain(w.domain, ctx)
    arrays = DeviceArrays(ctx, w.domain.num_dofs, 0, 8)
    rc, nnz = noalloc(ctx, dd, w.glq, F.HierPoly, F.FP64, arrays)
    assert rc == N.ERR_UNSUPPORTED, (rc, N.last_error())
    assert "FEM2D_NOALLOC_SCRATCH_BYTES" in N.last_error()
    arrays.free()
    dom, glq = cases.waveguide_small()
    check_same(ctx, dom, glq, F.HierPoly, F.FP64).free()
    ctx.close()
