"""Galerkin GEP assembly on the GPU: the drop-in for
``galerkin_sample_gep_hcurl::<BSpace, AI, BI>(&domain, glq_grid_dim)``
(reference src/fem_problem/galerkin.rs:33-187).

The assembly always runs in libfem2d_amd.so's HIP kernels; there is no CPU
fallback. Without a usable GPU, :class:`Context` raises NativeError."""

from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _native as N
from .domain import ContinuityCondition, Domain
from .errors import check

MIN_GLQ_ORDER = 4  # galerkin.rs:13

# precision of the assembled values (fem2d_precision): FP64 is the reference's; the
# mixed modes keep the FP64 arithmetic and round at the store: FP32 stores A and B as
# FP32 (12 B per entry), FP32R as FP32 plus a 16-bit residual each (16 B per entry,
# relative entry error <= 2^-39)
FP64 = N.PRECISION_FP64
FP32 = N.PRECISION_FP32
FP32R = N.PRECISION_FP32R


# -- generic arguments of galerkin_sample_gep_hcurl ------------------------
class HierCurlBasisFnSpace:
    """Marker base of the built-in shape-function spaces (basis.rs:19-44)."""

    code = -1


ShapeFn = HierCurlBasisFnSpace


class HierPoly(HierCurlBasisFnSpace):
    """Monomial N and "bubble" T functions (hierarchical_basis_fns.rs:15-198)."""

    code = 0


class KOLShapeFn(HierCurlBasisFnSpace):
    """Legendre-based maximally-orthogonal space (hierarchical_basis_fns.rs:201-647), with the
    order-10 coefficient row corrected so that T_n(+-1) = 0 for every n >= 2."""

    code = 1


class HierCurlIntegral:
    """Marker base of the built-in integrals (integration.rs:59-90)."""

    code = -1


class CurlCurl(HierCurlIntegral):
    """<curl u, curl v> / mu_r (integrals.rs:13-92)"""

    code = 0


class L2Inner(HierCurlIntegral):
    """eps_r <u, v> (integrals.rs:279-354)"""

    code = 1


class Context:
    """A GPU context (HIP stream, work buffers, timers) on one device."""

    def __init__(self, device: int = 0):
        self._lib = N.load()
        h = C.c_void_p()
        check(self._lib.fem2d_ctx_create(int(device), C.byref(h)))
        self._h = h
        self.device = device

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.fem2d_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graphs(self, enable: bool) -> None:
        """Record each assembly once as a HIP graph and replay it (default), or launch kernel by kernel."""
        check(self._lib.fem2d_ctx_set_graphs(self._h, 1 if enable else 0))

    def set_autotune(self, enable: bool) -> None:
        """Time the entry fill's chunk sizes on each Domain's first assembly and keep the fastest
        (default), or keep the default chunk size."""
        check(self._lib.fem2d_ctx_set_autotune(self._h, 1 if enable else 0))

    def set_fill_chunk(self, chunk: int) -> None:
        """Forces the fill's chunk size (a power of two from 64 to 4096 CSR entries) for later
        assemblies, with no tuning; 0 returns to the tuned or default size."""
        check(self._lib.fem2d_ctx_set_fill_chunk(self._h, int(chunk)))

    def last_eigen_stats(self) -> dict:
        """Statistics of the last sparse eigensolve on this context."""
        s = N.EigenStats()
        check(self._lib.fem2d_ctx_last_eigen_stats(self._h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in N.EigenStats._fields_}

    def synchronize(self) -> None:
        """Wait for all work on the context's device."""
        check(self._lib.fem2d_ctx_synchronize(self._h))

    @property
    def stream(self) -> int:
        return int(self._lib.fem2d_ctx_stream(self._h) or 0)

    def last_timing(self) -> dict:
        t = N.Timing()
        check(self._lib.fem2d_ctx_last_timing(self._h, C.byref(t)))
        return dict(total_ms=t.total_ms, prep_ms=t.prep_ms, symbolic_ms=t.symbolic_ms, fill_ms=t.fill_ms,
                    n_groups=t.n_groups, count_kernel_ms=t.count_kernel_ms, fill_kernel_ms=t.fill_kernel_ms,
                    nnz=t.nnz, n_large=t.n_large, fill_chunk=t.fill_chunk, fill_lib_lds=t.fill_lib_lds)


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class GEP:
    """Generalized eigenproblem A u = lambda B u (linalg.rs:29-57) as upper-triangular CSR.

    ``row_ptr``, ``col``, ``a_val``, ``b_val`` hold the stored upper triangle (diagonal
    included, explicit zeros kept); A and B share the pattern."""

    def __init__(self, n: int, row_ptr, col, a_val, b_val):
        self.n = int(n)
        self.row_ptr = row_ptr
        self.col = col
        self.a_val = a_val
        self.b_val = b_val

    @classmethod
    def _from_csr(cls, g: N.GepCsr) -> "GEP":
        n, nnz = g.n, g.nnz

        def take(ptr, count, dt):
            if count == 0:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dt, copy=True)

        return cls(n, take(g.row_ptr, n + 1, np.int64), take(g.col, nnz, np.int32), take(g.a_val, nnz, np.float64),
                   take(g.b_val, nnz, np.float64))

    @property
    def nnz(self) -> int:
        """Stored upper-triangular entries (len of the reference's BTreeMap)."""
        return int(self.col.size)

    def rows(self) -> np.ndarray:
        return np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.row_ptr))

    def num_entries(self) -> int:
        """SparseMatrix::num_entries (sparse_matrix.rs:32-35): full symmetric count."""
        return 2 * self.nnz - int(np.count_nonzero(self.rows() == self.col))

    def iter_upper_tri(self, which: str = "a"):
        vals = self.a_val if which == "a" else self.b_val
        rows = self.rows()
        for r, c, v in zip(rows, self.col, vals):
            yield (int(r), int(c)), float(v)

    def _as_csr(self) -> N.GepCsr:
        """A borrowed fem2d_gep_csr over this GEP's arrays (valid while self lives)."""
        self._keep = [np.ascontiguousarray(x) for x in (self.row_ptr, self.col, self.a_val, self.b_val)]
        rp, col, a, b = self._keep
        g = N.GepCsr()
        g.n, g.nnz = self.n, self.nnz
        g.row_ptr = rp.ctypes.data_as(C.POINTER(C.c_int64))
        g.col = col.ctypes.data_as(C.POINTER(C.c_int32))
        g.a_val = a.ctypes.data_as(C.POINTER(C.c_double))
        g.b_val = b.ctypes.data_as(C.POINTER(C.c_double))
        return g

    def print_to_petsc_binary_files(self, dir: str, prefix: str, ctx: Optional["Context"] = None) -> None:
        """GEP::print_to_petsc_binary_files (linalg.rs:44-52): writes {dir}/tmp/{prefix}_a.dat and
        {dir}/tmp/{prefix}_b.dat as PETSc AIJ binary matrices (both triangles, big-endian)."""
        ctx = ctx or default_context()
        g = self._as_csr()
        check(N.load().fem2d_gep_csr_print_to_petsc_binary_files(ctx._h, C.byref(g), str(dir).encode(),
                                                                 str(prefix).encode()))

    def to_dense(self):
        """Full symmetric dense (A, B) (sparse_matrix.rs:168-182)."""
        A = np.zeros((self.n, self.n))
        B = np.zeros((self.n, self.n))
        rows = self.rows()
        A[rows, self.col] = self.a_val
        A[self.col, rows] = self.a_val
        B[rows, self.col] = self.b_val
        B[self.col, rows] = self.b_val
        return A, B

    def to_scipy(self):
        """Full symmetric scipy.sparse CSR matrices (A, B)."""
        import scipy.sparse as sp

        rows = self.rows()
        off = rows != self.col
        r = np.concatenate([rows, self.col[off]])
        c = np.concatenate([self.col, rows[off]])
        A = sp.csr_matrix((np.concatenate([self.a_val, self.a_val[off]]), (r, c)), shape=(self.n, self.n))
        B = sp.csr_matrix((np.concatenate([self.b_val, self.b_val[off]]), (r, c)), shape=(self.n, self.n))
        return A, B


def _glq(glq_grid_dim: Optional[Sequence[int]]):
    if glq_grid_dim is None:
        return 0, 0
    u, v = glq_grid_dim
    return int(u), int(v)


def galerkin_sample_gep_hcurl(domain: Domain, glq_grid_dim: Optional[Sequence[int]] = None,
                              BSpace=HierPoly, AI=CurlCurl, BI=L2Inner, ctx: Optional[Context] = None,
                              precision: int = FP64) -> GEP:
    """Fill A (``AI``) and B (``BI``) over the Domain's basis space on the GPU (galerkin.rs:33-187).

    Raises WrongContinuityCondition, EmptyDOFSet or InvalidGLQSettings in the reference's order.
    ``precision`` FP32 stores A and B rounded to FP32, FP32R as FP32 plus a 16-bit residual (returned
    widened to float64)."""
    ctx = ctx or default_context()
    gu, gv = _glq(glq_grid_dim)
    out = N.GepCsr()
    lib = N.load()
    check(lib.fem2d_galerkin_sample_gep_hcurl(ctx._h, C.byref(domain.view()), BSpace.code, AI.code, BI.code, gu, gv,
                                              int(precision), C.byref(out)))
    try:
        return GEP._from_csr(out)
    finally:
        lib.fem2d_gep_csr_free(C.byref(out))


class DeviceDomain:
    """A Domain held in HBM for repeated device-resident assembly: uploaded from a host Domain,
    or built on the GPU from a Mesh (:meth:`from_mesh`)."""

    def __init__(self, domain: Optional[Domain], ctx: Optional[Context] = None, _handle=None):
        self.ctx = ctx or default_context()
        self._lib = N.load()
        if _handle is None:
            h = C.c_void_p()
            check(self._lib.fem2d_device_domain_upload(self.ctx._h, C.byref(domain.view()), C.byref(h)))
            _handle = h
        self._h = _handle
        n_elems, n_dofs, n_bs = C.c_int32(), C.c_int64(), C.c_int64()
        check(self._lib.fem2d_device_domain_info(self._h, C.byref(n_elems), C.byref(n_dofs), C.byref(n_bs)))
        self.n_elems, self.n_dofs, self.n_bs = n_elems.value, n_dofs.value, n_bs.value

    @classmethod
    def from_mesh(cls, mesh, cc: ContinuityCondition = ContinuityCondition.HCurl,
                  ctx: Optional[Context] = None) -> "DeviceDomain":
        """Domain::from_mesh (domain.rs:69-159) built on the GPU straight into HBM: the same device
        Domain as uploading ``Domain.from_mesh(mesh, cc)``, bit for bit."""
        ctx = ctx or default_context()
        lib = N.load()
        h = C.c_void_p()
        check(lib.fem2d_device_domain_from_mesh(ctx._h, mesh._h, int(cc), C.byref(h)))
        return cls(None, ctx, _handle=h)

    def download_view(self) -> dict:
        """The device Domain's view arrays (fem2d_domain_view without orders) as numpy arrays."""
        ne, nbs, nd = self.n_elems, self.n_bs, self.n_dofs
        out = {"parent": np.empty(ne, np.int32), "loc": np.empty(ne, np.uint8),
               "element_box": np.empty(4 * ne, np.float64), "materials": np.empty(2 * ne, np.float64),
               "bs_offsets": np.empty(ne + 1, np.int64), "bs_dir": np.empty(nbs, np.uint8),
               "bs_i": np.empty(nbs, np.uint8), "bs_j": np.empty(nbs, np.uint8),
               "bs_dof": np.empty(nbs, np.int32), "dof_bs": np.empty(2 * nd, np.int32)}
        ptr = [out[k].ctypes.data_as(C.c_void_p) for k in ("parent", "loc", "element_box", "materials", "bs_offsets",
                                                            "bs_dir", "bs_i", "bs_j", "bs_dof", "dof_bs")]
        check(self._lib.fem2d_device_domain_download_view(self.ctx._h, self._h, *ptr))
        return out

    def __del__(self):
        try:
            if self._h:
                self._lib.fem2d_device_domain_free(self._h)
        except Exception:
            pass

    def assemble(self, glq_grid_dim=None, BSpace=HierPoly, AI=CurlCurl, BI=L2Inner, precision: int = FP64
                 ) -> N.GepDevice:
        """Assemble into context-owned device buffers (valid until the next assembly)."""
        gu, gv = _glq(glq_grid_dim)
        out = N.GepDevice()
        check(self._lib.fem2d_assemble_device(self.ctx._h, self._h, BSpace.code, AI.code, BI.code, gu, gv,
                                              int(precision), C.byref(out)))
        return out

    def assembler(self, glq_grid_dim=None, BSpace=HierPoly, AI=CurlCurl, BI=L2Inner, precision: int = FP64):
        """A callable running :meth:`assemble` with these arguments, converted once: each call
        is one foreign call into the library and returns the same GepDevice object, refilled."""
        gu, gv = _glq(glq_grid_dim)
        out = N.GepDevice()
        fn = self._lib.fem2d_assemble_device
        args = (self.ctx._h, self._h, BSpace.code, AI.code, BI.code, gu, gv, int(precision), C.byref(out))

        def run() -> N.GepDevice:
            rc = fn(*args)
            if rc:
                check(rc)
            return out

        return run

    def print_to_petsc_binary_files(self, dev: N.GepDevice, dir: str, prefix: str) -> None:
        """PETSc AIJ binary files of a device-resident GEP (see GEP.print_to_petsc_binary_files)."""
        check(self._lib.fem2d_gep_device_print_to_petsc_binary_files(self.ctx._h, C.byref(dev), str(dir).encode(),
                                                                     str(prefix).encode()))

    def download(self, dev: N.GepDevice) -> GEP:
        out = N.GepCsr()
        check(self._lib.fem2d_gep_device_download(self.ctx._h, C.byref(dev), C.byref(out)))
        try:
            return GEP._from_csr(out)
        finally:
            self._lib.fem2d_gep_csr_free(C.byref(out))


def gauss_quadrature_points(n: int):
    """Gauss-Legendre nodes (ascending) and weights on (-1, 1) used by the library."""
    x = (C.c_double * n)()
    w = (C.c_double * n)()
    check(N.load().fem2d_gauss_legendre(int(n), x, w))
    return np.array(x[:]), np.array(w[:])
