"""The C ABI library loads and exports every function include/fem2d.h declares; host-side
error paths of the boundary work without a GPU."""

import ctypes as C
import os
import re

import pytest

import fem_2d_amd as F
from fem_2d_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fem2d.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fem2d_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    lib = N.load()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_header():
    assert set(declared_functions()) == set(N.SIGNATURES)


def _call(view):
    out = N.GepCsr()
    return N.load().fem2d_galerkin_sample_gep_hcurl(None, C.byref(view), 0, 0, 1, 8, 8, 0, C.byref(out))


def test_boundary_errors_before_device_work():
    """galerkin.rs:41-59: the three GalerkinSamplingError variants are returned before any
    device (or context) is touched, in the reference's order."""
    assert _call(F.Domain.blank(F.ContinuityCondition.HDiv).view()) == N.ERR_WRONG_CC
    assert _call(F.Domain.blank(F.ContinuityCondition.HCurl).view()) == N.ERR_EMPTY_DOFS
    d = F.Domain.from_mesh(F.Mesh.unit())
    m = F.Mesh.unit()
    m.set_global_expansion_orders(F.Orders(2, 2))
    d = F.Domain.from_mesh(m)
    v = d.view()
    out = N.GepCsr()
    lib = N.load()
    assert lib.fem2d_galerkin_sample_gep_hcurl(None, C.byref(v), 0, 0, 1, 3, 8, 0, C.byref(out)) == N.ERR_INVALID_GLQ
    assert lib.fem2d_galerkin_sample_gep_hcurl(None, C.byref(v), 0, 0, 1, 0, 0, 0, C.byref(out)) == N.ERR_INVALID_ARG
    assert lib.fem2d_galerkin_sample_gep_hcurl(None, C.byref(v), 7, 0, 1, 8, 8, 0, C.byref(out)) == N.ERR_INVALID_ARG
    # precision codes: FEM2D_PRECISION_FP64 (0) and FEM2D_PRECISION_FP32 (1)
    assert lib.fem2d_galerkin_sample_gep_hcurl(None, C.byref(v), 0, 0, 1, 8, 8, 5, C.byref(out)) == N.ERR_INVALID_ARG


def test_context_without_device_fails_loudly():
    """No CPU fallback: without a HIP device the context cannot be created."""
    try:
        ctx = F.Context(0)
    except F.NativeError:
        return
    ctx.close()
    pytest.skip("a GPU is present")


def test_error_messages():
    with pytest.raises(F.MeshAccessError):
        F.Mesh.unit().elem(5)
    assert "non-existent" in N.last_error()


def test_device_domain_builder_argument_errors_without_a_device():
    """fem2d_device_domain_from_mesh rejects null handles and unknown continuity conditions
    before it touches the (absent) device; fem2d_device_domain_info rejects a null Domain."""
    lib = N.load()
    m = F.Mesh.unit()
    out = C.c_void_p(1)
    assert lib.fem2d_device_domain_from_mesh(None, m._h, 0, C.byref(out)) == N.ERR_INVALID_ARG
    assert lib.fem2d_device_domain_from_mesh(C.c_void_p(1), None, 0, C.byref(out)) == N.ERR_INVALID_ARG
    assert lib.fem2d_device_domain_from_mesh(C.c_void_p(1), m._h, 0, None) == N.ERR_INVALID_ARG
    assert lib.fem2d_device_domain_from_mesh(C.c_void_p(1), m._h, 3, C.byref(out)) == N.ERR_INVALID_ARG
    assert out.value is None  # cleared before the condition is checked
    n = C.c_int64()
    assert lib.fem2d_device_domain_info(None, None, C.byref(n), None) == N.ERR_INVALID_ARG
