"""ctypes declarations of the C ABI in include/fem2d.h (libfem2d_amd.so)."""

from __future__ import annotations

import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfem2d_amd.so")

# status codes (enum fem2d_status)
OK = 0
ERR_WRONG_CC = 1
ERR_EMPTY_DOFS = 2
ERR_INVALID_GLQ = 3
ERR_HREF = 10
ERR_PREF = 11
ERR_MESH_ACCESS = 12
ERR_PANIC = 13
ERR_IO = 14
ERR_FAILED_TO_INVERT_B = 20
ERR_SPURIOUSLY_CONVERGED = 21
ERR_PROBLEM_TOO_LARGE = 22
ERR_NOT_CONVERGED = 23
ERR_MISMATCHED_SOLUTION_SIZE = 24
ERR_INVALID_ARG = 100
ERR_INVALID_VIEW = 101
ERR_UNSUPPORTED = 102
ERR_HIP = 110
ERR_OOM = 111


class DomainView(C.Structure):
    _fields_ = [
        ("cc", C.c_int32),
        ("n_elems", C.c_int32),
        ("n_dofs", C.c_int64),
        ("parent", C.POINTER(C.c_int32)),
        ("loc", C.POINTER(C.c_uint8)),
        ("element_box", C.POINTER(C.c_double)),
        ("materials", C.POINTER(C.c_double)),
        ("orders", C.POINTER(C.c_uint8)),
        ("bs_offsets", C.POINTER(C.c_int64)),
        ("bs_dir", C.POINTER(C.c_uint8)),
        ("bs_i", C.POINTER(C.c_uint8)),
        ("bs_j", C.POINTER(C.c_uint8)),
        ("bs_dof", C.POINTER(C.c_int32)),
        ("dof_bs", C.POINTER(C.c_int32)),
    ]


class GepCsr(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("nnz", C.c_int64),
        ("row_ptr", C.POINTER(C.c_int64)),
        ("col", C.POINTER(C.c_int32)),
        ("a_val", C.POINTER(C.c_double)),
        ("b_val", C.POINTER(C.c_double)),
    ]


PRECISION_FP64 = 0
PRECISION_FP32 = 1
PRECISION_FP32R = 2


class GepDevice(C.Structure):
    _fields_ = GepCsr._fields_ + [
        ("precision", C.c_int32),
        ("a_val32", C.POINTER(C.c_float)),
        ("b_val32", C.POINTER(C.c_float)),
        ("generation", C.c_uint64),
        ("res16", C.POINTER(C.c_uint32)),
    ]


class EigenPair(C.Structure):
    _fields_ = [("value", C.c_double), ("n", C.c_int64), ("vector", C.POINTER(C.c_double))]


NOALLOC_SCRATCH_BYTES = 64 << 20  # FEM2D_NOALLOC_SCRATCH_BYTES
NALGEBRA_MAX_DENSE_SIZE = 1000  # FEM2D_NALGEBRA_MAX_DENSE_SIZE (nalgebra_solve.rs:6)
DENSE_GEP_MAX_SIZE = 32768  # FEM2D_DENSE_GEP_MAX_SIZE


class Timing(C.Structure):
    _fields_ = [
        ("total_ms", C.c_double),
        ("prep_ms", C.c_double),
        ("symbolic_ms", C.c_double),
        ("fill_ms", C.c_double),
        ("n_groups", C.c_int64),
        ("count_kernel_ms", C.c_double),
        ("fill_kernel_ms", C.c_double),
        ("nnz", C.c_int64),
        ("n_large", C.c_int64),
        ("fill_chunk", C.c_int64),
        ("fill_lib_lds", C.c_int64),
    ]


class EigenStats(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("blocks", C.c_int32),
        ("max_block", C.c_int32),
        ("factor_bytes", C.c_double),
        ("iterations", C.c_int32),
        ("inner_iterations", C.c_int32),
        ("residual", C.c_double),
        ("max_inner_per_solve", C.c_double),
        ("symbolic_ms", C.c_double),
        ("factor_ms", C.c_double),
        ("iterate_ms", C.c_double),
        ("min_pivot", C.c_double),
        ("precond_error", C.c_double),
        ("bad_block", C.c_int32),
    ]


class ElemInfo(C.Structure):
    _fields_ = [
        ("id", C.c_int32),
        ("nodes", C.c_int32 * 4),
        ("edges", C.c_int32 * 4),
        ("element", C.c_int32),
        ("parent", C.c_int32),
        ("loc", C.c_int32),
        ("h_levels", C.c_int32 * 2),
        ("orders", C.c_int32 * 2),
        ("n_children", C.c_int32),
        ("children", C.c_int32 * 4),
        ("depth", C.c_int32),
    ]


class EdgeInfo(C.Structure):
    _fields_ = [
        ("id", C.c_int32),
        ("nodes", C.c_int32 * 2),
        ("boundary", C.c_int32),
        ("dir", C.c_int32),
        ("length", C.c_double),
        ("parent", C.c_int32),
        ("children", C.c_int32 * 2),
        ("child_node", C.c_int32),
        ("active", C.c_int32 * 2),
    ]


P = C.POINTER
vp = C.c_void_p
i32, i64, dbl = C.c_int32, C.c_int64, C.c_double

# name -> (restype, argtypes)
SIGNATURES = {
    "fem2d_ctx_create": (C.c_int, [C.c_int, P(vp)]),
    "fem2d_ctx_destroy": (None, [vp]),
    "fem2d_ctx_stream": (vp, [vp]),
    "fem2d_ctx_synchronize": (C.c_int, [vp]),
    "fem2d_ctx_set_autotune": (C.c_int, [vp, C.c_int]),
    "fem2d_ctx_set_fill_chunk": (C.c_int, [vp, C.c_int]),
    "fem2d_ctx_last_eigen_stats": (C.c_int, [vp, P(EigenStats)]),
    "fem2d_galerkin_sample_gep_hcurl": (
        C.c_int, [vp, P(DomainView), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P(GepCsr)]),
    "fem2d_gep_csr_free": (None, [P(GepCsr)]),
    "fem2d_device_domain_upload": (C.c_int, [vp, P(DomainView), P(vp)]),
    "fem2d_device_domain_from_mesh": (C.c_int, [vp, vp, i32, P(vp)]),
    "fem2d_device_domain_info": (C.c_int, [vp, P(i32), P(i64), P(i64)]),
    "fem2d_device_domain_download_view": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "fem2d_device_domain_free": (None, [vp]),
    "fem2d_assemble_device": (
        C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P(GepDevice)]),
    "fem2d_gep_device_download": (C.c_int, [vp, P(GepDevice), P(GepCsr)]),
    "fem2d_assemble_device_noalloc": (
        C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, i64, vp, vp, vp, vp, vp, P(i64)]),
    "fem2d_ctx_set_graphs": (C.c_int, [vp, C.c_int]),
    "fem2d_gep_device_print_to_petsc_binary_files": (C.c_int, [vp, P(GepDevice), C.c_char_p, C.c_char_p]),
    "fem2d_gep_csr_print_to_petsc_binary_files": (C.c_int, [vp, P(GepCsr), C.c_char_p, C.c_char_p]),
    "fem2d_device_domain_num_leaves": (C.c_int, [vp, P(C.c_int64)]),
    "fem2d_device_malloc": (C.c_int, [vp, C.c_int64, P(vp)]),
    "fem2d_device_free": (C.c_int, [vp, vp]),
    "fem2d_memcpy_to_device": (C.c_int, [vp, vp, vp, C.c_int64]),
    "fem2d_memcpy_to_host": (C.c_int, [vp, vp, vp, C.c_int64]),
    "fem2d_device_xy_fields": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int, vp, C.c_int64, vp, vp]),
    "fem2d_xy_fields": (C.c_int, [vp, P(DomainView), C.c_int, C.c_int, C.c_int, P(C.c_double), C.c_int64,
                                  P(C.c_double), P(C.c_double)]),
    "fem2d_eigenpair_free": (None, [P(EigenPair)]),
    "fem2d_nalgebra_solve_gep": (C.c_int, [vp, P(GepCsr), C.c_double, P(EigenPair)]),
    "fem2d_gep_device_nalgebra_solve": (C.c_int, [vp, P(GepDevice), C.c_double, P(EigenPair)]),
    "fem2d_solve_gep_dense": (C.c_int, [vp, P(GepCsr), C.c_double, P(EigenPair)]),
    "fem2d_gep_device_solve_dense": (C.c_int, [vp, P(GepDevice), C.c_double, P(EigenPair)]),
    "fem2d_solve_gep_sparse": (C.c_int, [vp, P(GepCsr), C.c_double, P(EigenPair)]),
    "fem2d_gep_device_solve_sparse": (C.c_int, [vp, P(GepDevice), C.c_double, P(EigenPair)]),
    "fem2d_ctx_last_timing": (C.c_int, [vp, P(Timing)]),
    "fem2d_last_error": (C.c_char_p, []),
    "fem2d_last_error_detail": (C.c_int, []),
    "fem2d_mesh_from_file": (C.c_int, [C.c_char_p, P(vp)]),
    "fem2d_mesh_from_json": (C.c_int, [C.c_char_p, P(vp)]),
    "fem2d_mesh_from_parts": (C.c_int, [i64, P(dbl), i64, P(i32), P(dbl), P(vp)]),
    "fem2d_mesh_unit": (C.c_int, [P(vp)]),
    "fem2d_mesh_blank": (C.c_int, [P(vp)]),
    "fem2d_mesh_clone": (C.c_int, [vp, P(vp)]),
    "fem2d_mesh_free": (None, [vp]),
    "fem2d_mesh_counts": (C.c_int, [vp, P(i64), P(i64), P(i64), P(i64)]),
    "fem2d_mesh_elem_info": (C.c_int, [vp, i32, P(ElemInfo)]),
    "fem2d_mesh_edge_info": (C.c_int, [vp, i32, P(EdgeInfo)]),
    "fem2d_mesh_node_info": (C.c_int, [vp, i32, P(dbl), P(dbl), P(i32)]),
    "fem2d_mesh_element_info": (C.c_int, [vp, i32, P(dbl), P(dbl)]),
    "fem2d_mesh_descendant_elems": (C.c_int, [vp, i32, C.c_int, P(i32), i64, P(i64)]),
    "fem2d_mesh_ancestor_elems": (C.c_int, [vp, i32, C.c_int, P(i32), i64, P(i64)]),
    "fem2d_mesh_descendant_edges": (C.c_int, [vp, i32, C.c_int, P(i32), i64, P(i64)]),
    "fem2d_mesh_max_expansion_orders": (C.c_int, [vp, P(i32)]),
    "fem2d_mesh_elem_is_h_refineable": (C.c_int, [vp, i32, P(i32)]),
    "fem2d_mesh_elem_p_refinement_window": (C.c_int, [vp, i32, P(i32)]),
    "fem2d_mesh_elem_parametric_range": (C.c_int, [vp, i32, i32, P(dbl)]),
    "fem2d_mesh_global_h_refinement": (C.c_int, [vp, i32, i32]),
    "fem2d_mesh_h_refine_elems": (C.c_int, [vp, P(i32), i64, i32, i32]),
    "fem2d_mesh_execute_h_refinements": (C.c_int, [vp, P(i32), P(i32), P(i32), i64]),
    "fem2d_mesh_execute_h_refinements_threads": (C.c_int, [vp, P(i32), P(i32), P(i32), i64, i32]),
    "fem2d_mesh_global_p_refinement": (C.c_int, [vp, i32, i32]),
    "fem2d_mesh_p_refine_elems": (C.c_int, [vp, P(i32), i64, i32, i32]),
    "fem2d_mesh_execute_p_refinements": (C.c_int, [vp, P(i32), P(i32), P(i32), i64]),
    "fem2d_mesh_set_global_expansion_orders": (C.c_int, [vp, i32, i32]),
    "fem2d_mesh_set_expansion_orders": (C.c_int, [vp, P(i32), P(i32), P(i32), i64]),
    "fem2d_domain_from_mesh": (C.c_int, [vp, i32, P(vp)]),
    "fem2d_domain_from_mesh_threads": (C.c_int, [vp, i32, i32, P(vp)]),
    "fem2d_domain_from_mesh_serial": (C.c_int, [vp, i32, P(vp)]),
    "fem2d_domain_blank": (C.c_int, [i32, P(vp)]),
    "fem2d_domain_free": (None, [vp]),
    "fem2d_domain_num_dofs": (C.c_int, [vp, P(i64)]),
    "fem2d_domain_mesh": (C.c_int, [vp, P(vp)]),
    "fem2d_domain_get_view": (C.c_int, [vp, P(DomainView)]),
    "fem2d_domain_dof_info": (C.c_int, [vp, i64, P(i32), P(i32), P(i32)]),
    "fem2d_gauss_legendre": (C.c_int, [i32, P(dbl), P(dbl)]),
}

_lib = None


def load() -> C.CDLL:
    """Load libfem2d_amd.so; raises ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m fem_2d_amd.build_native` "
            "(the assembly has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().fem2d_last_error()
    return msg.decode() if msg else ""


def last_error_detail() -> int:
    return int(load().fem2d_last_error_detail())
