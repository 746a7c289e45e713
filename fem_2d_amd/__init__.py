"""fem_2d_amd: MI355X-native Galerkin GEP assembly for fem_2d domains.

The prelude mirrors the reference crate's (src/lib.rs:14-39) names for the
assembly path: Mesh / HRef / PRef / Orders build the input, Domain numbers the
DoFs, galerkin_sample_gep_hcurl assembles A and B with HIP kernels on the
GPU, and nalgebra_solve_gep / solve_gep_dense solve the GEP there.
"""

from .domain import BasisDir, BasisSpec, ContinuityCondition, Domain
from .errors import (EmptyDOFSet, FailedToConverge, FailedToInvertB, Fem2dError, GalerkinSamplingError, HRefError,
                     InvalidGLQSettings, MeshAccessError, MeshPanic, MismatchedSolutionSize, MissingQuantity,
                     NalgebraGEPError, NativeError, PRefError, ProblemTooLarge, SpuriouslyConverged,
                     UniformFieldError, WrongContinuityCondition)
from .fields import UniformFieldSpace, xy_fields_array
from .galerkin import (FP32, FP32R, FP64, GEP, MIN_GLQ_ORDER, Context, CurlCurl, DeviceDomain, HierCurlBasisFnSpace,
                       HierCurlIntegral, HierPoly, KOLShapeFn, L2Inner, ShapeFn, default_context,
                       galerkin_sample_gep_hcurl, gauss_quadrature_points)
from .linalg import EigenPair, nalgebra_solve_gep, slepc_solve_gep, solve_gep_dense, solve_gep_sparse
from .mesh import (MAX_POLYNOMIAL_ORDER, MIN_EDGE_LENGTH, Edge, Elem, HLevels, HRef, Mesh, Node, Orders, PolyOrders,
                   PRef)

__all__ = [
    "BasisDir", "BasisSpec", "ContinuityCondition", "Domain", "EmptyDOFSet", "Fem2dError", "GalerkinSamplingError",
    "HRefError", "InvalidGLQSettings", "MeshAccessError", "MeshPanic", "NativeError", "PRefError",
    "WrongContinuityCondition", "FP32", "FP32R", "FP64", "GEP", "MIN_GLQ_ORDER", "Context", "CurlCurl", "DeviceDomain", "HierCurlBasisFnSpace",
    "HierCurlIntegral", "HierPoly", "KOLShapeFn", "L2Inner", "ShapeFn", "default_context",
    "galerkin_sample_gep_hcurl", "gauss_quadrature_points", "MAX_POLYNOMIAL_ORDER", "MIN_EDGE_LENGTH", "Edge", "Elem",
    "HLevels", "HRef", "Mesh", "Node", "Orders", "PolyOrders", "PRef", "EigenPair", "nalgebra_solve_gep",
    "solve_gep_dense", "solve_gep_sparse", "slepc_solve_gep", "NalgebraGEPError", "FailedToInvertB", "SpuriouslyConverged", "ProblemTooLarge",
    "FailedToConverge", "UniformFieldSpace", "xy_fields_array", "UniformFieldError", "MismatchedSolutionSize",
    "MissingQuantity",
]
