"""Error types mirroring the reference's error enums, raised from C ABI status codes."""

from __future__ import annotations

from . import _native as N


class Fem2dError(Exception):
    """Base class of all errors raised by fem_2d_amd."""


class GalerkinSamplingError(Fem2dError):
    """GalerkinSamplingError (galerkin.rs:191-195)."""


class WrongContinuityCondition(GalerkinSamplingError):
    pass


class EmptyDOFSet(GalerkinSamplingError):
    pass


class InvalidGLQSettings(GalerkinSamplingError):
    pass


class HRefError(Fem2dError):
    """HRefError (h_refinement.rs:284-296); ``kind`` names the variant."""

    KINDS = ["MinEdgeLength", "ElemHasChildren", "EdgeHasChildren", "UninitializedElem", "EdgeOnEqualPoints",
             "BisectionIdxExceeded", "ElemDoesNotExist", "ElemNotRefineable", "DuplicateElemIds"]

    def __init__(self, msg: str, detail: int):
        super().__init__(msg)
        self.kind = self.KINDS[detail - 1] if 1 <= detail <= len(self.KINDS) else "Unknown"


class PRefError(Fem2dError):
    """PRefError (p_refinement.rs:305-317); ``kind`` names the variant."""

    KINDS = ["NegExpansion", "ExceededMaxExpansion", "DuplicateElemIds", "ElemDoesNotExist", "RefinementOutOfBounds",
             "InvalidExpansionOrders"]

    def __init__(self, msg: str, detail: int):
        super().__init__(msg)
        self.kind = self.KINDS[detail - 1] if 1 <= detail <= len(self.KINDS) else "Unknown"


class MeshAccessError(Fem2dError):
    """MeshAccessError (mesh.rs:1799-1804)."""


class MeshPanic(Fem2dError):
    """A condition on which the reference panics (assert!, panic!, unwrap of an Err)."""


class NalgebraGEPError(Fem2dError):
    """NalgebraGEPError (nalgebra_solve.rs:52-58)."""


class FailedToInvertB(NalgebraGEPError):
    pass


class SpuriouslyConverged(NalgebraGEPError):
    pass


class ProblemTooLarge(NalgebraGEPError):
    pass


class FailedToConverge(Fem2dError):
    """The eigensolver's iteration did not converge (SlepcGEPError::FailedToConverge, slepc_solve.rs:70-81)."""


class UniformFieldError(Fem2dError):
    """UniformFieldError (fields.rs:407-411)."""


class MismatchedSolutionSize(UniformFieldError):
    pass


class MissingQuantity(UniformFieldError):
    pass


class NativeError(Fem2dError):
    """Library-level failure: HIP runtime, allocation, invalid domain view, unsupported size."""

    def __init__(self, msg: str, status: int):
        super().__init__(msg)
        self.status = status


def check(rc: int) -> None:
    """Raise the exception matching a C ABI status code."""
    if rc == N.OK:
        return
    msg = N.last_error()
    detail = N.last_error_detail()
    if rc == N.ERR_WRONG_CC:
        raise WrongContinuityCondition(msg)
    if rc == N.ERR_EMPTY_DOFS:
        raise EmptyDOFSet(msg)
    if rc == N.ERR_INVALID_GLQ:
        raise InvalidGLQSettings(msg)
    if rc == N.ERR_HREF:
        raise HRefError(msg, detail)
    if rc == N.ERR_PREF:
        raise PRefError(msg, detail)
    if rc == N.ERR_MESH_ACCESS:
        raise MeshAccessError(msg)
    if rc == N.ERR_PANIC:
        raise MeshPanic(msg)
    if rc == N.ERR_IO:
        raise FileNotFoundError(msg)
    if rc == N.ERR_FAILED_TO_INVERT_B:
        raise FailedToInvertB(msg)
    if rc == N.ERR_SPURIOUSLY_CONVERGED:
        raise SpuriouslyConverged(msg)
    if rc == N.ERR_PROBLEM_TOO_LARGE:
        raise ProblemTooLarge(msg)
    if rc == N.ERR_NOT_CONVERGED:
        raise FailedToConverge(msg)
    if rc == N.ERR_MISMATCHED_SOLUTION_SIZE:
        raise MismatchedSolutionSize(msg)
    if rc == N.ERR_INVALID_ARG:
        raise ValueError(msg)
    raise NativeError(f"status {rc}: {msg}", rc)
