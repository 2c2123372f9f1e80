"""researchuv_amd -- MI355X-native LSCM solve engine for ResearchUV's unfold path.

The compute lives in libruv_lscm.so (HIP kernels for gfx950 + a C++ host library, C ABI in
include/ruv_lscm.h).  This package is the Python mirror of the reference's interface for the path:
`solver` mirrors crates/researchuv-gpu (GpuSolver, CsrMatrix, Precond).
"""
from ._lib import RuvError, SolveInfo, lib  # noqa: F401
from .solver import CsrMatrix, GpuSolver, Plan, Precond  # noqa: F401

__all__ = ["CsrMatrix", "GpuSolver", "Plan", "Precond", "RuvError", "SolveInfo", "lib"]
