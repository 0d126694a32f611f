"""MI355X (gfx950) backend of the PlaBundle frame bundle-adjustment solve.

The compute path is libpb_hip.so (hand-written HIP kernels + C++20 host
controller behind the C ABI of include/pb_hip.h). This package holds the
Python mirror of the reference's Solver/Problem/SolveOptions surface used by
the tests and the benchmark, and seeded synthetic problem generators.
"""
from .options import Backend, ImageRobustLoss, IntrinsicParameter, Precision, SolveOptions  # noqa: F401
from .problem import BrownConvention, GaugePolicy, Problem, intrinsic_block  # noqa: F401
from .solver import Result, SolveStatus, Solver, SolverWorkspace  # noqa: F401

__all__ = [
    "Backend", "ImageRobustLoss", "IntrinsicParameter", "Precision", "SolveOptions", "BrownConvention",
    "GaugePolicy", "Problem", "intrinsic_block", "Result", "SolveStatus", "Solver", "SolverWorkspace",
]
