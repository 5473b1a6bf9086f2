"""MI355X-native SIMP state-solve backend (``backend="hip"``).

Public names mirror the reference package ``topogpu`` for the SolverV4 path.
GPU classes live in ``topogpu_amd.hip`` and load ``libtopohip.so`` lazily.
"""

from __future__ import annotations

__version__ = "0.1.0"

from .evidence import EvidenceBundle  # noqa: F401
from .gallery import CaseGallery, side_load_cantilever, tool_case  # noqa: F401
from .problem import TopologyProblem, cantilever_3d, multi_load_bracket, unit_cantilever  # noqa: F401
from .problem_spec import (  # noqa: F401
    DistributedLoad,
    EdgeSupport,
    PointLoad,
    PointSupport,
    ProblemSpec,
)
from .solver import OCOptimizer, OptimizationResult, SIMPModel, SIMPSolver, solve  # noqa: F401

__all__ = [
    "__version__", "EvidenceBundle", "TopologyProblem", "cantilever_3d", "multi_load_bracket", "unit_cantilever",
    "CaseGallery", "side_load_cantilever", "tool_case",
    "ProblemSpec",
    "PointSupport", "EdgeSupport", "PointLoad", "DistributedLoad", "SIMPSolver", "SIMPModel", "OCOptimizer",
    "OptimizationResult", "solve",
]
