"""phipict_amd: MI355X (gfx950) implementation of phiPICT's PISO time step and linear solvers.

The public names mirror the reference package ``phipict`` for this path (src/phipict/__init__.py):
the domain model and operator functions of ``phipict._C``, ``Simulation``/``MHDSimulation``,
``SolverTolerance``, hooks and boundary-condition helpers. All compute runs in the gfx950
library ``phipict_amd/_lib/libphipict_hip.so``; importing the package fails if it is missing.
"""
from . import _lib  # noqa: F401  (fails loudly without the gfx950 library)
from .domain import (AdvectionScheme, Block, Boundary, BoundaryConditionType, BoundaryType, ConnectedBoundary,
                     ConvergenceCriterion, CSRmatrix, Domain, FixedBoundary, PeriodicBoundary, PotentialBC)
from .ops import (AMGPCGSolve, ComputeCurrentDensityFaceBased, ComputeEpotRHS, CoordsToFaceTransforms,
                  CoordsToTransforms, CopyEpotResultFromBlocks, CopyEpotResultToBlocks, CopyPressureResultFromBlocks,
                  CopyPressureResultToBlocks, CopyScalarResultFromBlocks, CopyScalarResultToBlocks,
                  CopyVelocityResultFromBlocks, CopyVelocityResultToBlocks, CorrectVelocity,
                  GetKrylovSettings, LinearSolverResultInfo, ReleaseKrylovWorkspaces, SetKrylovSettings,
                  SetupAdvectionMatrix, SetupAdvectionScalar, SetupAdvectionVelocity, SetupEpotMatrix,
                  SetupPressureCorrection,
                  SetupPressureMatrix, SetupPressureRHS, SetupPressureRHSdiv, SolveLinear, SolveSegmented,
                  SparseOuterProduct)
from .hooks import Hook, Hooks
from .simulation import LinsolveError, Simulation
from .mhd import MHDSimulation
from .solvers.tolerance import SolverTolerance
from .bc import Face, Potential, set_bc

__all__ = [n for n in dir() if not n.startswith("_")]
