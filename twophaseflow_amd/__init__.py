"""twophaseflow_amd -- MI355X-native CCD operator and variable-density PPE.

Drop-in replacements for the two hot paths of TwoPhaseFlow behind its own
Python API: ``CCDSolver`` (ccd/ccd_solver.py) and the ``IPPESolver`` plugins
``PPESolverFVMMatrixFree`` (ppe/fvm_matrixfree.py) and
``PPESolverFCCDMatrixFree`` (ppe/fccd_matrixfree.py, default configuration).  The compute runs in
hand-written HIP kernels for gfx950 in ``libtpf.so`` (C ABI in
include/tpf.h), reached through ctypes; PyTorch provides device memory and
the stream only.  There is no CPU fallback.
"""

from .config import FluidConfig, GridConfig, NumericsConfig, SimulationConfig, SolverConfig
from .grid import Grid
from .boundary import BoundarySpec, axis_kinds, pin_dof
from .ccd import CCDBatch, CCDSolver
from .fccd import FCCDFaceJet, FCCDSolver
from .defect_correction import PPESolverDefectCorrection
from .rhie_chow import RhieChowInterpolator
from .dccd import DissipativeCCDAdvection
from .curvature import CurvatureCalculatorPsi
from .reinit import SplitReinitializer
from .ppe import (IPPESolver, MatrixAssemblyUnavailable, PPEBatch, PPESolverFCCDMatrixFree,
                  PPESolverFVMMatrixFree, create_ppe_solver, register_ppe_solver)
from .ns_terms import (BoundaryConditionHandler, CFLCalculator, ClosestPointExtender,
                       NullFieldExtender, Predictor, UCCD6ConvectionTerm, VelocityCorrector,
                       ViscousTerm)
from .simulation import (BatchedSimulation, SimulationBuilder, TwoPhaseSimulation,
                         bubble_simulation)

__all__ = [
    "FluidConfig", "GridConfig", "NumericsConfig", "SimulationConfig", "SolverConfig", "Grid",
    "BoundarySpec", "axis_kinds", "pin_dof", "CCDSolver", "CCDBatch", "FCCDSolver",
    "FCCDFaceJet", "PPESolverDefectCorrection", "RhieChowInterpolator",
    "DissipativeCCDAdvection", "CurvatureCalculatorPsi", "SplitReinitializer", "IPPESolver",
    "MatrixAssemblyUnavailable", "PPESolverFVMMatrixFree", "PPESolverFCCDMatrixFree",
    "create_ppe_solver",
    "register_ppe_solver", "BoundaryConditionHandler", "CFLCalculator", "ClosestPointExtender",
    "NullFieldExtender", "Predictor", "UCCD6ConvectionTerm", "VelocityCorrector", "ViscousTerm",
    "SimulationBuilder", "TwoPhaseSimulation", "PPEBatch", "BatchedSimulation",
    "bubble_simulation",
]
