"""Configuration dataclasses with the reference's field names (subset that the
CCD and PPE paths read).

Mirrors ``config.py`` of the reference: ``GridConfig`` (grid axes),
``NumericsConfig.bc_type`` (:239), ``SolverConfig`` (:322-358) and
``SimulationConfig`` (:363-381).  Extra attributes the reference reads with
``getattr`` (``ppe_preconditioner``, ``ppe_restart``, ``ppe_pcr_stages``,
fvm_matrixfree.py:104-127) may be set dynamically the same way; this build
adds ``ppe_precision`` ("fp64" | "mixed") and ``ppe_check_every``.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Tuple


@dataclass
class GridConfig:
    ndim: int = 2
    N: Tuple[int, ...] = (64, 64)
    L: Tuple[float, ...] = (1.0, 1.0)
    alpha_grid: float = 1.0  # 1.0 = uniform; other values are not supported here

    def __post_init__(self):
        if len(self.N) != self.ndim or len(self.L) != self.ndim:
            raise ValueError("GridConfig: len(N) and len(L) must equal ndim")


@dataclass
class NumericsConfig:
    bc_type: str = "wall"
    cfl_number: float = 0.3


@dataclass
class SolverConfig:
    ppe_solver_type: str = "fvm_iterative"
    pseudo_tol: float = 1e-8
    pseudo_maxiter: int = 500
    pseudo_c_tau: float = 2.0
    ppe_iteration_method: str = "adi"

    def __post_init__(self):
        aliases = {"fvm_matrixfree": "fvm_iterative", "fvm_hip": "fvm_iterative"}
        self.ppe_solver_type = aliases.get(self.ppe_solver_type, self.ppe_solver_type)


@dataclass
class SimulationConfig:
    grid: GridConfig = field(default_factory=GridConfig)
    numerics: NumericsConfig = field(default_factory=NumericsConfig)
    solver: SolverConfig = field(default_factory=SolverConfig)
    use_gpu: bool = True
