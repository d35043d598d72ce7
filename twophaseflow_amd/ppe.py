"""PPE plugin boundary and the MI355X matrix-free FVM solver.

Mirrors the reference plugin interface (``ppe/interfaces.py:24-88``,
``ppe/factory.py:31-120``) and ``PPESolverFVMMatrixFree``
(``ppe/fvm_matrixfree.py:46-363``) in CG form: the system
``(-W L) p = -W rhs`` with the centre-node pin, Jacobi preconditioning, the
recursively updated ``||r|| <= tol ||b||`` stop, ``RuntimeError`` when it
does not converge, and ``p[pin] = 0`` on return.  The Krylov loop runs on
the device (libtpf, hipGraph chunks); nothing iterates in Python.

Options read from the config like the reference (``getattr``):
``pseudo_tol``, ``pseudo_maxiter``, ``ppe_iteration_method`` (only "cg" is
built: the reference's GMRES + line-PCR route is not), ``ppe_preconditioner``
("jacobi"), plus ``ppe_precision`` ("fp64" | "mixed") and
``ppe_check_every``.
"""

from __future__ import annotations

import ctypes as C
from abc import ABC, abstractmethod
from typing import Callable, Dict

import numpy as np
import torch

from . import _native as nat
from .boundary import BoundarySpec, axis_kinds, pin_dof
from .device import Handoff, ptr, require_gpu, stream_handle


class MatrixAssemblyUnavailable(RuntimeError):
    """Raised when a PPE solver does not expose a sparse matrix
    (ppe/interfaces.py:24-25)."""


class IPPESolver(ABC):
    """PPE plugin interface (ppe/interfaces.py:28-88)."""

    @abstractmethod
    def solve(self, rhs, rho, dt: float, p_init=None):
        """Solve div(1/rho grad p) = rhs; return p of grid.shape."""

    def get_matrix(self, rho):
        raise MatrixAssemblyUnavailable(
            f"{type(self).__name__} does not provide an assembled PPE matrix")

    def update_grid(self, grid) -> None:
        return None

    def invalidate_cache(self) -> None:
        return None

    def clear_interface_jump_context(self) -> None:
        return None


class PPESolverFVMMatrixFree(IPPESolver):
    """Matrix-free variable-density FVM PPE on the GPU (CG form, Jacobi)."""

    scheme_names = ("fvm_iterative",)
    _scheme_aliases = {"fvm_matrixfree": "fvm_iterative", "matrixfree": "fvm_iterative"}

    def __init__(self, backend, config, grid, bc_type: str = "wall", bc_spec=None):
        self.backend = backend
        self.xp = np
        self.grid = grid
        self.ndim = grid.ndim
        self.bc_type = bc_type
        solver_cfg = getattr(config, "solver", config)
        if any(k != "wall" for k in axis_kinds(bc_type, self.ndim)):
            raise ValueError(
                "PPESolverFVMMatrixFree supports bc_type='wall' only; "
                f"got {bc_type!r}.")
        self.tol = float(getattr(solver_cfg, "pseudo_tol", 1e-8))
        self.maxiter = int(getattr(solver_cfg, "pseudo_maxiter", 500))
        method = str(getattr(solver_cfg, "ppe_iteration_method", "cg")).strip().lower()
        if method in {"explicit", "gauss_seidel", "adi"}:
            method = "cg"
        if method != "cg":
            raise ValueError(
                "twophaseflow_amd builds the CG form of the FVM PPE only; "
                f"ppe_iteration_method={method!r} is not available")
        self.iteration_method = method
        self.preconditioner = str(getattr(solver_cfg, "ppe_preconditioner", "jacobi")).lower()
        if self.preconditioner != "jacobi":
            raise ValueError("ppe_iteration_method='cg' supports preconditioner='jacobi' here")
        self.precision = str(getattr(solver_cfg, "ppe_precision", "fp64")).lower()
        if self.precision not in ("fp64", "mixed"):
            raise ValueError("ppe_precision must be 'fp64' or 'mixed'")
        self.check_every = int(getattr(solver_cfg, "ppe_check_every", 32))
        self.inner_tol = float(getattr(solver_cfg, "ppe_inner_tol", 0.0))
        spec = bc_spec if bc_spec is not None else BoundarySpec(bc_type, grid.shape, grid.N)
        self._pin_dof = spec.pin_dof
        self.shape = tuple(int(n) + 1 for n in grid.N)
        self.last_diagnostics: dict = {}
        self._lib = nat.lib()
        self._handle = None
        self._rho = None
        self._create()

    def _create(self):
        require_gpu()
        h = C.c_void_p()
        nat.check(self._lib.tpf_ppe_create(C.byref(h), self.ndim, nat.i64arr(self.shape),
                                           nat.darr(self.grid.L), int(self._pin_dof),
                                           stream_handle()))
        self._handle = h

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            self._lib.tpf_ppe_destroy(h)
            self._handle = None

    def update_grid(self, grid=None) -> None:
        if grid is not None:
            self.grid = grid
            self.ndim = grid.ndim
            self.shape = tuple(int(n) + 1 for n in grid.N)
            self._pin_dof = pin_dof(grid.N, self.bc_type)
        self.__del__()
        self._rho = None
        self._create()

    def invalidate_cache(self) -> None:
        self._rho = None

    # -- operator --------------------------------------------------------------
    def prepare_operator(self, rho) -> None:
        """Keep rho on the device and build the Jacobi diagonal
        (fvm_matrixfree.py:262-279, 352-363)."""
        io = Handoff(rho)
        t = io.to_device(rho)
        if tuple(t.shape) != self.shape:
            raise ValueError(f"rho shape {tuple(t.shape)} != grid.shape {self.shape}")
        self._rho = t
        nat.check(self._lib.tpf_ppe_set_density(self._handle, ptr(t)))

    def apply(self, p):
        """Pinned operator L p with out[pin] = p[pin] (fvm_matrixfree.py:281-300)."""
        if self._rho is None:
            raise RuntimeError("prepare_operator(rho) must be called before apply()")
        io = Handoff(p)
        t = io.to_device(p)
        out = torch.empty_like(t)
        nat.check(self._lib.tpf_ppe_apply(self._handle, ptr(t), ptr(out)))
        return io.back(out)

    # -- solve -----------------------------------------------------------------
    def solve_into(self, rhs: torch.Tensor, rho: torch.Tensor | None, p: torch.Tensor,
                   use_x0: bool = False) -> dict:
        """Zero-copy solve on CUDA tensors; p holds p_init when use_x0."""
        if rho is not None:
            if self._rho is None or rho.data_ptr() != self._rho.data_ptr():
                self.prepare_operator(rho)
        elif self._rho is None:
            raise RuntimeError("no density set")
        opts = nat.SolveOpts(tol=self.tol, maxiter=self.maxiter, precond=1,
                             check_every=self.check_every, use_x0=1 if use_x0 else 0,
                             mixed=1 if self.precision == "mixed" else 0,
                             inner_tol=self.inner_tol)
        rep = nat.SolveReport()
        status = self._lib.tpf_ppe_solve(self._handle, ptr(rhs), ptr(p), C.byref(opts),
                                         C.byref(rep))
        key = f"ppe_low_order_{self.iteration_method}"
        self.last_diagnostics = {
            f"{key}_info": 0 if status == nat.TPF_OK else int(rep.iterations),
            f"{key}_tol": self.tol,
            f"{key}_maxiter": self.maxiter,
            "iterations": int(rep.iterations),
            "relres": float(rep.relres),
            "true_relres": float(rep.true_relres),
            "refinements": int(rep.refinements),
            "precision": self.precision,
        }
        if status != nat.TPF_OK:
            msg = self._lib.tpf_last_error().decode(errors="replace")
            if status == nat.TPF_ERR_RUNTIME:
                raise RuntimeError(
                    f"{type(self).__name__} did not converge cleanly: {msg}; no alternate "
                    "PPE solver was applied.")
            nat.check(status)
        return self.last_diagnostics

    def solve(self, rhs, rho, dt: float = 0.0, p_init=None):
        io = Handoff(rhs)
        rhs_t = io.to_device(rhs)
        if tuple(rhs_t.shape) != self.shape:
            raise ValueError(f"rhs shape {tuple(rhs_t.shape)} != grid.shape {self.shape}")
        rho_t = io.to_device(rho)
        if p_init is None:
            p = torch.empty_like(rhs_t)
            use_x0 = False
        else:
            p = io.to_device(p_init).clone()
            use_x0 = True
        self.solve_into(rhs_t, rho_t, p, use_x0=use_x0)
        return io.back(p)

    def true_relres(self, rhs, p) -> float:
        """FP64 ||b - A p|| / ||b|| on the CG system, recomputed on the device."""
        io = Handoff(rhs)
        r = io.to_device(rhs)
        x = io.to_device(p)
        out = C.c_double()
        nat.check(self._lib.tpf_ppe_true_relres(self._handle, ptr(r), ptr(x), C.byref(out)))
        return out.value


# -- factory (ppe/factory.py:31-120) ------------------------------------------
_SOLVER_REGISTRY: Dict[str, Callable] = {}


def register_ppe_solver(name: str, factory_fn: Callable) -> None:
    _SOLVER_REGISTRY[name] = factory_fn


def _create_fvm_matrixfree(config, backend, grid, ccd, bc_spec):
    return PPESolverFVMMatrixFree(backend, config, grid,
                                  bc_type=config.numerics.bc_type, bc_spec=bc_spec)


register_ppe_solver("fvm_iterative", _create_fvm_matrixfree)
register_ppe_solver("fvm_matrixfree", _create_fvm_matrixfree)


def create_ppe_solver(config, backend, grid, ccd=None, bc_spec=None) -> IPPESolver:
    if bc_spec is None:
        bc_spec = BoundarySpec(config.numerics.bc_type, tuple(n + 1 for n in config.grid.N),
                               config.grid.N)
    solver_type = config.solver.ppe_solver_type
    fn = _SOLVER_REGISTRY.get(solver_type)
    if fn is None:
        raise ValueError(f"Unknown ppe_solver_type: '{solver_type}'. "
                         f"Available: {', '.join(sorted(_SOLVER_REGISTRY))}")
    return fn(config, backend, grid, ccd, bc_spec)
