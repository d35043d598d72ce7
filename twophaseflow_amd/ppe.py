"""PPE plugin boundary and the MI355X matrix-free FVM solver.

Mirrors the reference plugin interface (``ppe/interfaces.py:24-88``,
``ppe/factory.py:31-120``) and ``PPESolverFVMMatrixFree``
(``ppe/fvm_matrixfree.py:46-363``) in both of its forms:

* GMRES (the reference default, also chosen by "explicit" / "gauss_seidel" /
  "adi"): ``L p = rhs`` with the pin row = identity, left-preconditioned
  restarted GMRES with SciPy's stopping rules, preconditioner "line_pcr"
  (default), "jacobi" or "none";
* CG: ``(-W L) p = -W rhs``, Jacobi preconditioning, the recursively updated
  ``||r|| <= tol ||b||`` stop (FP64 or FP32-inner/FP64-refined); or, with
  this build's ``ppe_preconditioner="mg"``, flexible CG preconditioned by one
  FP32 multigrid cycle per iteration (``ppe_mg_cycle`` "W" | "V",
  ``ppe_mg_smooth``, ``ppe_mg_coarse_sweeps``, ``ppe_mg_w_levels``: W
  recursion on the top two levels by default), judged by the FP64 true
  residual.

``PPESolverFCCDMatrixFree`` mirrors the FCCD matrix-free PPE
(``ppe/fccd_matrixfree.py``) with the phase_density and phase_separated
coefficient schemes and the none / jump_decomposition / affine_jump
interface couplings: the operator is the CCD second derivative per axis plus
one fused face-gradient / weighted divergence kernel, solved with the same
device GMRES.

All raise ``RuntimeError`` when they do not converge and return
``p[pin] = 0``.  The Krylov loops run on the device (libtpf); nothing
iterates in Python.

Options read from the config like the reference (``getattr``):
``pseudo_tol``, ``pseudo_maxiter``, ``pseudo_c_tau``, ``ppe_iteration_method``,
``ppe_preconditioner``, ``ppe_restart`` (default ``min(80, max(20,
maxiter))``), ``ppe_pcr_stages`` (default 4; the line solves use that many
PCR stages, as the reference's GPU backend does), plus this build's
``ppe_line_solver`` ("pcr" | "thomas": exact Thomas, the reference's CPU
backend), ``ppe_precision`` ("fp64" | "mixed", CG) and ``ppe_check_every``.
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
from .grid import require_uniform


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
    """Matrix-free variable-density FVM PPE on the GPU (GMRES or CG form)."""

    scheme_names = ("fvm_iterative",)
    _scheme_aliases = {"fvm_matrixfree": "fvm_iterative", "matrixfree": "fvm_iterative"}

    def __init__(self, backend, config, grid, bc_type: str = "wall", bc_spec=None):
        require_uniform(grid, "PPESolverFVMMatrixFree")
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
        self.c_tau = float(getattr(solver_cfg, "pseudo_c_tau", 2.0))
        method = str(getattr(solver_cfg, "ppe_iteration_method", "gmres")).strip().lower()
        if method in {"explicit", "gauss_seidel", "adi"}:
            method = "gmres"
        if method not in {"gmres", "cg"}:
            raise ValueError(
                "PPESolverFVMMatrixFree supports ppe_iteration_method='gmres' or 'cg' "
                f"today, got {method!r}")
        self.iteration_method = method
        restart = getattr(solver_cfg, "ppe_restart", None)
        self.restart = int(restart) if restart is not None else min(80, max(20, self.maxiter))
        default_pre = "jacobi" if method == "cg" else "line_pcr"
        self.preconditioner = str(
            getattr(solver_cfg, "ppe_preconditioner", default_pre)).strip().lower()
        if self.preconditioner not in {"jacobi", "line_pcr", "none", "mg"}:
            raise ValueError("ppe_preconditioner must be 'jacobi', 'line_pcr', 'none' or "
                             f"'mg', got {self.preconditioner!r}")
        if self.preconditioner == "mg" and method != "cg":
            raise ValueError("ppe_preconditioner='mg' is built for ppe_iteration_method='cg'")
        if method == "cg" and self.preconditioner == "line_pcr":
            raise ValueError(
                "ppe_iteration_method='cg' supports preconditioner='jacobi' or 'none'; "
                "line_pcr is a nonsymmetric GMRES preconditioner.")
        if method == "cg" and self.preconditioner == "none":
            raise ValueError("unpreconditioned CG is not built; use 'jacobi' or 'mg'")
        pcr_stages = getattr(solver_cfg, "ppe_pcr_stages", 4)
        self.precond_pcr_stages = 4 if pcr_stages is None else int(pcr_stages)
        self.line_solver = str(getattr(solver_cfg, "ppe_line_solver", "pcr")).lower()
        if self.line_solver not in ("pcr", "thomas"):
            raise ValueError("ppe_line_solver must be 'pcr' or 'thomas'")
        self.precision = str(getattr(solver_cfg, "ppe_precision", "fp64")).lower()
        if self.precision not in ("fp64", "mixed"):
            raise ValueError("ppe_precision must be 'fp64' or 'mixed'")
        if self.precision == "mixed" and method != "cg":
            raise ValueError("ppe_precision='mixed' is built for ppe_iteration_method='cg'")
        self.check_every = int(getattr(solver_cfg, "ppe_check_every", 32))
        cyc = str(getattr(solver_cfg, "ppe_mg_cycle", "W")).upper()
        if cyc not in ("V", "W"):
            raise ValueError("ppe_mg_cycle must be 'V' or 'W'")
        self.mg_gamma = 1 if cyc == "V" else 2
        self.mg_smooth = int(getattr(solver_cfg, "ppe_mg_smooth", 2))
        self.mg_coarse_sweeps = int(getattr(solver_cfg, "ppe_mg_coarse_sweeps", 20))
        self.mg_w_levels = int(getattr(solver_cfg, "ppe_mg_w_levels", 2))
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
        if rho is not None:  # the reference prepares the operator on every solve
            self.prepare_operator(rho)
        elif self._rho is None:
            raise RuntimeError("no density set")
        rep = nat.SolveReport()
        if self.iteration_method == "gmres":
            gopts = self._gmres_opts(use_x0)
            status = self._lib.tpf_ppe_solve_gmres(self._handle, ptr(rhs), ptr(p),
                                                   C.byref(gopts), C.byref(rep))
        elif self.preconditioner == "mg":
            mopts = nat.MgOpts(tol=self.tol, maxiter=self.maxiter, use_x0=1 if use_x0 else 0,
                               nu=self.mg_smooth, gamma=self.mg_gamma,
                               coarse_sweeps=self.mg_coarse_sweeps, max_levels=0,
                               w_levels=self.mg_w_levels)
            status = self._lib.tpf_ppe_solve_mg(self._handle, ptr(rhs), ptr(p), C.byref(mopts),
                                                C.byref(rep))
        else:
            opts = nat.SolveOpts(tol=self.tol, maxiter=self.maxiter, precond=1,
                                 check_every=self.check_every, use_x0=1 if use_x0 else 0,
                                 mixed=1 if self.precision == "mixed" else 0,
                                 inner_tol=self.inner_tol)
            status = self._lib.tpf_ppe_solve(self._handle, ptr(rhs), ptr(p), C.byref(opts),
                                             C.byref(rep))
        key = f"ppe_low_order_{self.iteration_method}"
        self.last_diagnostics = {
            f"{key}_info": 0 if status == nat.TPF_OK else (
                self.maxiter if self.iteration_method == "gmres" else int(rep.iterations)),
            f"{key}_tol": self.tol,
            f"{key}_maxiter": self.maxiter,
            "iterations": int(rep.iterations),
            "relres": float(rep.relres),
            "true_relres": float(rep.true_relres),
            "refinements": int(rep.refinements),
            "precision": self.precision,
            "preconditioner": self.preconditioner,
        }
        if status != nat.TPF_OK:
            msg = self._lib.tpf_last_error().decode(errors="replace")
            if status == nat.TPF_ERR_RUNTIME:
                raise RuntimeError(
                    f"{type(self).__name__} did not converge cleanly: {msg}; no alternate "
                    "PPE solver was applied.")
            nat.check(status)
        return self.last_diagnostics

    def _gmres_opts(self, use_x0: bool = False) -> "nat.GmresOpts":
        pre = {"none": 0, "jacobi": 1, "line_pcr": 2}[self.preconditioner]
        stages = 0 if self.line_solver == "thomas" else max(1, self.precond_pcr_stages)
        return nat.GmresOpts(tol=self.tol, maxiter=self.maxiter, restart=self.restart,
                             precond=pre, pcr_stages=stages, use_x0=1 if use_x0 else 0,
                             c_tau=self.c_tau)

    def apply_preconditioner(self, r):
        """GMRES-form preconditioner z = M^-1 r with z[pin] = 0
        (fvm_matrixfree.py:302-330)."""
        if self._rho is None:
            raise RuntimeError("prepare_operator(rho) must be called first")
        if self.iteration_method != "gmres" or self.preconditioner == "none":
            raise RuntimeError("no GMRES-form preconditioner is configured")
        io = Handoff(r)
        t = io.to_device(r)
        out = torch.empty_like(t)
        nat.check(self._lib.tpf_ppe_gmres_precond(self._handle, ptr(t), ptr(out),
                                                  C.byref(self._gmres_opts())))
        return io.back(out)

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


class PPEBatch:
    """Several ``PPESolverFVMMatrixFree`` solvers (CG form) solved at once;
    each problem's result equals its own ``solve_into`` bit for bit.  The
    reference solves such problems one after another
    (``fvm_matrixfree.py:150-260`` per simulation).

    Two modes, chosen from the solvers:

    - Jacobi-PCG in FP64 (``ppe_preconditioner="jacobi"``): one launch per
      PCG step covers every problem (``tpf_ppe_batch_solve``).
    - multigrid FCG (``ppe_preconditioner="mg"``): every problem keeps its
      own stream and hierarchy, and the solves advance together round by
      round (``tpf_ppe_solve_mg_many``).

    The solvers must share tol, maxiter and the rest of their iteration
    options.  Densities are set per solver with ``prepare_operator`` before
    ``solve_into``."""

    def __init__(self, solvers):
        solvers = list(solvers)
        if not solvers:
            raise ValueError("PPEBatch needs at least one solver")
        s0 = solvers[0]
        for s in solvers:
            if not isinstance(s, PPESolverFVMMatrixFree):
                raise TypeError("PPEBatch takes PPESolverFVMMatrixFree solvers")
        self.mode = "mg" if s0.preconditioner == "mg" else "jacobi"
        for s in solvers:
            if self.mode == "mg":
                if s.preconditioner != "mg":
                    raise ValueError("solvers of one PPEBatch must all use ppe_preconditioner"
                                     "='mg' or none of them")
                key = lambda t: (t.tol, t.maxiter, t.mg_smooth, t.mg_gamma,  # noqa: E731
                                 t.mg_coarse_sweeps, t.mg_w_levels)
                what = ("pseudo_tol, pseudo_maxiter and the ppe_mg_* cycle options")
            else:
                if s.iteration_method != "cg" or s.preconditioner != "jacobi" or \
                        s.precision != "fp64":
                    raise ValueError("PPEBatch runs the CG form with Jacobi preconditioning in "
                                     "FP64 (ppe_iteration_method='cg', ppe_preconditioner="
                                     "'jacobi', ppe_precision='fp64') or multigrid FCG "
                                     "(ppe_preconditioner='mg')")
                key = lambda t: (t.tol, t.maxiter, t.check_every)  # noqa: E731
                what = "pseudo_tol, pseudo_maxiter and ppe_check_every"
            if key(s) != key(s0):
                raise ValueError(f"solvers of one PPEBatch must share {what}")
        if len({id(s) for s in solvers}) != len(solvers):
            raise ValueError("a solver appears twice in one PPEBatch")
        self.solvers = solvers
        self._lib = nat.lib()
        self._handle = None
        if self.mode == "jacobi":
            h = C.c_void_p()
            nat.check(self._lib.tpf_ppe_batch_create(
                C.byref(h), nat.parr([s._handle.value for s in solvers]), len(solvers),
                stream_handle()))
            self._handle = h
        self.last_reports: list = []

    def __len__(self):
        return len(self.solvers)

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            self._lib.tpf_ppe_batch_destroy(h)
            self._handle = None

    def solve_into(self, rhs, p, use_x0: bool = False, raise_on_failure: bool = True):
        """Solves problem i for p[i] (CUDA tensors of solver i's grid shape;
        rhs[i] = p[i] = None leaves problem i out of this call).  Returns the
        per-problem diagnostics (the keys of ``PPESolverFVMMatrixFree.solve_into``
        plus ``converged``) and stores them in each solver's
        ``last_diagnostics``.  With ``raise_on_failure`` a problem that ended
        above tol raises RuntimeError naming it; otherwise its diagnostics
        carry ``converged=False`` and the caller decides."""
        n = len(self.solvers)
        if len(rhs) != n or len(p) != n:
            raise ValueError(f"expected {n} right-hand sides and solutions")
        for s, r, x in zip(self.solvers, rhs, p):
            if (r is None) != (x is None):
                raise ValueError("rhs[i] and p[i] must both be given or both be None")
            if r is None:
                continue
            for t in (r, x):
                if tuple(t.shape) != s.shape or t.dtype != torch.float64 or not t.is_cuda \
                        or not t.is_contiguous():
                    raise ValueError(f"batch tensors must be contiguous CUDA float64 of "
                                     f"shape {s.shape}")
            if s._rho is None:
                raise RuntimeError("prepare_operator(rho) must be called on every solver")
        s0 = self.solvers[0]
        reps = (nat.SolveReport * n)()
        if self.mode == "mg":
            on = [i for i in range(n) if rhs[i] is not None]
            sub = (nat.SolveReport * max(1, len(on)))()
            status = nat.TPF_OK
            if on:
                mopts = nat.MgOpts(tol=s0.tol, maxiter=s0.maxiter, use_x0=1 if use_x0 else 0,
                                   nu=s0.mg_smooth, gamma=s0.mg_gamma,
                                   coarse_sweeps=s0.mg_coarse_sweeps, max_levels=0,
                                   w_levels=s0.mg_w_levels)
                status = self._lib.tpf_ppe_solve_mg_many(
                    nat.parr([self.solvers[i]._handle.value for i in on]), len(on),
                    nat.parr([ptr(rhs[i]) for i in on]), nat.parr([ptr(p[i]) for i in on]),
                    C.byref(mopts), sub)
            for j, i in enumerate(on):
                reps[i] = sub[j]
        else:
            opts = nat.SolveOpts(tol=s0.tol, maxiter=s0.maxiter, precond=1,
                                 check_every=s0.check_every, use_x0=1 if use_x0 else 0,
                                 mixed=0, inner_tol=0.0)
            status = self._lib.tpf_ppe_batch_solve(self._handle, nat.parr([ptr(t) for t in rhs]),
                                                   nat.parr([ptr(t) for t in p]), C.byref(opts),
                                                   reps)
        if status not in (nat.TPF_OK, nat.TPF_ERR_RUNTIME):
            nat.check(status)
        out = []
        for s, r, t in zip(self.solvers, reps, rhs):
            if t is None:
                out.append(None)
                continue
            ok = bool(r.converged) if self.mode == "mg" else r.relres <= s.tol
            d = {"ppe_low_order_cg_info": 0 if ok else int(r.iterations),
                 "ppe_low_order_cg_tol": s.tol, "ppe_low_order_cg_maxiter": s.maxiter,
                 "iterations": int(r.iterations), "relres": float(r.relres),
                 "true_relres": float(r.true_relres), "refinements": int(r.refinements),
                 "precision": s.precision, "preconditioner": s.preconditioner,
                 "converged": ok}
            s.last_diagnostics = d
            out.append(d)
        self.last_reports = out
        if status == nat.TPF_ERR_RUNTIME and raise_on_failure:
            msg = self._lib.tpf_last_error().decode(errors="replace")
            raise RuntimeError(f"PPEBatch did not converge cleanly: {msg}; no alternate PPE "
                               "solver was applied.")
        return out


class PPESolverFCCDMatrixFree(IPPESolver):
    """Matrix-free FCCD PPE on the GPU: L p = D_f[(1/rho)_f G_f p] with the
    FCCD face gradient (ppe/fccd_matrixfree.py:108-600, ccd/fccd.py:243-277),
    solved by SciPy-semantics GMRES on the device.

    Coefficient schemes ``"phase_density"`` (the class default) and
    ``"phase_separated"`` (the pipeline default, ns_pipeline.py:195-196) with
    the interface couplings ``"none"``, ``"jump_decomposition"`` and
    ``"affine_jump"`` (the pipeline default); interface data through
    ``set_interface_jump_context`` (nodal_cut_face route).  Built for
    ``boundary_face_space = "full_face"``, the legacy
    ``scalar_operator_pairing``, uniform grids, wall or periodic axes; the
    other face spaces, pairings and curvature routes raise
    ``NotImplementedError``; invalid names raise ``ValueError`` as in the
    reference (fccd_matrixfree.py:141-186).

    ``ppe_preconditioner`` takes the reference's "none" (default) and
    "jacobi", plus this build's "mg" (wall grids): one FP32 W(2,2) cycle of the
    FVM operator's multigrid hierarchy per GMRES step, the low-order operator
    as preconditioner of the high-order one (the Krylov form of the
    reference's defect correction, ppe/defect_correction.py:212-371)."""

    scheme_names = ("fccd_iterative",)
    _scheme_aliases = {"fccd_matrixfree": "fccd_iterative", "fccd": "fccd_iterative"}

    def __init__(self, backend, config, grid, fccd=None, *, bc_spec=None, bc_type=None):
        require_uniform(grid, "PPESolverFCCDMatrixFree")
        self.backend = backend
        self.xp = np
        self.grid = grid
        self.ndim = grid.ndim
        self.fccd = fccd
        if bc_type is None:
            bc_type = getattr(fccd, "bc_type", None)
        if bc_type is None:
            bc_type = getattr(getattr(config, "numerics", None), "bc_type", "wall")
        self.bc_type = bc_type
        self._kinds = axis_kinds(bc_type, self.ndim)
        solver_cfg = getattr(config, "solver", config)
        self.tol = float(getattr(solver_cfg, "pseudo_tol", 1.0e-8))
        self.maxiter = int(getattr(solver_cfg, "pseudo_maxiter", 500))
        restart = getattr(solver_cfg, "ppe_restart", None)
        self.restart = int(restart) if restart is not None else None
        self.preconditioner = str(getattr(solver_cfg, "ppe_preconditioner", "none")).strip().lower()
        self.coefficient_scheme = str(
            getattr(solver_cfg, "ppe_coefficient_scheme", "phase_density")).strip().lower()
        self.interface_coupling_scheme = str(
            getattr(solver_cfg, "ppe_interface_coupling_scheme", "none")).strip().lower()
        contract = str(getattr(solver_cfg, "pressure_force_contract", None)
                       or "raw_compact_gradient").strip().lower()
        pairing = str(getattr(solver_cfg, "scalar_operator_pairing", None)
                      or "legacy").strip().lower()
        pairing = {"raw": "legacy", "old": "legacy", "certified": "require_certified",
                   "require": "require_certified", "variational": "variational_operator",
                   "l_var": "variational_operator"}.get(pairing, pairing)
        face_space = str(getattr(solver_cfg, "boundary_face_space", None)
                         or "full_face").strip().lower().replace("-", "_")
        face_space = {"full": "full_face", "unconstrained": "full_face"}.get(face_space,
                                                                            face_space)
        if pairing == "require_certified" and contract in ("variational_adjoint", "variational",
                                                           "adjoint", "pressure_adjoint"):
            raise RuntimeError(
                "scalar_operator_pairing='require_certified' is fail-closed for FCCD until "
                "PPE.apply(p)==D(Gp) certification is implemented; use "
                "scalar_operator_pairing='variational_operator' to select L_var=D(Gp).")
        if self.coefficient_scheme not in {"phase_density", "phase_separated"}:
            raise ValueError("FCCD PPE supports ppe_coefficient_scheme="
                             "'phase_density'|'phase_separated'")
        if self.interface_coupling_scheme not in {"none", "jump_decomposition", "affine_jump"}:
            raise ValueError("FCCD PPE supports ppe_interface_coupling_scheme="
                             "'none'|'jump_decomposition'|'affine_jump'")
        if self.coefficient_scheme == "phase_density" and self.interface_coupling_scheme != "none":
            raise ValueError("phase_density PPE requires interface_coupling='none'")
        if self.preconditioner not in {"jacobi", "none", "mg"}:
            raise ValueError("FCCD PPE supports preconditioner='jacobi'|'none' (and this "
                             "build's 'mg')")
        if self.preconditioner == "mg" and any(k != "wall" for k in self._kinds):
            raise ValueError("ppe_preconditioner='mg' needs wall axes")
        if pairing == "variational_operator":
            raise NotImplementedError(
                "this build's FCCD PPE covers the legacy scalar_operator_pairing")
        if face_space != "full_face":
            raise NotImplementedError("this build's FCCD PPE covers boundary_face_space="
                                      "'full_face'")
        if not getattr(grid, "uniform", True):
            raise ValueError("this build's FCCD PPE needs a uniform grid")
        self.shape = tuple(int(n) + 1 for n in grid.N)
        if bc_spec is not None:
            self._pin_dof = int(bc_spec.pin_dof)
        else:  # fccd_matrixfree.py:187-192
            self._pin_dof = int(np.ravel_multi_index(tuple(n // 2 for n in grid.N), self.shape))
        self._pin_dofs = (self._pin_dof,)
        self.last_base_pressure = None
        self.last_diagnostics: dict = {}
        self._lib = nat.lib()
        self._handle = None
        self._rho = None
        self._interface = None   # device tensors of the interface data (kept alive)
        self._create()

    _SCHEMES = {"phase_density": 0, "phase_separated": 1}
    _COUPLINGS = {"none": 0, "jump_decomposition": 1, "affine_jump": 2}

    def _create(self):
        require_gpu()
        h = C.c_void_p()
        bc = [nat.BC_PERIODIC if k == "periodic" else nat.BC_WALL for k in self._kinds]
        nat.check(self._lib.tpf_fccd_ppe_create(C.byref(h), self.ndim, nat.i64arr(self.shape),
                                                nat.darr(self.grid.L), nat.iarr(bc),
                                                int(self._pin_dof), stream_handle()))
        self._handle = h
        nat.check(self._lib.tpf_fccd_ppe_set_scheme(
            h, self._SCHEMES[self.coefficient_scheme],
            self._COUPLINGS[self.interface_coupling_scheme]))
        if self._interface is not None:
            self._push_interface()

    # -- interface data (fccd_matrixfree.py:230-304) --------------------------------
    def set_interface_jump_context(self, *, psi, kappa=None, sigma: float = 0.0,
                                   psi_previous=None, pressure_jump_gas_minus_liquid=None,
                                   phase_threshold: float = 0.5,
                                   face_curvature_method: str = "nodal_cut_face",
                                   transport_variational_nodal_covector=None,
                                   transport_variational_psi=None,
                                   transport_variational_previous_surface_energy=None) -> None:
        """Young-Laplace j_gl = p_gas - p_liquid = -sigma kappa_lg (kappa at the
        psi cut point on cut faces; J = j (1 - psi) for the jump
        decomposition), or a given nodal jump (face-averaged; the jump
        decomposition stays off, as the reference's sigma = 0 context).
        Only the nodal_cut_face curvature route is built."""
        if face_curvature_method != "nodal_cut_face":
            raise NotImplementedError(
                f"face_curvature_method={face_curvature_method!r} is not built; this build "
                "covers 'nodal_cut_face'")
        if any(x is not None for x in (transport_variational_nodal_covector,
                                       transport_variational_psi,
                                       transport_variational_previous_surface_energy)):
            raise NotImplementedError("the transport-variational jump routes are not built")
        io = Handoff(psi)
        psi_t = io.to_device(psi).clone()
        if tuple(psi_t.shape) != self.shape:
            raise ValueError(f"psi shape {tuple(psi_t.shape)} != grid.shape {self.shape}")
        if pressure_jump_gas_minus_liquid is None:
            if kappa is None:
                raise ValueError("kappa is required without pressure_jump_gas_minus_liquid")
            self._interface = {"psi": psi_t, "kappa": io.to_device(kappa).clone(), "jump": None,
                               "sigma": float(sigma), "threshold": float(phase_threshold)}
        else:
            self._interface = {"psi": psi_t, "kappa": None,
                               "jump": io.to_device(pressure_jump_gas_minus_liquid).clone(),
                               "sigma": 0.0, "threshold": float(phase_threshold)}
        self._push_interface()

    def _push_interface(self) -> None:
        c = self._interface
        nat.check(self._lib.tpf_fccd_ppe_set_interface(
            self._handle, ptr(c["psi"]), ptr(c["kappa"]), ptr(c["jump"]), c["sigma"],
            c["threshold"]))

    def clear_interface_jump_context(self) -> None:
        self._interface = None
        nat.check(self._lib.tpf_fccd_ppe_set_interface(self._handle, None, None, None, 0.0, 0.5))

    @property
    def _phase_threshold(self):
        g, t = C.c_int(), C.c_double()
        nat.check(self._lib.tpf_fccd_ppe_gauge_info(self._handle, C.byref(g), C.byref(t)))
        return float(t.value) if g.value else None

    def _uses_phase_mean_gauge(self) -> bool:
        return self.coefficient_scheme == "phase_separated" and self._phase_threshold is not None

    def face_coefficients(self, axis: int) -> torch.Tensor:
        """alpha_f of `axis` at each face's lower node (PPESolverFCCDMatrixFree.
        _coeff_face; the last node of a line holds 0)."""
        if self._rho is None:
            raise RuntimeError("prepare_operator(rho) must be called first")
        out = torch.empty(self.shape, dtype=torch.float64, device="cuda")
        nat.check(self._lib.tpf_fccd_ppe_face_coeff(self._handle, int(axis), ptr(out)))
        return out

    def apply_interface_jump(self, pressure):
        """p + J of the jump decomposition, else p (fccd_matrixfree.py:644-654)."""
        io = Handoff(pressure)
        t = io.to_device(pressure)
        out = torch.empty_like(t)
        nat.check(self._lib.tpf_fccd_ppe_interface_jump(self._handle, ptr(t), ptr(out)))
        return io.back(out)

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            self._lib.tpf_fccd_ppe_destroy(h)
            self._handle = None

    def update_grid(self, grid=None) -> None:
        if grid is not None:
            self.grid = grid
            self.ndim = grid.ndim
            self._kinds = axis_kinds(self.bc_type, self.ndim)
            self.shape = tuple(int(n) + 1 for n in grid.N)
            self._pin_dof = int(np.ravel_multi_index(tuple(n // 2 for n in grid.N), self.shape))
            self._pin_dofs = (self._pin_dof,)
            self._images_np = None
        self.__del__()
        self._rho = None
        self._create()

    def invalidate_cache(self) -> None:
        self._rho = None

    # -- operator --------------------------------------------------------------
    def prepare_operator(self, rho) -> None:
        """Keep rho on the device and build the Jacobi inverse
        (fccd_matrixfree_lifecycle.py:44-96)."""
        io = Handoff(rho)
        t = io.to_device(rho)
        if tuple(t.shape) != self.shape:
            raise ValueError(f"rho shape {tuple(t.shape)} != grid.shape {self.shape}")
        self._rho = t
        nat.check(self._lib.tpf_fccd_ppe_set_density(self._handle, ptr(t)))

    def apply_into(self, p: torch.Tensor, out: torch.Tensor) -> None:
        """out = L p on CUDA tensors (asynchronous)."""
        if self._rho is None:
            raise RuntimeError("prepare_operator(rho) must be called before apply(p)")
        nat.check(self._lib.tpf_fccd_ppe_apply(self._handle, ptr(p), ptr(out)))

    def apply(self, p):
        """L p with the pin row and the periodic image rows
        (fccd_matrixfree.py:378-397)."""
        io = Handoff(p)
        t = io.to_device(p)
        out = torch.empty_like(t)
        self.apply_into(t, out)
        return io.back(out)

    # -- solve -----------------------------------------------------------------
    def _gmres_opts(self, use_x0: bool) -> "nat.GmresOpts":
        # SciPy's gmres takes restart=None as 20 (gmres_helpers.py:16-42)
        restart = 20 if self.restart is None else self.restart
        return nat.GmresOpts(tol=self.tol, maxiter=self.maxiter, restart=restart,
                             precond={"none": 0, "jacobi": 1, "mg": 2}[self.preconditioner],
                             pcr_stages=0, use_x0=1 if use_x0 else 0, c_tau=0.0)

    def system_rhs(self, rhs: torch.Tensor) -> torch.Tensor:
        """The right-hand side GMRES iterates on (fccd_matrixfree.py:506-517):
        rhs - L J (jump decomposition), + the affine drive, the per-phase
        mean projection or rhs[pin] = 0, zero periodic image rows."""
        if self._rho is None:
            raise RuntimeError("prepare_operator(rho) must be called first")
        b = torch.empty_like(rhs)
        nat.check(self._lib.tpf_fccd_ppe_system_rhs(self._handle, ptr(rhs), ptr(b)))
        return b

    # -- gauge helpers used by defect correction (fccd_matrixfree.py:300-376,
    #    682-706); NumPy arrays or CUDA tensors -----------------------------------
    def _image_rows(self, like):
        """(images, sources) flat indices: the first periodic axis on which a
        node sits at index N maps it to index 0 (fccd_matrixfree.py:300-345)."""
        if not any(k == "periodic" for k in self._kinds):
            return None
        if getattr(self, "_images_np", None) is None:
            table = {}
            idx = np.indices(self.shape).reshape(self.ndim, -1)
            for a, kind in enumerate(self._kinds):
                if kind != "periodic":
                    continue
                sel = idx[a] == self.shape[a] - 1
                src = idx[:, sel].copy()
                src[a] = 0
                for i, j in zip(np.ravel_multi_index(idx[:, sel], self.shape).tolist(),
                                np.ravel_multi_index(src, self.shape).tolist()):
                    table.setdefault(i, j)
            img = np.array(sorted(table), dtype=np.int64)
            self._images_np = (img, np.array([table[i] for i in img], dtype=np.int64))
            self._images_t = None
        if isinstance(like, torch.Tensor):
            if self._images_t is None or self._images_t[0].device != like.device:
                self._images_t = tuple(torch.from_numpy(a).to(like.device)
                                       for a in self._images_np)
            return self._images_t
        return self._images_np

    def _sync_periodic_images(self, arr):
        rows = self._image_rows(arr)
        if rows is None:
            return arr
        out = arr.copy() if isinstance(arr, np.ndarray) else arr.clone()
        flat = out.reshape(-1)
        flat[rows[0]] = flat[rows[1]]
        return out

    def _zero_periodic_image_rows(self, arr):
        rows = self._image_rows(arr)
        if rows is not None:
            arr.reshape(-1)[rows[0]] = 0.0
        return arr

    def _project_rhs_compatibility(self, rhs, *, record_stats: bool = True):
        """The pinned gauge: a copy with rhs[pin] = 0
        (fccd_matrixfree_helpers.py:337-375); the per-phase mean gauge is
        applied on the device by system_rhs and the solve."""
        if self._rho is not None and self._uses_phase_mean_gauge():
            raise NotImplementedError("the per-phase mean gauge projects on the device "
                                      "(system_rhs); this host helper covers the pinned gauge")
        out = rhs.copy() if isinstance(rhs, np.ndarray) else rhs.clone()
        out.reshape(-1)[self._pin_dof] = 0.0
        return out

    def solve_into(self, rhs: torch.Tensor, rho, p: torch.Tensor, use_x0: bool = False) -> dict:
        """Zero-copy solve on CUDA tensors; p holds p_init when use_x0."""
        if rho is not None:  # the reference prepares the operator on every solve
            self.prepare_operator(rho)
        elif self._rho is None:
            raise RuntimeError("no density set")
        rep = nat.SolveReport()
        status = self._lib.tpf_fccd_ppe_solve_gmres(self._handle, ptr(rhs), ptr(p),
                                                    C.byref(self._gmres_opts(use_x0)),
                                                    C.byref(rep))
        if status not in (nat.TPF_OK, nat.TPF_ERR_RUNTIME):
            nat.check(status)
        msg = self._lib.tpf_last_error().decode(errors="replace") if status else ""
        b = self.system_rhs(rhs)
        base = p
        if self.interface_coupling_scheme == "jump_decomposition" and \
                self._interface is not None and self._interface["sigma"] > 0.0:
            # the solve returned sync(p + J): the residual is the base pressure's
            # (image nodes of the base equal their sources)
            jp = torch.zeros_like(p)
            nat.check(self._lib.tpf_fccd_ppe_interface_jump(self._handle, ptr(jp), ptr(jp)))
            base = self._sync_periodic_images(p - jp)
        res = torch.empty_like(p)
        self.apply_into(base.contiguous(), res)
        res -= b
        rhs_l2 = float(torch.linalg.vector_norm(b))
        res_l2 = float(torch.linalg.vector_norm(res))
        mean_gauge = self._uses_phase_mean_gauge()
        self.last_diagnostics = {
            "ppe_phase_count": 2.0 if mean_gauge else 1.0,
            "ppe_pin_count": 0.0 if mean_gauge else 1.0,
            "ppe_mean_gauge": float(mean_gauge),
            "ppe_interface_coupling_jump": float(
                self.interface_coupling_scheme == "jump_decomposition"),
            "ppe_interface_coupling_affine_jump": float(
                self.interface_coupling_scheme == "affine_jump"),
            "ppe_linear_info": float(0 if status == nat.TPF_OK else self.maxiter),
            "ppe_linear_converged": float(status == nat.TPF_OK),
            "ppe_linear_rhs_l2": rhs_l2,
            "ppe_linear_residual_l2": res_l2,
            "ppe_linear_relative_l2": res_l2 / rhs_l2 if rhs_l2 > 0 else res_l2,
            "ppe_linear_residual_linf": float(res.abs().max()) if res.numel() else 0.0,
            "iterations": int(rep.iterations),
            "restart_cycles": int(rep.refinements),
            "true_relres": float(rep.true_relres),
        }
        if status != nat.TPF_OK:
            d = self.last_diagnostics
            raise RuntimeError(
                "PPESolverFCCDMatrixFree did not converge: "
                f"info={self.maxiter}, relative_l2={d['ppe_linear_relative_l2']:.3e}, "
                f"linf={d['ppe_linear_residual_linf']:.3e}, maxiter={self.maxiter} ({msg})")
        self.last_base_pressure = base.clone()
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


# -- factory (ppe/factory.py:31-120) ------------------------------------------
_SOLVER_REGISTRY: Dict[str, Callable] = {}


def register_ppe_solver(name: str, factory_fn: Callable) -> None:
    _SOLVER_REGISTRY[name] = factory_fn


def _create_fvm_matrixfree(config, backend, grid, ccd, bc_spec):
    return PPESolverFVMMatrixFree(backend, config, grid,
                                  bc_type=config.numerics.bc_type, bc_spec=bc_spec)


def _create_fccd_matrixfree(config, backend, grid, ccd, bc_spec):
    return PPESolverFCCDMatrixFree(backend, config, grid, None, bc_spec=bc_spec,
                                   bc_type=config.numerics.bc_type)


register_ppe_solver("fvm_iterative", _create_fvm_matrixfree)
register_ppe_solver("fvm_matrixfree", _create_fvm_matrixfree)
for _name in ("fccd_iterative", "fccd_matrixfree", "fccd"):
    register_ppe_solver(_name, _create_fccd_matrixfree)


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
