"""Defect correction: a high-order PPE operator corrected by a low-order solve.

Mirrors ``PPESolverDefectCorrection`` (ppe/defect_correction.py:28-439): a
base solve with the low-order solver L_L, then up to ``max_corrections``
passes of
    r = b - L_H p;  stop when ||r|| <= tol * max(||b||, 1);
    d = L_L^-1 r;   p += alpha d,
with alpha from the reference's line search on ||r - alpha L_H d||: the
minimiser alpha* first, then relaxation * 2^-k (k < 12); the first
candidate that lowers the residual is taken, and the loop stops when none
does.  The gauge (pin, periodic images) and the right-hand-side
compatibility come from the operator, as in the reference.

In this build L_H is typically ``PPESolverFCCDMatrixFree`` and L_L
``PPESolverFVMMatrixFree`` (multigrid-FCG): both run on the device, and the
wrapper's own work is three reductions and one update per correction, done
on the arrays the solvers return (CUDA tensors for CUDA inputs, NumPy arrays
for NumPy inputs).  ``PPESolverFCCDMatrixFree(ppe_preconditioner="mg")`` is
the Krylov-accelerated form of the same pairing.
"""

from __future__ import annotations

import math

import numpy as np

from .ppe import IPPESolver


def _copy(x):
    return x.copy() if isinstance(x, np.ndarray) else x.clone()


def _dot(a, b) -> float:
    return float((a * b).sum())


_OPERATOR_KEYS = ("grid", "coefficient_scheme", "interface_coupling_scheme", "bc_type")


def _operator_signature(solver):
    """What fixes the linear operator a PPE solver discretises: its class and
    the grid / coefficient / coupling / boundary settings."""
    return type(solver), tuple(getattr(solver, k, None) for k in _OPERATOR_KEYS)


def _can_collapse_same_operator(base_solver, operator) -> bool:
    """True when the low-order base solver and the high-order operator are the
    same discretisation (a defect correction that would iterate on itself);
    the rule of ppe/defect_correction.py:431-439."""
    (kind_a, keys_a), (kind_b, keys_b) = (_operator_signature(base_solver),
                                          _operator_signature(operator))
    same = kind_a is kind_b and all(a is b or a == b for a, b in zip(keys_a, keys_b))
    return same and hasattr(base_solver, "solve") and hasattr(operator, "apply")


class PPESolverDefectCorrection(IPPESolver):
    """Outer residual-correction loop for a matrix-free PPE operator."""

    def __init__(self, backend, grid, base_solver, operator, *, max_corrections: int = 3,
                 tolerance: float = 1.0e-8, relaxation: float = 1.0,
                 fail_on_nonconvergence: bool = False) -> None:
        if max_corrections <= 0:
            raise ValueError("max_corrections must be > 0")
        if tolerance <= 0.0:
            raise ValueError("tolerance must be > 0")
        if relaxation <= 0.0:
            raise ValueError("relaxation must be > 0")
        if not hasattr(operator, "prepare_operator") or not hasattr(operator, "apply"):
            raise TypeError("operator must provide prepare_operator(rho) and apply(p)")
        if _can_collapse_same_operator(base_solver, operator):
            raise ValueError(
                "PPESolverDefectCorrection requires a distinct low-order base solver L_L; "
                "wrapping the same high-order operator would bypass the paper's "
                "defect-correction contract.")
        self.backend = backend
        self.grid = grid
        self.base_solver = base_solver
        self.operator = operator
        self.max_corrections = int(max_corrections)
        self.tolerance = float(tolerance)
        self.relaxation = float(relaxation)
        self.fail_on_nonconvergence = bool(fail_on_nonconvergence)
        self._pin_dof = operator._pin_dof
        self._pin_dofs = getattr(operator, "_pin_dofs", (self._pin_dof,))
        self.last_base_pressure = None
        self.last_diagnostics: dict = {}
        self.last_residual_history: list[float] = []
        self.last_stalled = False

    def update_grid(self, grid=None) -> None:
        if grid is not None:
            self.grid = grid
        self.base_solver.update_grid(self.grid)
        self.operator.update_grid(self.grid)
        self._pin_dof = self.operator._pin_dof
        self._pin_dofs = getattr(self.operator, "_pin_dofs", (self._pin_dof,))

    def invalidate_cache(self) -> None:
        self.base_solver.invalidate_cache()
        self.operator.invalidate_cache()

    # -- gauge and compatibility, delegated to the operator (:170-211) ----------
    def _enforce_periodic_pressure(self, p):
        sync = getattr(self.operator, "_sync_periodic_images", None)
        return sync(p) if callable(sync) else p

    def _enforce_periodic_rhs(self, rhs):
        zero = getattr(self.operator, "_zero_periodic_image_rows", None)
        return zero(rhs) if callable(zero) else rhs

    def _enforce_pressure_gauge(self, p):
        p = self._enforce_periodic_pressure(p)
        flat = p.reshape(-1)
        for dof in self._pin_dofs:
            flat[dof] = 0.0
        return self._enforce_periodic_pressure(p)

    def _enforce_rhs_compatibility(self, rhs):
        project = getattr(self.operator, "_project_rhs_compatibility", None)
        if callable(project):
            return self._enforce_periodic_rhs(project(rhs, record_stats=False))
        out = _copy(rhs)
        flat = out.reshape(-1)
        for dof in self._pin_dofs:
            flat[dof] = 0.0
        return self._enforce_periodic_rhs(out)

    # -- solve (:212-371) ------------------------------------------------------------
    def solve(self, rhs, rho, dt: float = 0.0, p_init=None):
        self.operator.prepare_operator(rho)
        if hasattr(self.base_solver, "prepare_operator"):
            self.base_solver.prepare_operator(rho)
        b = self._enforce_rhs_compatibility(rhs)
        p = self.base_solver.solve(b, rho, dt=dt, p_init=p_init)
        p = self._enforce_pressure_gauge(p)
        rhs_norm = None
        scale = None
        history: list[float] = []
        broke = False
        applied = 0
        final_norm = 0.0
        final_res = None
        for _ in range(self.max_corrections):
            r = b - self.operator.apply(p)
            r = self._enforce_rhs_compatibility(r)
            rn = math.sqrt(_dot(r, r))
            if scale is None:
                rhs_norm = math.sqrt(_dot(b, b))
                scale = max(rhs_norm, 1.0)
            history.append(rn)
            if rn <= self.tolerance * scale:
                broke, final_res, final_norm = True, r, rn
                break
            d = self.base_solver.solve(r, rho, dt=dt, p_init=None)
            d = self._enforce_pressure_gauge(d)
            Ad = self._enforce_rhs_compatibility(self.operator.apply(d))
            rr = _dot(r, r)
            num = _dot(r, Ad)
            den = _dot(Ad, Ad)
            alpha_opt = num / (den if den > 0.0 else 1.0)
            accepted = False
            for alpha in [alpha_opt] + [self.relaxation * 0.5 ** k for k in range(12)]:
                trial = rr - 2.0 * alpha * num + alpha * alpha * den
                if alpha > 0.0 and trial < rr:
                    p = self._enforce_pressure_gauge(p + alpha * d)
                    accepted = True
                    applied += 1
                    break
            if not accepted:
                final_res, final_norm = r, rn
                break
        if not broke:
            final_res = self._enforce_rhs_compatibility(b - self.operator.apply(p))
            final_norm = math.sqrt(_dot(final_res, final_res))
            history.append(final_norm)
            broke = final_norm <= self.tolerance * scale
        final_linf = float(abs(final_res).max()) if final_res is not None else 0.0
        self.last_residual_history = history
        self.last_stalled = not broke
        self.last_base_pressure = _copy(p)
        diag = dict(getattr(self.operator, "last_diagnostics", {}) or {})
        diag.update({
            "ppe_dc_iterations": float(applied),
            "ppe_dc_max_corrections": float(self.max_corrections),
            "ppe_dc_tolerance": float(self.tolerance),
            "ppe_dc_relaxation": float(self.relaxation),
            "ppe_dc_rhs_l2": float(rhs_norm),
            "ppe_dc_initial_residual_l2": float(history[0]),
            "ppe_dc_final_residual_l2": float(final_norm),
            "ppe_dc_final_relative_l2": float(final_norm / rhs_norm if rhs_norm > 0.0
                                              else final_norm),
            "ppe_dc_final_residual_linf": final_linf,
            "ppe_dc_converged": float(not self.last_stalled),
            "ppe_dc_stalled": float(self.last_stalled),
        })
        self.last_diagnostics = diag
        if self.last_stalled and self.fail_on_nonconvergence:
            raise RuntimeError(
                "PPESolverDefectCorrection did not converge: "
                f"relative_l2={final_norm / scale:.3e}, linf={final_linf:.3e}, "
                f"max_corrections={self.max_corrections}")
        return self._enforce_periodic_pressure(p)
