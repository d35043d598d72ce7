"""The legacy two-phase simulation on the GPU: ``SimulationBuilder`` and
``TwoPhaseSimulation`` with the reference's surface
(``simulation/builder.py:46-117``, ``simulation/_core.py:61-224``,
``simulation/legacy_*.py``).

``SimulationBuilder(cfg).build()`` wires, as the legacy builder does
(``simulation/legacy_builder_runtime.py:34-199``): the wall / periodic
``CCDSolver``, ``DissipativeCCDAdvection`` with mass correction, the split
reinitialisation, ``CurvatureCalculatorPsi``, the UCCD6 + Picard CN
``Predictor``, the FVM matrix-free PPE (``create_ppe_solver``), the
Rhie-Chow right-hand side, the CCD velocity corrector, the CFL calculator,
the wall boundary handler and the closest-point (or null) field extender.
Every field lives on the device as a float64 tensor; each step is a sequence
of HIP launches through ``libtpf`` plus the PPE solve.

``step_forward(dt)`` is the 7-step algorithm (_core.py:126-166); ``run``
and ``_step_with_retry`` keep the run loop's CFL time step, snapshot,
rollback and dt halving (legacy_run_loop.py:13-82).
"""

from __future__ import annotations

import ctypes as C
from typing import Callable, List, Optional

import numpy as np
import torch

from . import _native as nat
from .boundary import BoundarySpec, axis_kinds
from .ccd import CCDSolver
from .curvature import CurvatureCalculatorPsi
from .dccd import DissipativeCCDAdvection
from .device import ptr, require_gpu, stream_handle
from .grid import Grid, require_uniform
from .ns_terms import (BoundaryConditionHandler, CFLCalculator, ClosestPointExtender, FlowState,
                       NullFieldExtender, Predictor, VelocityCorrector)
from .ppe import create_ppe_solver
from .reinit import SplitReinitializer
from .rhie_chow import RhieChowInterpolator


def _to_device(value, shape) -> torch.Tensor:
    """A contiguous CUDA float64 copy of value, broadcast to shape."""
    if isinstance(value, torch.Tensor):
        t = value.to(device="cuda", dtype=torch.float64)
    else:
        t = torch.from_numpy(np.asarray(value, dtype=np.float64)).cuda()
    return t.expand(tuple(shape)).contiguous().clone()


class ScalarField:
    """core/field.py:18-45 on the device: ``data`` is a CUDA float64 tensor;
    assigning NumPy or torch values copies them in."""

    def __init__(self, grid):
        self.grid = grid
        self._data = torch.zeros(grid.shape, dtype=torch.float64, device="cuda")

    @property
    def data(self) -> torch.Tensor:
        return self._data

    @data.setter
    def data(self, value) -> None:
        self._data = _to_device(value, self.grid.shape)

    def __repr__(self) -> str:
        return f"ScalarField(shape={tuple(self._data.shape)})"


class VectorField:
    """core/field.py: one device component per axis."""

    def __init__(self, grid):
        self.grid = grid
        self._c = [torch.zeros(grid.shape, dtype=torch.float64, device="cuda")
                   for _ in range(grid.ndim)]

    def __getitem__(self, i):
        return self._c[i]

    def __setitem__(self, i, value):
        self._c[i] = _to_device(value, self.grid.shape)

    def __len__(self):
        return len(self._c)


class DiagnosticsReporter:
    """simulation/diagnostics.py:34-65: max |sum_a D1_a u_a| and
    vol(psi) = sum psi * prod(L/N), printed."""

    def __init__(self, backend, grid):
        require_uniform(grid, "DiagnosticsReporter")
        self.grid = grid

    def divergence_max(self, sim) -> float:
        from .fields import divergence
        div = divergence(sim.ccd, [sim.velocity[a] for a in range(self.grid.ndim)])
        out = C.c_double()
        w = torch.empty(3 * 1024 + 8, dtype=torch.float64, device="cuda")
        nat.check(nat.lib().tpf_ns_absmax(1, nat.parr([ptr(div)]), div.numel(), ptr(w),
                                          C.byref(out), stream_handle()))
        return out.value

    def volume(self, sim) -> float:
        """sum psi times the uniform cell volume prod(L/N): the nodal mass
        with full-h weights on every axis."""
        m = C.c_double()
        nd = self.grid.ndim
        nat.check(nat.lib().tpf_fields_mass(nd, nat.i64arr(self.grid.shape),
                                            nat.darr(self.grid.L), nat.iarr([nat.BC_WALL] * nd),
                                            ptr(sim.psi.data), C.byref(m), stream_handle()))
        return m.value

    def report(self, sim, dt: float) -> None:
        print(f"  t={sim.time:.5f}  dt={dt:.3e}  |∇·u|_∞={self.divergence_max(sim):.3e}  "
              f"vol(ψ)={self.volume(sim):.6f}")


class TwoPhaseSimulation:
    """The legacy two-phase solver; build it with ``SimulationBuilder``."""

    @classmethod
    def _from_components(cls, comp: dict) -> "TwoPhaseSimulation":
        obj = cls.__new__(cls)
        for k, v in comp.items():
            setattr(obj, k, v)
        g = obj.grid
        obj.psi, obj.rho, obj.mu = ScalarField(g), ScalarField(g), ScalarField(g)
        obj.kappa, obj.phi, obj.pressure = ScalarField(g), ScalarField(g), ScalarField(g)
        obj.velocity, obj.vel_star = VectorField(g), VectorField(g)
        obj.time, obj.step = 0.0, 0
        obj._rho_l, obj._rho_g = 1.0, obj.config.fluid.rho_ratio
        obj._mu_l, obj._mu_g = 1.0, obj.config.fluid.mu_ratio
        obj._needs_phi = not obj._field_ext.is_null_extender
        obj._nf_work = torch.empty(8 + 64, dtype=torch.float64, device="cuda")
        return obj

    # -- public API (_core.py:88-166) ------------------------------------------
    def run(self, t_end: Optional[float] = None, output_interval: int = 10,
            verbose: bool = True, callback: Optional[Callable] = None) -> None:
        """legacy_run_loop.py:13-42."""
        if t_end is None:
            t_end = self.config.numerics.t_end
        self._update_properties()
        self._update_curvature()
        while self.time < t_end:
            dt_nominal = self.cfl_calc.compute(
                [self.velocity[a] for a in range(self.config.grid.ndim)], self.mu.data,
                self.rho.data)
            dt_nominal = min(dt_nominal, t_end - self.time)
            dt = self._step_with_retry(dt_nominal, verbose=verbose)
            if verbose and self.step % output_interval == 0:
                self._diagnostics.report(self, dt)
            if callback is not None and self.step % output_interval == 0:
                callback(self)
        if verbose:
            print(f"simulation finished t={self.time:.6f}, step={self.step}")

    def _snapshot(self):
        nd = self.config.grid.ndim
        return dict(psi=self.psi.data.clone(), rho=self.rho.data.clone(),
                    mu=self.mu.data.clone(), kappa=self.kappa.data.clone(),
                    phi=self.phi.data.clone(), pressure=self.pressure.data.clone(),
                    velocity=[self.velocity[a].clone() for a in range(nd)],
                    vel_star=[self.vel_star[a].clone() for a in range(nd)],
                    time=float(self.time), step=int(self.step))

    def _restore(self, s) -> None:
        for name in ("psi", "rho", "mu", "kappa", "phi", "pressure"):
            getattr(self, name)._data = s[name].clone()
        for a in range(self.config.grid.ndim):
            self.velocity._c[a] = s["velocity"][a].clone()
            self.vel_star._c[a] = s["vel_star"][a].clone()
        self.time, self.step = s["time"], s["step"]

    def _step_with_retry(self, dt_nominal: float, verbose: bool = False,
                         first_retry: int = 0, snap=None) -> float:
        """legacy_run_loop.py:45-67: rollback and halve dt, at most 12 times.
        ``first_retry`` > 0 continues a sequence whose earlier attempts ran
        elsewhere (a batched step): dt_nominal is then the nominal dt and
        ``snap`` the state before the first attempt."""
        max_retries = 12
        dt_try = float(dt_nominal) * 0.5 ** first_retry
        if snap is None:
            snap = self._snapshot()
        for retry in range(first_retry, max_retries + 1):
            try:
                self.step_forward(dt_try)
                if self._has_nonfinite_state():
                    raise FloatingPointError("non-finite state after step")
                return dt_try
            except Exception:
                self._restore(snap)
                if retry >= max_retries:
                    raise
                dt_try *= 0.5
                if verbose:
                    print(f"[run] step retry {retry + 1}/{max_retries}: "
                          f"reducing dt to {dt_try:.3e}")
        return dt_try

    def _has_nonfinite_state(self) -> bool:
        """legacy_run_loop.py:70-82."""
        nd = self.config.grid.ndim
        arrs = [self.psi.data, self.rho.data, self.mu.data, self.kappa.data, self.phi.data,
                self.pressure.data] + [self.velocity[a] for a in range(nd)] + \
            [self.vel_star[a] for a in range(nd)]
        cnt = C.c_int64()
        nat.check(nat.lib().tpf_ns_nonfinite(len(arrs), nat.parr([ptr(a) for a in arrs]),
                                             arrs[0].numel(), ptr(self._nf_work), C.byref(cnt),
                                             stream_handle()))
        return cnt.value > 0

    def step_forward(self, dt: float) -> None:
        """The 7-step algorithm (_core.py:126-166)."""
        vel_star, rhs = self._step_pre_ppe(dt)
        # 6: PPE for dp (legacy_flow_helpers.py:38-58)
        delta_p = self.ppe_solver.solve(rhs, self.rho.data, dt, p_init=None)
        self._step_post_ppe(delta_p, vel_star, dt)

    def _step_pre_ppe(self, dt: float):
        """Steps 1-5 and the PPE right-hand side; returns (u*, rhs)."""
        nd = self.config.grid.ndim
        # 1-2: CLS advection + reinitialisation (legacy_flow_helpers.py:23-27)
        vel = [self.velocity[a] for a in range(nd)]
        psi_adv = self.ls_advect.advance(self.psi.data, vel, dt)
        self.psi._data = self.ls_reinit.reinitialize(psi_adv)
        # 3-4
        self._update_properties()
        self._update_curvature()
        # 5: predictor with the extended p^n (IPC)
        n_hat = self._field_ext.compute_normal(self.phi.data)
        p_ext = self._field_ext.extend(self.pressure.data, self.phi.data, n_hat)
        state = FlowState(velocity=vel, psi=self.psi.data, rho=self.rho.data, mu=self.mu.data,
                          kappa=self.kappa.data, pressure=p_ext)
        vel_star = self.predictor.compute(state, dt)
        for a in range(nd):
            self.vel_star._c[a] = vel_star[a]
        # 5b: wall BC on u* before the PPE right-hand side reads it (_core.py:151)
        self._bc_handler.apply(self.vel_star)
        vel_star = [self.vel_star[a] for a in range(nd)]
        # 6a: Rhie-Chow divergence / dt (legacy_flow_helpers.py:38-52)
        div_rc = self.rhie_chow.face_velocity_divergence(vel_star, self.pressure.data,
                                                         self.rho.data, dt)
        rhs = torch.empty_like(div_rc)
        nat.check(nat.lib().tpf_ns_div_scalar(div_rc.numel(), ptr(div_rc), float(dt), ptr(rhs),
                                              stream_handle()))
        return vel_star, rhs

    def _step_post_ppe(self, delta_p: torch.Tensor, vel_star: List[torch.Tensor],
                       dt: float) -> None:
        """p += dp, then step 7: corrector with the extended dp
        (legacy_flow_helpers.py:55-67) and the wall BC."""
        nd = self.config.grid.ndim
        p_new = torch.empty_like(delta_p)
        nat.check(nat.lib().tpf_ns_add(delta_p.numel(), ptr(self.pressure.data), ptr(delta_p),
                                       ptr(p_new), stream_handle()))
        self.pressure._data = p_new
        n_hat = self._field_ext.compute_normal(self.phi.data)
        dp_ext = self._field_ext.extend(delta_p, self.phi.data, n_hat)
        vel_new = self.vel_corrector.correct(vel_star, dp_ext, self.rho.data, dt)
        for a in range(nd):
            self.velocity._c[a] = vel_new[a]
        self._bc_handler.apply(self.velocity)
        self.time += dt
        self.step += 1

    def _update_properties(self) -> None:
        """legacy_flow_helpers.py:70-80 (heaviside.py:167-187)."""
        psi = self.psi.data
        rho, mu = torch.empty_like(psi), torch.empty_like(psi)
        nat.check(nat.lib().tpf_ns_properties(psi.numel(), ptr(psi), self._rho_l, self._rho_g,
                                              self._mu_l, self._mu_g, 0.0, ptr(rho), ptr(mu),
                                              None, stream_handle()))
        self.rho._data, self.mu._data = rho, mu

    def _update_curvature(self) -> None:
        """legacy_flow_helpers.py:83-86."""
        self.kappa._data = self.curvature_calc.compute(self.psi.data)
        if self._needs_phi:
            psi = self.psi.data
            phi, tr, tm = (torch.empty_like(psi) for _ in range(3))
            nat.check(nat.lib().tpf_ns_properties(psi.numel(), ptr(psi), 1.0, 0.0, 1.0, 0.0,
                                                  self.eps, ptr(tr), ptr(tm), ptr(phi),
                                                  stream_handle()))
            self.phi._data = phi


class SimulationBuilder:
    """simulation/builder.py:46-117."""

    def __init__(self, config):
        self._config = config
        self._ppe_solver = None
        self._convection = None
        self._viscous = None
        self._gravity = None
        self._surface_tension = None

    def with_ppe_solver(self, solver) -> "SimulationBuilder":
        self._ppe_solver = solver
        return self

    def with_convection(self, term) -> "SimulationBuilder":
        self._convection = term
        return self

    def with_viscous(self, term) -> "SimulationBuilder":
        self._viscous = term
        return self

    def with_gravity(self, term) -> "SimulationBuilder":
        self._gravity = term
        return self

    def with_surface_tension(self, term) -> "SimulationBuilder":
        self._surface_tension = term
        return self

    def build(self) -> TwoPhaseSimulation:
        """legacy_builder_runtime.py:34-199 for the schemes this build has."""
        cfg = self._config
        num = cfg.numerics
        if num.advection_scheme != "dissipative_ccd":
            raise NotImplementedError(
                f"advection_scheme={num.advection_scheme!r} is not built; use 'dissipative_ccd'")
        if num.convection_scheme != "uccd6":
            raise NotImplementedError(
                f"convection_scheme={num.convection_scheme!r} is not built; use 'uccd6'")
        if num.reinit_method != "split":
            raise NotImplementedError(f"reinit_method={num.reinit_method!r} is not built")
        if num.surface_tension_model != "csf":
            raise NotImplementedError("surface_tension_model='gfm' is not built")
        if num.extension_method == "upwind" and num.n_extend > 0:
            raise NotImplementedError("extension_method='upwind' is not built")
        require_gpu()
        grid = Grid(cfg.grid)
        nd = cfg.grid.ndim
        bc = num.bc_type
        bc_spec = BoundarySpec(bc, grid.shape, grid.N)
        dx_min = min(cfg.grid.L[a] / cfg.grid.N[a] for a in range(nd))
        eps = num.epsilon_factor * dx_min
        ccd = CCDSolver(grid, None, bc_type=bc)
        ls_bc = tuple("periodic" if k == "periodic" else "neumann" for k in axis_kinds(bc, nd))
        ls_advect = DissipativeCCDAdvection(None, grid, ccd, bc=ls_bc, mass_correction=True)
        ls_reinit = SplitReinitializer(None, grid, ccd, eps, num.reinit_steps, bc=ls_bc)
        curvature_calc = CurvatureCalculatorPsi(None, ccd)
        predictor = Predictor(None, cfg, ccd, convection=self._convection, viscous=self._viscous,
                              gravity=self._gravity, surface_tension=self._surface_tension)
        ppe = self._ppe_solver or create_ppe_solver(cfg, None, grid, ccd=ccd, bc_spec=bc_spec)
        if num.extension_method == "hermite":
            ext = ClosestPointExtender(None, grid, ccd)
        else:
            ext = NullFieldExtender()
        comp = dict(config=cfg, backend=None, grid=grid, eps=eps, ccd=ccd, ls_advect=ls_advect,
                    ls_reinit=ls_reinit, curvature_calc=curvature_calc, predictor=predictor,
                    ppe_solver=ppe,
                    rhie_chow=RhieChowInterpolator(None, grid, ccd, bc_type=bc),
                    vel_corrector=VelocityCorrector(None, grid, ccd),
                    cfl_calc=CFLCalculator(None, grid, num.cfl_number, We=cfg.fluid.We,
                                           rho_ratio=cfg.fluid.rho_ratio,
                                           cn_viscous=num.cn_viscous),
                    _bc_handler=BoundaryConditionHandler(cfg),
                    _diagnostics=DiagnosticsReporter(None, grid), _field_ext=ext)
        return TwoPhaseSimulation._from_components(comp)


class BatchedSimulation:
    """Steps independent legacy simulations together (uneven grids allowed).

    Each step runs, per simulation, the CFL time step and steps 1-5 plus the
    Rhie-Chow right-hand side (``_step_pre_ppe``), then ONE batched solve of
    every simulation's PPE (``PPEBatch``: with Jacobi-PCG one launch per PCG
    step covers all problems and chunks of iterations replay as a captured
    graph; with ``ppe_preconditioner="mg"`` the multigrid-FCG solves advance
    together, each on its own stream), then the corrector
    (``_step_post_ppe``) and the non-finite check.
    A simulation whose attempt fails (PPE above tol, non-finite state, an
    exception) is rolled back and continues the run loop's dt-halving
    sequence on its own from the second attempt (legacy_run_loop.py:45-67),
    so every simulation follows exactly the trajectory of its own ``run``:
    the batched PPE solve equals the single solve bit for bit.

    The reference steps such simulations one after another; the simulations
    must use the CG form of the FVM matrix-free PPE, all with Jacobi
    preconditioning in FP64 or all with multigrid FCG, and share its
    tolerance settings."""

    def __init__(self, sims):
        from .ppe import PPEBatch
        self.sims = list(sims)
        if not self.sims:
            raise ValueError("BatchedSimulation needs at least one simulation")
        self.ppe_batch = PPEBatch([s.ppe_solver for s in self.sims])
        self.retries = 0          # simulations that needed the dt-halving path
        self.steps_taken = 0      # simulation steps completed

    def __len__(self):
        return len(self.sims)

    def initialize(self) -> None:
        """Properties and curvature of the initial state (as ``run`` does)."""
        for s in self.sims:
            s._update_properties()
            s._update_curvature()

    def step(self, t_end: Optional[float] = None) -> List[Optional[float]]:
        """One step of every simulation with its own CFL dt (clamped to
        ``t_end - time`` when given; simulations at t_end sit out).  Returns
        the dt each simulation took (None for one that sat out)."""
        sims = self.sims
        n = len(sims)
        dts: List[Optional[float]] = [None] * n
        snaps = [None] * n
        rhs: List[Optional[torch.Tensor]] = [None] * n
        sol: List[Optional[torch.Tensor]] = [None] * n
        vstar: List[Optional[list]] = [None] * n
        failed = [False] * n
        for i, s in enumerate(sims):
            if t_end is not None and s.time >= t_end:
                continue
            nd = s.config.grid.ndim
            dt = s.cfl_calc.compute([s.velocity[a] for a in range(nd)], s.mu.data, s.rho.data)
            if t_end is not None:
                dt = min(dt, t_end - s.time)
            dts[i] = float(dt)
            snaps[i] = s._snapshot()
            try:
                vstar[i], rhs[i] = s._step_pre_ppe(dts[i])
                s.ppe_solver.prepare_operator(s.rho.data)
                sol[i] = torch.empty_like(rhs[i])
            except Exception:
                failed[i] = True
                rhs[i] = sol[i] = None
        reps = self.ppe_batch.solve_into(rhs, sol, raise_on_failure=False)
        for i, s in enumerate(sims):
            if rhs[i] is None:
                continue
            if not reps[i]["converged"]:
                failed[i] = True
                continue
            try:
                s._step_post_ppe(sol[i], vstar[i], dts[i])
                if s._has_nonfinite_state():
                    failed[i] = True
            except Exception:
                failed[i] = True
        for i, s in enumerate(sims):
            if dts[i] is None:
                continue
            if failed[i]:
                s._restore(snaps[i])
                self.retries += 1
                dts[i] = s._step_with_retry(dts[i], first_retry=1, snap=snaps[i])
            self.steps_taken += 1
        return dts

    def run(self, t_end: float, max_steps: Optional[int] = None) -> None:
        """Steps until every simulation reaches t_end (or max_steps batched
        steps)."""
        k = 0
        while any(s.time < t_end for s in self.sims):
            if max_steps is not None and k >= max_steps:
                break
            self.step(t_end)
            k += 1


def bubble_simulation(case: dict, pseudo_tol: float = 1e-10, pseudo_maxiter: int = 20000,
                      ppe_preconditioner: str = "jacobi") -> TwoPhaseSimulation:
    """A legacy 2-D rising-bubble simulation of one ``uneven_bubble_batch``
    case: the bubble_2d parameters of the reference (configs/bubble_2d.yaml:
    Re 35, Fr 1, mu ratio 0.1, CFL 0.25, epsilon_factor 1.5, DCCD
    advection, PPE tolerance 1e-10) on the case's N x N grid of [0,1]^2,
    with the case's radius, density ratio and We; the PPE in CG form with
    `ppe_preconditioner` ("jacobi": the reference's Jacobi-PCG; "mg":
    flexible CG with one multigrid cycle per iteration) and iteration cap
    `pseudo_maxiter`, the null field extension."""
    from .config import FluidConfig, GridConfig, NumericsConfig, SimulationConfig, SolverConfig
    from .fields import bubble_psi_2d
    N = int(case["N"])
    cfg = SimulationConfig(
        grid=GridConfig(ndim=2, N=(N, N), L=(1.0, 1.0)),
        fluid=FluidConfig(Re=35.0, Fr=1.0, We=float(case["We"]),
                          rho_ratio=1.0 / float(case["density_ratio"]), mu_ratio=0.1),
        numerics=NumericsConfig(advection_scheme="dissipative_ccd", cfl_number=0.25,
                                extension_method="none"),
        solver=SolverConfig(ppe_solver_type="fvm_iterative", pseudo_tol=pseudo_tol,
                            pseudo_maxiter=pseudo_maxiter))
    cfg.solver.ppe_iteration_method = "cg"
    cfg.solver.ppe_preconditioner = ppe_preconditioner
    sim = SimulationBuilder(cfg).build()
    sim.psi.data = bubble_psi_2d(N, float(case["radius"]))
    return sim
