"""Writes the golden vectors of the legacy two-phase step from the reference.

Run in a container that holds the reference source (pure Python, NumPy and
SciPy); the GPU tests only read the .npz files this writes:

    python tests/golden/make_legacy_step_fixtures.py /path/to/reference/src

1. ``legacy_components.npz``: seeded smooth inputs on a 33 x 25 wall grid
   (L = 1.0 x 0.8) and the reference outputs of the step's building blocks:
   ``update_properties``, ``invert_heaviside``, ``UCCD6ConvectionTerm``,
   ``ViscousTerm`` (ccd_bulk body), ``GravityTerm``, ``SurfaceTensionTerm``,
   ``ccd_pressure_gradient``, ``ClosestPointExtender`` (normal and
   extension), ``Predictor.compute`` (an Euler call then an AB2 call, Picard
   CN viscous), ``CFLCalculator.compute`` and ``BoundaryConditionHandler``.
2. ``legacy_config1_trajectory.npz``: the 2-D 64 x 64 rising-bubble case
   (Re 100, Fr 1, We 10, density and viscosity ratio 0.1, CFL 0.3, DCCD
   advection, FVM matrix-free PPE by Jacobi-CG to 1e-10, bubble of radius
   0.25 at (0.5, 0.5), psi = 1/(1 + exp(-(r - 0.25)/(1.5 h)))) with
   ``extension_method="none"``: the reference run loop's first 20 steps
   (CFL dt, retry), per step dt, the divergence max |sum_a D1 u_a|,
   sum psi dV, ||p||_2 and max |u|, and the fields psi, u, v, p after steps
   1 and 20.
3. ``legacy_config1_hermite.npz``: the same case with the reference's
   default ``extension_method="hermite"`` (closest-point extension of p and
   dp), whose run diverges (max |div u| 74 after step 1, 8.6e5 after step
   2): the first two steps, recorded the same way.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def smooth(X, Y, seed, amp=1.0):
    rng = np.random.default_rng(seed)
    out = np.zeros_like(X)
    for _ in range(4):
        kx, ky = rng.integers(1, 4, size=2)
        a = rng.standard_normal()
        px, py = rng.uniform(0, 2 * np.pi, 2)
        out += a * np.sin(np.pi * kx * X + px) * np.cos(np.pi * ky * Y + py)
    return amp * out


def components(tp):
    from twophase.backend import Backend
    from twophase.ccd.ccd_solver import CCDSolver
    from twophase.config import (FluidConfig, GridConfig, NumericsConfig, SimulationConfig,
                                 SolverConfig)
    from twophase.core.flow_state import FlowState
    from twophase.core.grid import Grid
    from twophase.coupling.velocity_corrector import ccd_pressure_gradient
    from twophase.levelset.closest_point_extender import ClosestPointExtender
    from twophase.levelset.heaviside import invert_heaviside, update_properties
    from twophase.ns_terms.context import NSComputeContext
    from twophase.ns_terms.gravity import GravityTerm
    from twophase.ns_terms.surface_tension import SurfaceTensionTerm
    from twophase.ns_terms.uccd6_convection import UCCD6ConvectionTerm
    from twophase.ns_terms.viscous import ViscousTerm
    from twophase.simulation.boundary_condition import BoundaryConditionHandler
    from twophase.time_integration.ab2_predictor import Predictor
    from twophase.time_integration.cfl import CFLCalculator

    N, L = (32, 24), (1.0, 0.8)
    cfg = SimulationConfig(grid=GridConfig(ndim=2, N=N, L=L),
                           fluid=FluidConfig(Re=100.0, Fr=1.0, We=10.0, rho_ratio=0.1,
                                             mu_ratio=0.1),
                           numerics=NumericsConfig(advection_scheme="dissipative_ccd",
                                                   cfl_number=0.3),
                           solver=SolverConfig(pseudo_tol=1e-10))
    be = Backend(use_gpu=False)
    grid = Grid(cfg.grid, be)
    ccd = CCDSolver(grid, be, bc_type="wall")
    X, Y = np.meshgrid(np.linspace(0, L[0], N[0] + 1), np.linspace(0, L[1], N[1] + 1),
                       indexing="ij")
    h = min(L[0] / N[0], L[1] / N[1])
    eps = 1.5 * h
    r = np.sqrt((X - 0.45) ** 2 + (Y - 0.4) ** 2)
    psi = 1.0 / (1.0 + np.exp(-(r - 0.2) / eps))
    u, v = smooth(X, Y, 1), smooth(X, Y, 2)
    u2, v2 = u * 0.9 + smooth(X, Y, 3, 0.1), v * 0.9 + smooth(X, Y, 4, 0.1)
    p = smooth(X, Y, 5, 0.3)
    kappa = smooth(X, Y, 6, 5.0)
    rho, mu = update_properties(np, psi, 1.0, 0.1, 1.0, 0.1)
    phi = invert_heaviside(np, psi, eps)
    ctx = NSComputeContext(velocity=[u, v], ccd=ccd, rho=rho, mu=mu, kappa=kappa, psi=psi)
    conv = UCCD6ConvectionTerm(be, grid, ccd, sigma=1.0e-3).compute(ctx)
    visc = ViscousTerm(be, 100.0, True)._evaluate([u, v], mu, rho, ccd)
    grav = GravityTerm(be, 1.0, 2).compute(ctx)
    st = SurfaceTensionTerm(be, 10.0).compute(ctx)
    gradp = ccd_pressure_gradient(ccd, p, 2)
    ext = ClosestPointExtender(be, grid, ccd)
    n_hat = ext.compute_normal(phi)
    p_ext = ext.extend(p, phi, n_hat)
    pred = Predictor(be, cfg, ccd, convection=UCCD6ConvectionTerm(be, grid, ccd, sigma=1.0e-3))
    dt = 2.0e-3
    s1 = FlowState(velocity=[u, v], psi=psi, rho=rho, mu=mu, kappa=kappa, pressure=p)
    us1 = pred.compute(s1, dt)
    s2 = FlowState(velocity=[u2, v2], psi=psi, rho=rho, mu=mu, kappa=kappa, pressure=p)
    us2 = pred.compute(s2, dt)
    cfl = CFLCalculator(be, grid, 0.3, We=10.0, rho_ratio=0.1, cn_viscous=True)
    dt_cfl = cfl.compute([u, v], mu, rho)
    bc = [u.copy(), v.copy()]
    BoundaryConditionHandler(cfg).apply(bc)
    out = dict(N=np.array(N), L=np.array(L), eps=eps, psi=psi, u=u, v=v, u2=u2, v2=v2, p=p,
               kappa=kappa, rho=rho, mu=mu, phi=phi, conv_u=conv[0], conv_v=conv[1],
               visc_u=visc[0], visc_v=visc[1], grav_u=grav[0], grav_v=grav[1], st_u=st[0],
               st_v=st[1], gradp_x=gradp[0], gradp_y=gradp[1], nhat_x=n_hat[0],
               nhat_y=n_hat[1], p_ext=p_ext, dt=dt, ustar1_u=us1[0], ustar1_v=us1[1],
               ustar2_u=us2[0], ustar2_v=us2[1], dt_cfl=dt_cfl, bc_u=bc[0], bc_v=bc[1])
    np.savez_compressed(os.path.join(HERE, "legacy_components.npz"), **out)
    print("legacy_components.npz:", len(out), "arrays")


def trajectory(tp, steps=20, extension="none", name="legacy_config1_trajectory.npz"):
    from twophase.config import (FluidConfig, GridConfig, NumericsConfig, SimulationConfig,
                                 SolverConfig)
    from twophase.simulation.builder import SimulationBuilder

    n = 64
    cfg = SimulationConfig(grid=GridConfig(ndim=2, N=(n, n), L=(1.0, 1.0)),
                           fluid=FluidConfig(Re=100.0, Fr=1.0, We=10.0, rho_ratio=0.1,
                                             mu_ratio=0.1),
                           numerics=NumericsConfig(advection_scheme="dissipative_ccd",
                                                   cfl_number=0.3, extension_method=extension),
                           solver=SolverConfig(ppe_solver_type="fvm_iterative",
                                               pseudo_tol=1e-10))
    cfg.solver.ppe_iteration_method = "cg"
    sim = SimulationBuilder(cfg).build()
    X, Y = sim.grid.meshgrid()
    sim.psi.data[:] = 1.0 / (1.0 + np.exp(-(np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2) - 0.25)
                                          / (1.5 / n)))
    psi0 = np.array(sim.psi.data)
    sim._update_properties()
    sim._update_curvature()
    h = 1.0 / n
    rec = {k: [] for k in ("dt", "div_max", "mass", "p_l2", "umax")}
    fields = {}
    for k in range(steps):
        dt = sim.cfl_calc.compute([sim.velocity[a] for a in range(2)], sim.mu.data,
                                  sim.rho.data)
        dt = sim._step_with_retry(dt)
        div = sum(sim.ccd.differentiate(sim.velocity[a], a)[0] for a in range(2))
        rec["dt"].append(dt)
        rec["div_max"].append(float(np.max(np.abs(div))))
        rec["mass"].append(float(np.sum(sim.psi.data)) * h * h)
        rec["p_l2"].append(float(np.linalg.norm(sim.pressure.data)))
        rec["umax"].append(float(max(np.max(np.abs(sim.velocity[a])) for a in range(2))))
        if k + 1 in (1, steps):
            for fname, arr in (("psi", sim.psi.data), ("u", sim.velocity[0]),
                              ("v", sim.velocity[1]), ("p", sim.pressure.data)):
                fields[f"{fname}_step{k + 1}"] = np.array(arr)
        print(f"step {k + 1}: dt={dt:.4e} div={rec['div_max'][-1]:.3e} "
              f"mass={rec['mass'][-1]:.12f} |p|={rec['p_l2'][-1]:.6e}", flush=True)
    np.savez_compressed(os.path.join(HERE, name), psi0=psi0,
                        **{k: np.array(v) for k, v in rec.items()}, **fields)


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else None
    if not src or not os.path.isdir(os.path.join(src, "twophase")):
        sys.exit("usage: make_legacy_step_fixtures.py <reference src directory>")
    sys.path.insert(0, src)
    import twophase as tp
    components(tp)
    trajectory(tp)
    trajectory(tp, steps=2, extension="hermite", name="legacy_config1_hermite.npz")
