"""Oracle: the legacy two-phase time step (config-1 loop), FP64, NumPy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates, on the CCD oracle:

* ``update_properties`` and ``invert_heaviside`` (``levelset/heaviside.py:
  90-118, 167-187``);
* ``UCCD6ConvectionTerm.compute`` (``ns_terms/uccd6_convection.py:88-128``):
  skew form -(u_k D1 u_j + D1(u_k u_j))/2 minus sigma max|u_k| h_k^7 D2^4 u_j;
* the viscous body of ``ViscousTerm`` with the default ``ccd_bulk`` scheme
  (``ns_terms/viscous_spatial.py:39-73, 93-130, 156-169``): CCD velocity
  gradients, low-order (one-sided at the ends, three-point inside)
  divergence of mu (grad u + grad u^T), over Re rho;
* ``GravityTerm`` / ``SurfaceTensionTerm`` (``ns_terms/gravity.py:57-86``,
  ``ns_terms/surface_tension.py:55-96``) and ``ccd_pressure_gradient``
  (``coupling/velocity_corrector.py:39-50``);
* ``ClosestPointExtender`` in 2-D (``levelset/closest_point_extender.py:
  101-228``): CCD normal of phi, closest point, bracket search with the
  three one-sided shifts, tensor-product quintic Hermite;
* ``Predictor.compute`` with AB2 start-up and the Picard CN viscous advance
  (``time_integration/ab2_predictor.py:110-204``,
  ``time_integration/cn_advance/picard_cn.py:44-96``);
* ``CFLCalculator.compute`` (``time_integration/cfl.py:85-137``) and
  ``BoundaryConditionHandler.apply`` (``simulation/boundary_condition.py:
  36-87``);
* ``TwoPhaseSimulation.step_forward`` and the run loop's first steps
  (``simulation/_core.py:126-166``, ``legacy_flow_helpers.py:23-86``,
  ``legacy_run_loop.py:13-67``) composed from the DCCD, reinitialisation,
  curvature, Rhie-Chow, PPE and corrector oracles.

Pinning (tests/test_oracle_legacy_step.py): the reference's own outputs on
seeded inputs (``tests/golden/legacy_components.npz``) for every building
block, and its 20-step trajectory (``legacy_config1_trajectory.npz``,
``legacy_config1_hermite.npz``); closed forms: the extension of a quadratic
field across a straight interface is exact, the viscous body of a linear
velocity and viscosity is exact, UCCD6 of a constant field is zero.
"""

from __future__ import annotations

import math

import numpy as np

from .ccd import OracleCCD
from .curvature import OracleCurvaturePsi
from .dccd import OracleDCCD
from .fields import correct_velocity
from .ppe import OraclePPE
from .reinit import OracleSplitReinit
from .rhie_chow import OracleRhieChow


def update_properties(psi, rho_l, rho_g, mu_l, mu_g):
    """heaviside.py:167-187."""
    return rho_g + (rho_l - rho_g) * psi, mu_g + (mu_l - mu_g) * psi


def invert_heaviside(psi, eps):
    """heaviside.py:90-118: eps ln(psi / (1 - psi)) after clipping psi into
    [1e-6, 1 - 1e-6]."""
    p = np.clip(psi, 1e-6, 1.0 - 1e-6)
    return eps * np.log(p / (1.0 - p))


def low_order_derivative(arr, axis, x):
    """viscous_spatial.py:93-130: first-order one-sided at the two ends,
    the three-point (non-uniform) central formula inside."""
    a = np.moveaxis(np.asarray(arr, dtype=np.float64), axis, 0)
    dx = np.diff(x)
    out = np.empty_like(a)
    out[0] = (a[1] - a[0]) / dx[0]
    out[-1] = (a[-1] - a[-2]) / dx[-1]
    shp = (-1,) + (1,) * (a.ndim - 1)
    hl, hr = dx[:-1].reshape(shp), dx[1:].reshape(shp)
    out[1:-1] = (-(hr / (hl * (hl + hr))) * a[:-2] + ((hr - hl) / (hl * hr)) * a[1:-1]
                 + (hl / (hr * (hl + hr))) * a[2:])
    return np.moveaxis(out, 0, axis)


def uccd6_convection(ccd: OracleCCD, vel, sigma=1.0e-3):
    """uccd6_convection.py:88-128."""
    nd = len(vel)
    h7 = [(ccd.L[k] / ccd.N[k]) ** 7 for k in range(nd)]
    out = []
    for j in range(nd):
        acc = np.zeros_like(vel[j])
        for k in range(nd):
            d1, d2 = ccd.differentiate(vel[j], k)
            cons = ccd.first_derivative(vel[k] * vel[j], k)
            acc = acc - 0.5 * (vel[k] * d1 + cons)
            d4 = ccd.second_derivative(d2, k)
            d6 = ccd.second_derivative(d4, k)
            d8 = ccd.second_derivative(d6, k)
            acc = acc - (sigma * h7[k]) * (np.max(np.abs(vel[k])) * d8)
        out.append(acc)
    return out


def viscous_ccd_bulk(ccd: OracleCCD, vel, mu, rho, Re):
    """viscous_spatial.py:156-169 with the 1/(Re rho) of :39-73."""
    nd = len(vel)
    xs = [np.linspace(0.0, ccd.L[a], ccd.N[a] + 1) for a in range(nd)]
    out = []
    for al in range(nd):
        tot = np.zeros_like(vel[al])
        for be in range(nd):
            s = mu * (ccd.first_derivative(vel[al], be) + ccd.first_derivative(vel[be], al))
            tot = tot + low_order_derivative(s, be, xs[be])
        out.append(tot / (Re * rho))
    return out


def gravity(rho, Fr, nd):
    g = [np.zeros_like(rho) for _ in range(nd)]
    g[nd - 1] = -rho / Fr ** 2
    return g


def surface_tension(ccd: OracleCCD, kappa, psi, We):
    return [kappa * ccd.first_derivative(psi, a) / We for a in range(psi.ndim)]


def pressure_gradient(ccd: OracleCCD, p):
    """velocity_corrector.py:39-50: CCD D1 with wall-normal planes zeroed."""
    out = []
    for a in range(p.ndim):
        g = ccd.first_derivative(p, a).copy()
        if ccd.kinds[a] == "wall":
            lo = tuple(0 if k == a else slice(None) for k in range(p.ndim))
            hi = tuple(-1 if k == a else slice(None) for k in range(p.ndim))
            g[lo] = 0.0
            g[hi] = 0.0
        out.append(g)
    return out


def wall_bc(vel, kinds):
    """boundary_condition.py:36-76: every component zero on both planes of
    every wall axis; the image plane copies the root on periodic axes."""
    for a, k in enumerate(kinds):
        lo = tuple(0 if j == a else slice(None) for j in range(vel[0].ndim))
        hi = tuple(-1 if j == a else slice(None) for j in range(vel[0].ndim))
        for u in vel:
            if k == "wall":
                u[lo] = 0.0
                u[hi] = 0.0
            else:
                u[hi] = u[lo]
    return vel


def quintic_hermite(t, Fa, Ga, Ha, Fb, Gb, Hb):
    """closest_point_extender.py:47-65 (G = h f', H = h^2 f'')."""
    c3 = 10 * (Fb - Fa) - 6 * Ga - 4 * Gb + 0.5 * (Hb - 3 * Ha)
    c4 = 15 * (Fa - Fb) + 8 * Ga + 7 * Gb + 0.5 * (3 * Ha - 2 * Hb)
    c5 = 6 * (Fb - Fa) - 3 * (Ga + Gb) + 0.5 * (Hb - Ha)
    return Fa + t * (Ga + t * (0.5 * Ha + t * (c3 + t * (c4 + t * c5))))


class ClosestPointExtender2D:
    """closest_point_extender.py:101-228 (ndim = 2 only, as the reference)."""

    def __init__(self, ccd: OracleCCD):
        if ccd.ndim != 2:
            raise NotImplementedError("ClosestPointExtender: only ndim=2 implemented")
        self.ccd = ccd
        self.x = [np.linspace(0.0, ccd.L[a], ccd.N[a] + 1) for a in range(2)]

    def compute_normal(self, phi):
        d = [self.ccd.first_derivative(phi, a) for a in range(2)]
        gs = d[0] * d[0] + d[1] * d[1]
        nrm = np.maximum(np.sqrt(np.maximum(gs, 1e-28)), 1e-14)
        return [g / nrm for g in d]

    def extend(self, q, phi, n_hat=None):
        c = self.ccd
        if n_hat is None:
            n_hat = self.compute_normal(phi)
        xs, ys = self.x
        nx, ny = len(xs), len(ys)
        qx, qxx = c.differentiate(q, 0)
        qy, qyy = c.differentiate(q, 1)
        qxy, qxxy = c.differentiate(qy, 0)
        qxyy, qxxyy = c.differentiate(qyy, 0)
        XX, YY = np.meshgrid(xs, ys, indexing="ij")
        xc = np.clip(XX - phi * n_hat[0], xs[0], xs[-1])
        yc = np.clip(YY - phi * n_hat[1], ys[0], ys[-1])
        ix = np.clip(np.searchsorted(xs, xc.ravel(), side="right").reshape(nx, ny) - 1, 0, nx - 2)
        iy = np.clip(np.searchsorted(ys, yc.ravel(), side="right").reshape(nx, ny) - 1, 0, ny - 2)
        for _ in range(3):
            ix1 = np.clip(ix + 1, 0, nx - 1)
            ix = np.where(phi[ix1, iy] >= 0, np.clip(ix - 1, 0, nx - 2), ix)
        ix1 = np.clip(ix + 1, 0, nx - 1)
        for _ in range(3):
            iy1 = np.clip(iy + 1, 0, ny - 1)
            iy = np.where(phi[ix, iy1] >= 0, np.clip(iy - 1, 0, ny - 2), iy)
        iy1 = np.clip(iy + 1, 0, ny - 1)
        hx = xs[ix1] - xs[ix]
        hy = ys[iy1] - ys[iy]
        tx = (xc - xs[ix]) / np.maximum(hx, 1e-30)
        ty = (yc - ys[iy]) / np.maximum(hy, 1e-30)

        def row(F, G, H, jy):
            return quintic_hermite(tx, F[ix, jy], hx * G[ix, jy], hx * hx * H[ix, jy],
                                   F[ix1, jy], hx * G[ix1, jy], hx * hx * H[ix1, jy])

        v0, d0, dd0 = row(q, qx, qxx, iy), row(qy, qxy, qxxy, iy), row(qyy, qxyy, qxxyy, iy)
        v1, d1, dd1 = row(q, qx, qxx, iy1), row(qy, qxy, qxxy, iy1), row(qyy, qxyy, qxxyy, iy1)
        qi = quintic_hermite(ty, v0, hy * d0, hy * hy * dd0, v1, hy * d1, hy * hy * dd1)
        return np.where(phi >= 0, qi, q)


class OraclePredictor:
    """ab2_predictor.py:110-204 with the Picard CN advance (picard_cn.py:44-96)."""

    def __init__(self, ccd: OracleCCD, Re, Fr, We, sigma=1.0e-3, cn_viscous=True):
        self.ccd, self.Re, self.Fr, self.We = ccd, Re, Fr, We
        self.sigma, self.cn = sigma, cn_viscous
        self.conv_prev = None

    def compute(self, vel, psi, rho, mu, kappa, p, dt):
        nd = len(vel)
        conv = uccd6_convection(self.ccd, vel, self.sigma)
        if self.conv_prev is not None:
            ab2 = [1.5 * conv[c] - 0.5 * self.conv_prev[c] for c in range(nd)]
        else:
            ab2 = conv
        grav = gravity(rho, self.Fr, nd)
        st = surface_tension(self.ccd, kappa, psi, self.We)
        gp = pressure_gradient(self.ccd, p)
        rhs = [rho * ab2[c] + grav[c] + st[c] + (-gp[c]) for c in range(nd)]
        vn = viscous_ccd_bulk(self.ccd, vel, mu, rho, self.Re)
        if self.cn:
            up = [vel[c] + dt * (rhs[c] / rho + vn[c]) for c in range(nd)]
            vs = viscous_ccd_bulk(self.ccd, up, mu, rho, self.Re)
            out = [vel[c] + dt * (rhs[c] / rho + 0.5 * vn[c] + 0.5 * vs[c]) for c in range(nd)]
        else:
            out = [vel[c] + dt * (rhs[c] + rho * vn[c]) / rho for c in range(nd)]
        self.conv_prev = [c.copy() for c in conv]
        return out


def cfl_dt(vel, mu, rho, N, L, cfl, We=None, rho_ratio=0.0, cn_viscous=True):
    """cfl.py:72-137."""
    h = min(L[a] / N[a] for a in range(len(N)))
    usum = sum(float(np.max(np.abs(u))) for u in vel)
    cands = [cfl * h / max(usum, 1e-14)]
    if not cn_viscous:
        cands.append(cfl * h * h / (4.0 * max(float(np.max(mu / rho)), 1e-14)))
    if We is not None and We > 0.0:
        cands.append(cfl * math.sqrt((1.0 + rho_ratio) * We * h ** 3 / (2.0 * math.pi)))
    return max(min(cands), 1e-8)


class OracleLegacySimulation:
    """TwoPhaseSimulation (simulation/_core.py:126-166) on uniform wall
    grids with the legacy builder's defaults (legacy_builder_runtime.py:
    34-199): DCCD advection with mass correction, split reinitialisation,
    psi curvature, UCCD6 convection, Picard CN viscous, CSF, Rhie-Chow
    right-hand side, Jacobi-CG FVM PPE, CCD corrector; extension "none" or
    "hermite"."""

    def __init__(self, N, L, Re=100.0, Fr=1.0, We=10.0, rho_ratio=0.1, mu_ratio=0.1,
                 cfl=0.3, eps_factor=1.5, reinit_steps=4, tol=1e-10, maxiter=500,
                 extension="none", sigma=1.0e-3):
        self.N, self.L = tuple(N), tuple(float(x) for x in L)
        nd = len(N)
        self.ccd = OracleCCD(self.N, self.L, "wall")
        self.kinds = self.ccd.kinds
        self.eps = eps_factor * min(self.L[a] / self.N[a] for a in range(nd))
        ls = ("neumann",) * nd
        self.adv = OracleDCCD(self.N, self.L, "wall", ls, 0.05, mass_correction=True)
        self.reinit = OracleSplitReinit(self.N, self.L, "wall", self.eps, reinit_steps, ls)
        self.curv = OracleCurvaturePsi(self.N, self.L, "wall")
        self.pred = OraclePredictor(self.ccd, Re, Fr, We, sigma)
        self.rc = OracleRhieChow(self.N, self.L, "wall")
        self.ppe = OraclePPE(self.N, self.L)
        self.ext = ClosestPointExtender2D(self.ccd) if extension == "hermite" else None
        self.props = (1.0, rho_ratio, 1.0, mu_ratio)
        self.cfl, self.We, self.rho_ratio = cfl, We, rho_ratio
        self.tol, self.maxiter = tol, maxiter
        shape = tuple(n + 1 for n in self.N)
        self.psi = np.zeros(shape)
        self.p = np.zeros(shape)
        self.vel = [np.zeros(shape) for _ in range(nd)]
        self.time, self.step = 0.0, 0

    def update_properties(self):
        self.rho, self.mu = update_properties(self.psi, *self.props)

    def update_curvature(self):
        self.kappa = self.curv.compute(self.psi)
        if self.ext is not None:
            self.phi = invert_heaviside(self.psi, self.eps)

    def _extend(self, q):
        if self.ext is None:
            return q
        return self.ext.extend(q, self.phi, self.ext.compute_normal(self.phi))

    def cfl_dt(self):
        return cfl_dt(self.vel, self.mu, self.rho, self.N, self.L, self.cfl, self.We,
                      self.rho_ratio, True)

    def step_forward(self, dt):
        self.psi = self.reinit.reinitialize(self.adv.advance(self.psi, self.vel, dt))
        self.update_properties()
        self.update_curvature()
        us = self.pred.compute(self.vel, self.psi, self.rho, self.mu, self.kappa,
                               self._extend(self.p), dt)
        us = wall_bc([u.copy() for u in us], self.kinds)
        rhs = self.rc.face_velocity_divergence(us, self.p, self.rho, dt) / dt
        dp, _ = self.ppe.solve_cg(rhs, self.rho, tol=self.tol, maxiter=self.maxiter)
        self.p = self.p + dp
        self.vel = wall_bc(correct_velocity(self.ccd, us, self._extend(dp), self.rho, dt),
                           self.kinds)
        self.time += dt
        self.step += 1

    def divergence_max(self):
        return float(np.max(np.abs(sum(self.ccd.first_derivative(self.vel[a], a)
                                       for a in range(len(self.N))))))
