"""Oracle: Rhie-Chow face-velocity divergence, FP64, NumPy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``RhieChowInterpolator.face_velocity_divergence``
(``spatial/rhie_chow.py:55-284``, paths relative to ``src/twophase``): per
axis the CCD cell gradient of p (one-sided on wall nodes, :140-150), the
face fluxes u_bar - dt (2/(rho_L+rho_R)) [dp_face - dp_bar] on faces
1..N with face 0 the periodic wrap or 0 (:172-234), the optional
balanced-force bracket (:152-204) and the nodal flux divergence (:238-284).

Pinning (tests/test_oracle_rhie_chow.py): a linear velocity with p = 0 gives
the closed-form face averages; a cubic p gives the closed-form bracket
(-h^2/2 inside, the one-sided wall values at the ends); at the balanced-force
equilibrium p = kappa psi / We the bracket vanishes and the result equals the
p = 0 divergence.
"""

from __future__ import annotations

import numpy as np

from .ccd import OracleCCD, axis_kinds


def _sl(ndim, axis, idx):
    s = [slice(None)] * ndim
    s[axis] = idx
    return tuple(s)


class OracleRhieChow:
    def __init__(self, N, L, bc_type="wall"):
        self.N = tuple(int(n) for n in N)
        self.L = tuple(float(x) for x in L)
        self.ndim = len(self.N)
        self.kinds = axis_kinds(bc_type, self.ndim)
        self.ccd = OracleCCD(self.N, self.L, bc_type)

    def _flux(self, u, p, rho, axis, dt, h, kappa=None, psi=None, we=None):
        nd, n = self.ndim, self.N[axis]
        sl = lambda idx: _sl(nd, axis, idx)
        shape = list(u.shape)
        shape[axis] = n + 1
        flux = np.zeros(shape)
        dp = self.ccd.first_derivative(p, axis).copy()
        wall = self.kinds[axis] == "wall"
        if wall:
            dp[sl(0)] = (p[sl(1)] - p[sl(0)]) / h
            dp[sl(n)] = (p[sl(n)] - p[sl(n - 1)]) / h
        bf = kappa is not None and psi is not None and we is not None
        if bf:
            fs = kappa * self.ccd.first_derivative(psi, axis) / we
            if wall:
                fs[sl(0)] = kappa[sl(0)] * (psi[sl(1)] - psi[sl(0)]) / (h * we)
                fs[sl(n)] = kappa[sl(n)] * (psi[sl(n)] - psi[sl(n - 1)]) / (h * we)

        def face(L_, R_):
            u_bar = 0.5 * (u[L_] + u[R_])
            br = (p[R_] - p[L_]) / h - 0.5 * (dp[L_] + dp[R_])
            if bf:
                kf = 0.5 * (kappa[L_] + kappa[R_])
                br = br - (kf * (psi[R_] - psi[L_]) / (h * we) - 0.5 * (fs[L_] + fs[R_]))
            return u_bar - dt * (2.0 / (rho[L_] + rho[R_])) * br

        flux[sl(slice(1, n + 1))] = face(sl(slice(0, n)), sl(slice(1, n + 1)))
        if not wall:
            flux[sl(0)] = face(sl(n), sl(0))
        return flux

    def face_velocity_divergence(self, vel, p, rho, dt, kappa=None, psi=None, we=None):
        out = np.zeros(np.asarray(p).shape)
        for axis in range(self.ndim):
            h = self.L[axis] / self.N[axis]
            F = self._flux(np.asarray(vel[axis], float), np.asarray(p, float),
                           np.asarray(rho, float), axis, dt, h, kappa, psi, we)
            nd, n = self.ndim, self.N[axis]
            d = np.empty_like(out)
            d[_sl(nd, axis, slice(0, n))] = (F[_sl(nd, axis, slice(1, None))]
                                             - F[_sl(nd, axis, slice(0, n))]) / h
            if self.kinds[axis] == "periodic":
                d[_sl(nd, axis, n)] = (F[_sl(nd, axis, 0)] - F[_sl(nd, axis, n)]) / h
            else:
                d[_sl(nd, axis, n)] = -F[_sl(nd, axis, n)] / h
            out += d
        return out
