"""Oracle: dissipative-CCD level-set advection (TVD-RK3), FP64, NumPy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``DissipativeCCDAdvection.advance`` / ``_rhs``
(``levelset/advection_dccd.py:84-152``), the adaptive filter
(``levelset/advection_kernels.py:31-34``), ghost padding
(``core/boundary.py:249-286``), TVD-RK3 (``time_integration/tvd_rk3.py:
26-48``), the image synchronisation (``core/boundary.py:116-133``) and the
mass correction (``levelset/heaviside.py:118-149``), on the CCD oracle.

Pinning (tests/test_oracle_dccd.py): with eps_d = 0 and constant velocity a
single Fourier mode is multiplied by the TVD-RK3 polynomial 1 + z + z^2/2 +
z^3/6 of z = -i u S1(theta) dt (S1 the CCD f' symbol); with a constant psi
the filter multiplies the derivative by 1 - 4 eps_d (2 psi - 1)^2
sin^2(theta/2); non-periodic end nodes are unfiltered; the mass correction
restores sum psi dV.
"""

from __future__ import annotations

import numpy as np

from .ccd import OracleCCD


class OracleDCCD:
    def __init__(self, N, L, ccd_bc="periodic", ls_bc="periodic", eps_d=0.05,
                 mass_correction=False):
        self.N = tuple(int(n) for n in N)
        self.L = tuple(float(x) for x in L)
        self.ndim = len(self.N)
        self.ccd = OracleCCD(self.N, self.L, ccd_bc)
        self.bc = tuple(ls_bc) if isinstance(ls_bc, (tuple, list)) else (ls_bc,) * self.ndim
        self.eps_d = float(eps_d)
        self.mass_correction = mass_correction

    def sync(self, arr):
        for a, b in enumerate(self.bc):
            if b == "periodic":
                src = [slice(None)] * arr.ndim
                img = [slice(None)] * arr.ndim
                src[a], img[a] = 0, -1
                arr[tuple(img)] = arr[tuple(src)]
        return arr

    def _pad(self, arr, axis):
        n = arr.shape[axis]
        sl = lambda a, b: tuple(slice(a, b) if k == axis else slice(None) for k in range(arr.ndim))
        if self.bc[axis] == "periodic":
            return np.concatenate([arr[sl(n - 2, n - 1)], arr, arr[sl(1, 2)]], axis=axis)
        return np.concatenate([np.flip(arr[sl(0, 1)], axis=axis), arr,
                               np.flip(arr[sl(n - 1, n)], axis=axis)], axis=axis)

    def rhs(self, psi, vel):
        out = np.zeros_like(psi)
        for ax in range(self.ndim):
            fp = self.ccd.first_derivative(psi * vel[ax], ax)
            n = fp.shape[ax]
            pad = self._pad(fp, ax)
            sl = lambda a, b: tuple(slice(a, b) if k == ax else slice(None) for k in range(fp.ndim))
            sw = (2.0 * psi - 1.0) * (2.0 * psi - 1.0)
            Ft = fp + self.eps_d * sw * (pad[sl(2, n + 2)] - 2.0 * fp + pad[sl(0, n)])
            if self.bc[ax] != "periodic":
                lo = tuple(0 if k == ax else slice(None) for k in range(fp.ndim))
                hi = tuple(n - 1 if k == ax else slice(None) for k in range(fp.ndim))
                Ft[lo], Ft[hi] = fp[lo], fp[hi]
            out -= Ft
        return out

    def dV(self):
        w = []
        for a in range(self.ndim):
            h = self.L[a] / self.N[a]
            v = np.full(self.N[a] + 1, h)
            if self.bc[a] == "periodic":
                v[-1] = 0.0
            w.append(v)
        vol = w[0]
        for a in range(1, self.ndim):
            vol = np.expand_dims(vol, axis=a) * w[a].reshape((1,) * a + (-1,))
        return vol

    def advance(self, psi, vel, dt, clip_bounds=(0.0, 1.0)):
        psi = self.sync(np.array(psi, dtype=np.float64))
        dV = self.dV() if self.mass_correction else None
        m_old = float(np.sum(psi * dV)) if self.mass_correction else None

        def P(q):
            if clip_bounds is not None:
                q = np.clip(q, clip_bounds[0], clip_bounds[1])
            return self.sync(q)

        q0 = psi
        q1 = P(q0 + dt * self.rhs(q0, vel))
        q2 = P(0.75 * q0 + 0.25 * (q1 + dt * self.rhs(q1, vel)))
        qn = P((1.0 / 3.0) * q0 + (2.0 / 3.0) * (q2 + dt * self.rhs(q2, vel)))
        if self.mass_correction:
            M_new = np.sum(qn * dV)
            w = 4.0 * qn * (1.0 - qn)
            W = np.sum(w * dV)
            ratio = (m_old - M_new) / (W if W > 1e-12 else 1.0)
            qn = self.sync(np.clip(qn + ((1.0 if W > 1e-12 else 0.0) * ratio) * w, 0.0, 1.0))
        return qn
