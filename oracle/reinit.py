"""Oracle: operator-split CLS reinitialisation, FP64, NumPy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``SplitReinitializer.reinitialize`` (``levelset/reinit_split.py:
72-102``) with ``compute_dtau``, ``compute_gradient_normal``,
``filtered_divergence``, ``dccd_compression_div``, ``build_cn_factors`` and
``cn_diffusion_axis`` on their CPU (Thomas) path (``levelset/reinit_ops.py:
36-204``), ghost padding (``core/boundary.py:249-293``), image
synchronisation and the mass correction (``levelset/heaviside.py:118-149``),
on the CCD oracle.

Pinning (tests/test_oracle_reinit.py): one CN-ADI sweep multiplies the
mirror-symmetric cosine modes by the closed-form factor (1 + 2 b c - 6 mu (1 -
c)) / (1 + 2 b c + 6 mu (1 - c)) (b = -1/8, c = cos theta; M2 and B2 share
those eigenvectors); dtau follows its formula; the mass correction restores
sum psi dV; a tanh profile of width eps is close to a fixed point.  The
compression step: the filter multiplies a cosine mode's CCD derivative by
1 - 4 eps_d sin^2(theta/2); on linear fluxes the filtered derivative is the
slope inside and slope * (1 - eps_d) (zero ghost) or the slope (neumann /
outflow ghost) at the walls; on a monotone tanh profile (n = +1) the
compression divergence is d/dx[psi (1 - psi)] = (1 - 2 psi) psi'.
"""

from __future__ import annotations

import numpy as np

from .ccd import OracleCCD

BETA2 = -1.0 / 8.0


def compute_dtau(N, L, eps):
    dx = min(L[a] / N[a] for a in range(len(N)))
    return min(0.5 * dx ** 2 / (2.0 * len(N) * eps), 0.5 * dx)


def cn_factors(n, mu):
    dL, cL = 1.0 + 6.0 * mu, BETA2 - 3.0 * mu
    main = np.full(n, dL)
    sup = np.full(n - 1, cL)
    sub = np.full(n - 1, cL)
    sup[0] = 2.0 * cL
    sub[-1] = 2.0 * cL
    m = main.copy()
    f = np.empty(n - 1)
    for i in range(1, n):
        f[i - 1] = sub[i - 1] / m[i - 1]
        m[i] -= f[i - 1] * sup[i - 1]
    return f, m, sup


def cn_sweep(psi, axis, mu, factors):
    dR, cR = 1.0 - 6.0 * mu, BETA2 + 3.0 * mu
    p = np.moveaxis(psi, axis, 0)
    n = p.shape[0]
    rhs = np.empty_like(p)
    rhs[1:-1] = cR * p[:-2] + dR * p[1:-1] + cR * p[2:]
    rhs[0] = dR * p[0] + 2.0 * cR * p[1]
    rhs[-1] = 2.0 * cR * p[-2] + dR * p[-1]
    f, m, sup = factors
    d = rhs.copy()
    for i in range(1, n):
        d[i] = d[i] - f[i - 1] * d[i - 1]
    x = np.empty_like(d)
    x[-1] = d[-1] / m[-1]
    for i in range(n - 2, -1, -1):
        x[i] = (d[i] - sup[i] * x[i + 1]) / m[i]
    return np.moveaxis(x, 0, axis)


class OracleSplitReinit:
    def __init__(self, N, L, ccd_bc, eps, n_steps=4, bc="zero", eps_d_comp=0.05,
                 mass_correction=True):
        self.N = tuple(int(n) for n in N)
        self.L = tuple(float(x) for x in L)
        self.ndim = len(self.N)
        self.ccd = OracleCCD(self.N, self.L, ccd_bc)
        self.eps, self.n_steps, self.eps_d = float(eps), int(n_steps), float(eps_d_comp)
        self.bc = tuple(bc) if isinstance(bc, (tuple, list)) else (bc,) * self.ndim
        self.mass_correction = mass_correction
        self.dtau = compute_dtau(self.N, self.L, self.eps)
        self.h = [self.L[a] / self.N[a] for a in range(self.ndim)]
        self.factors = [cn_factors(self.N[a] + 1, self.mu(a)) for a in range(self.ndim)]

    def mu(self, a):
        return self.eps * self.dtau / (2.0 * self.h[a] ** 2)

    def sync(self, q):
        for a, b in enumerate(self.bc):
            if b == "periodic":
                s = [slice(None)] * q.ndim
                i = [slice(None)] * q.ndim
                s[a], i[a] = 0, -1
                q[tuple(i)] = q[tuple(s)]
        return q

    def _pad(self, g, ax):
        n = g.shape[ax]
        sl = lambda a, b: tuple(slice(a, b) if k == ax else slice(None) for k in range(g.ndim))
        b = self.bc[ax]
        if b == "periodic":
            return np.concatenate([g[sl(n - 2, n - 1)], g, g[sl(1, 2)]], axis=ax)
        if b in ("neumann", "outflow"):
            return np.concatenate([g[sl(0, 1)], g, g[sl(n - 1, n)]], axis=ax)
        z = np.zeros_like(g[sl(0, 1)])
        return np.concatenate([z, g, z], axis=ax)

    def filtered_divergence(self, flux, a):
        """CCD f' of one flux component plus the 3-point dissipative filter
        on one ghost per side (reinit_ops.py:85-108, uniform grids)."""
        gp = self.ccd.first_derivative(flux, a)
        pad = self._pad(gp, a)
        n = gp.shape[a]
        sl = lambda s, e: tuple(slice(s, e) if k == a else slice(None) for k in range(flux.ndim))
        c = pad[sl(1, n + 1)]
        return c + self.eps_d * (pad[sl(2, n + 2)] - 2.0 * c + pad[sl(0, n)])

    def compression_div(self, q):
        """sum_a filtered D1[q (1 - q) n_a], n = grad q / max(|grad q|, floors)
        (reinit_ops.py:53-83, 111-119)."""
        d = [self.ccd.first_derivative(q, a) for a in range(self.ndim)]
        gs = sum(g * g for g in d)
        safe = np.maximum(np.sqrt(np.maximum(gs, 1e-12)), 1e-6)
        w = q * (1.0 - q)
        div = np.zeros_like(q)
        for a in range(self.ndim):
            div = div + self.filtered_divergence(w * (d[a] / safe), a)
        return div

    def dV(self):
        w = []
        for a in range(self.ndim):
            v = np.full(self.N[a] + 1, self.h[a])
            if self.bc[a] == "periodic":
                v[-1] = 0.0
            w.append(v)
        vol = w[0]
        for a in range(1, self.ndim):
            vol = np.expand_dims(vol, axis=a) * w[a].reshape((1,) * a + (-1,))
        return vol

    def reinitialize(self, psi):
        q = self.sync(np.array(psi, dtype=np.float64))
        dV = self.dV()
        m_old = np.sum(q * dV)
        for _ in range(self.n_steps):
            qs = self.sync(np.clip(q - self.dtau * self.compression_div(q), 0.0, 1.0))
            qn = qs
            for a in range(self.ndim):
                qn = self.sync(cn_sweep(qn, a, self.mu(a), self.factors[a]))
            q = self.sync(np.clip(qn, 0.0, 1.0))
        if self.mass_correction:
            M_new = np.sum(q * dV)
            w = 4.0 * q * (1.0 - q)
            W = np.sum(w * dV)
            ratio = (m_old - M_new) / (W if W > 1e-12 else 1.0)
            q = self.sync(np.clip(q + ((1.0 if W > 1e-12 else 0.0) * ratio) * w, 0.0, 1.0))
        return q
