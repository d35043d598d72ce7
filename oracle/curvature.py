"""Oracle: interface curvature from psi, FP64, NumPy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``CurvatureCalculatorPsi.compute`` (``levelset/curvature_psi.py:
79-168``), the DCCD filter (``levelset/curvature.py:60-128``) and the 3x3
Gaussian (``levelset/curvature_psi.py:171-196``) on the CCD oracle.

Pinning (tests/test_oracle_curvature.py): quadratic psi whose level sets are
circles / spheres gives kappa = -1/r (2-D) and -2/r (3-D) exactly (the CCD is
exact for quadratics); a rotated ellipse (non-zero psi_xy) equals the
axis-aligned ellipse's closed form at the rotated point; the filter leaves
linear fields unchanged inside and restores wall planes.
"""

from __future__ import annotations

import numpy as np

from .ccd import OracleCCD, axis_kinds


def _filter_1d(f, ax, N, coeff, periodic):
    nd = f.ndim
    res = f.copy()
    sl = lambda a, b: tuple(slice(a, b) if k == ax else slice(None) for k in range(nd))
    if periodic:
        u = f[sl(0, N)]
        c = coeff[sl(0, N)]
        res[sl(0, N)] = u + c * (np.roll(u, -1, axis=ax) - 2.0 * u + np.roll(u, 1, axis=ax))
        res[sl(N, N + 1)] = res[sl(0, 1)]
        return res
    c = coeff[sl(1, N)]
    res[sl(1, N)] = f[sl(1, N)] + c * (f[sl(2, N + 1)] - 2.0 * f[sl(1, N)] + f[sl(0, N - 1)])
    return res


def dccd_filter_nd(f, N, kinds, eps, psi):
    coeff = eps * ((2.0 * psi - 1.0) ** 2)
    res = f
    for ax in range(f.ndim):
        res = _filter_1d(res, ax, N[ax], coeff, kinds[ax] == "periodic")
    res = res.copy()
    for ax in range(f.ndim):
        if kinds[ax] == "periodic":
            continue
        lo = tuple(0 if k == ax else slice(None) for k in range(f.ndim))
        hi = tuple(N[ax] if k == ax else slice(None) for k in range(f.ndim))
        res[lo], res[hi] = f[lo], f[hi]
    return res


def gaussian_3x3(k):
    p = np.pad(k, 1, mode="edge")
    return (1.0 * p[:-2, :-2] + 2.0 * p[:-2, 1:-1] + 1.0 * p[:-2, 2:]
            + 2.0 * p[1:-1, :-2] + 4.0 * p[1:-1, 1:-1] + 2.0 * p[1:-1, 2:]
            + 1.0 * p[2:, :-2] + 2.0 * p[2:, 1:-1] + 1.0 * p[2:, 2:]) / 16.0


class OracleCurvaturePsi:
    def __init__(self, N, L, bc_type="wall", psi_min=0.01, gaussian_filter=False, dccd_eps=0.05):
        self.N = tuple(int(n) for n in N)
        self.ndim = len(self.N)
        self.kinds = axis_kinds(bc_type, self.ndim)
        self.ccd = OracleCCD(self.N, tuple(float(x) for x in L), bc_type)
        self.psi_min, self.gaussian, self.eps = psi_min, gaussian_filter, dccd_eps

    def compute(self, psi):
        psi = np.asarray(psi, dtype=np.float64)
        if self.eps > 0:
            def filt(f):
                return dccd_filter_nd(f, self.N, self.kinds, self.eps, psi)
        else:
            def filt(f):
                return f
        d1, d2 = [], []
        for ax in range(self.ndim):
            g1, g2 = self.ccd.differentiate(psi, ax)
            d1.append(filt(g1))
            d2.append(filt(g2))
        grad_sq = sum(g * g for g in d1)
        if self.ndim == 2:
            pxy = filt(self.ccd.first_derivative(d1[0], 1))
            num = d1[1] ** 2 * d2[0] - 2.0 * d1[0] * d1[1] * pxy + d1[0] ** 2 * d2[1]
            kappa = -num / (grad_sq + 1e-30) ** 1.5
        else:
            lap = sum(d2)
            con = sum(d1[i] ** 2 * d2[i] for i in range(3))
            for i in range(3):
                for j in range(i + 1, 3):
                    con = con + 2.0 * d1[i] * d1[j] * self.ccd.first_derivative(d1[i], j)
            gs = grad_sq + 1e-30
            kappa = -(lap - con / gs) / np.sqrt(gs)
        kappa[(psi <= self.psi_min) | (psi >= 1.0 - self.psi_min)] = 0.0
        if self.gaussian and self.ndim == 2:
            kappa = gaussian_3x3(kappa)
        return kappa
