"""Oracle: FCCD matrix-free PPE operator and its GMRES solve, FP64, NumPy/SciPy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``PPESolverFCCDMatrixFree`` in its default configuration
(``ppe_coefficient_scheme="phase_density"``, no interface coupling,
``boundary_face_space="full_face"``, legacy operator pairing) on uniform grids
(paths relative to ``src/twophase``):

* face gradient           ``ccd/fccd.py:243-277`` with ``_face_gradient_kernel``
                          (``ccd/fccd.py:95-98``) and ``build_axis_weights``
                          (``ccd/fccd_helpers.py:17-35``)
* face coefficients       ``ppe/fccd_matrixfree_helpers.py:207-248`` (2/(rho_lo+rho_hi))
* node widths             ``ppe/fccd_matrixfree_helpers.py:35-51``
* weighted divergence     ``ppe/fccd_matrixfree.py:769-818`` (wall) and
                          ``ccd/fccd.py:464-502`` (periodic)
* apply (pin, image rows) ``ppe/fccd_matrixfree.py:344-397``
* Jacobi inverse          ``ppe/fccd_matrixfree_helpers.py:251-264``
* solve                   ``ppe/fccd_matrixfree.py:505-600`` with
                          ``ppe/gmres_helpers.py:16-42`` (SciPy gmres,
                          restart None -> SciPy's 20)

The nodal second derivative q along each axis is the CCD oracle's
(``oracle/ccd.py``), as ``FCCDSolver.face_gradient`` takes it from
``CCDSolver.second_derivative``.

``OracleFCCD`` restates FCCDSolver's face primitives (``ccd/fccd.py:95-131,
243-521``): face gradient, face value, face second derivative, upwind Taylor
face state, R4 Hermite node gradient and face divergence.

Pinning: ``fccd_symbol`` (closed form from the CCD symbol and the face
stencils) checks the periodic operator; cubic exactness of the face gradient
at face midpoints and the discrete divergence theorem (sum of w_i out_i over a
wall line is zero) check the wall operator; the GMRES solve is checked
against a dense direct solve of the assembled operator.  The face primitives
are pinned by exactness for cubics (face value, node gradient) and quadratics
(upwind state), linear face fields (divergence), and the reference's own
truncated symbol of the periodic face gradient, met to O((wH)^6).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse.linalg as spla

from .ccd import OracleCCD, axis_kinds, periodic_symbol


def _front(a, axis):
    return np.moveaxis(a, axis, 0)


class OracleFCCDPPE:
    """CPU FCCD PPE with the reference call surface (uniform grids)."""

    def __init__(self, N, L=None, bc_type="wall", preconditioner="none", pin=None):
        self.N = tuple(int(n) for n in N)
        self.ndim = len(self.N)
        self.L = tuple(float(x) for x in (L if L is not None else (1.0,) * self.ndim))
        self.kinds = axis_kinds(bc_type, self.ndim)
        self.shape = tuple(n + 1 for n in self.N)
        self.ccd = OracleCCD(self.N, self.L, bc_type)
        # centre node, fccd_matrixfree.py:187-192
        self.pin = int(np.ravel_multi_index(tuple(n // 2 for n in self.N), self.shape)
                       if pin is None else pin)
        self.preconditioner = preconditioner
        self.images, self.sources = self._image_rows()
        self.rho = None

    # fccd_matrixfree.py:300-345: image node -> source node with that axis index 0;
    # the first periodic axis listing a node decides its source
    def _image_rows(self):
        table = {}
        for axis in range(self.ndim):
            if self.kinds[axis] != "periodic":
                continue
            idx = np.indices(self.shape).reshape(self.ndim, -1)
            sel = idx[axis] == self.N[axis]
            img = np.ravel_multi_index(idx[:, sel], self.shape)
            src_idx = idx[:, sel].copy()
            src_idx[axis] = 0
            src = np.ravel_multi_index(src_idx, self.shape)
            for i, s in zip(img.tolist(), src.tolist()):
                table.setdefault(i, s)
        if not table:
            return None, None
        images = np.array(sorted(table), dtype=np.intp)
        return images, np.array([table[i] for i in images], dtype=np.intp)

    def h(self, axis):
        return self.L[axis] / self.N[axis]

    def node_width(self, axis):
        """fccd_matrixfree_helpers.py:35-51: half cells at the ends."""
        x = np.linspace(0.0, self.L[axis], self.N[axis] + 1)
        fw = x[1:] - x[:-1]
        w = np.empty_like(x)
        w[0], w[-1] = 0.5 * fw[0], 0.5 * fw[-1]
        w[1:-1] = 0.5 * (x[2:] - x[:-2])
        return w

    def prepare_operator(self, rho):
        self.rho = np.asarray(rho, dtype=np.float64)

    def face_coeff(self, axis):
        """2 / (rho_lo + rho_hi) on faces j+1/2, j = 0..N-1 (no wrap)."""
        r = _front(self.rho, axis)
        return 2.0 / (r[:-1] + r[1:])

    def face_gradient(self, p, axis, q=None):
        """(u_hi - u_lo)/H - (H/24)(q_hi - q_lo); periodic axes wrap the
        unique nodes (fccd.py:212-232, 243-277)."""
        if q is None:
            q = self.ccd.second_derivative(p, axis)
        H = self.h(axis)
        u, qq = _front(np.asarray(p, dtype=np.float64), axis), _front(q, axis)
        if self.kinds[axis] == "periodic":
            n = self.N[axis]
            u_lo, u_hi = u[:n], np.roll(u[:n], -1, axis=0)
            q_lo, q_hi = qq[:n], np.roll(qq[:n], -1, axis=0)
        else:
            u_lo, u_hi, q_lo, q_hi = u[:-1], u[1:], qq[:-1], qq[1:]
        g = (u_hi - u_lo) * (1.0 / H) - (H / 24.0) * (q_hi - q_lo)
        return np.moveaxis(g, 0, axis)

    def core(self, p):
        """D_f[(1/rho)_f G_f p] without gauge rows (fccd_matrixfree.py:399-417,
        769-818; periodic axes: fccd.py:490-502)."""
        out = np.zeros(self.shape)
        for axis in range(self.ndim):
            g = _front(self.face_gradient(p, axis), axis)
            c = self.face_coeff(axis)
            o = _front(out, axis)
            n = self.N[axis]
            if self.kinds[axis] == "periodic":
                F = c * g
                div = (F - np.roll(F, 1, axis=0)) * (1.0 / self.h(axis))
                o[:n] += div
                o[n] += div[0]
            else:
                inv_w = (1.0 / self.node_width(axis)).reshape((-1,) + (1,) * (self.ndim - 1))
                o[1:n] += (c[1:] * g[1:] - c[:-1] * g[:-1]) * inv_w[1:n]
                o[0] += c[0] * g[0] * inv_w[0]
                o[n] -= c[n - 1] * g[n - 1] * inv_w[n]
        return out

    def apply(self, p):
        """Pin row identity, periodic image rows p_img - p_src
        (fccd_matrixfree.py:378-397)."""
        p = np.asarray(p, dtype=np.float64)
        ps = p.copy()
        if self.images is not None:  # _sync_periodic_images
            ps.ravel()[self.images] = ps.ravel()[self.sources]
        out = self.core(ps)
        out.ravel()[self.pin] = p.ravel()[self.pin]
        if self.images is not None:
            out.ravel()[self.images] = p.ravel()[self.images] - p.ravel()[self.sources]
        return out

    def jacobi_inv(self):
        """fccd_matrixfree_helpers.py:251-264 (h_min per axis = face width)."""
        diag = np.zeros(self.shape)
        for axis in range(self.ndim):
            x = np.linspace(0.0, self.L[axis], self.N[axis] + 1)
            h = float(np.min(x[1:] - x[:-1]))
            diag -= 2.0 / (self.rho * h * h)
        diag.ravel()[self.pin] = 1.0
        return 1.0 / np.where(np.abs(diag) > 1e-30, diag, 1.0)

    def system_rhs(self, rhs):
        """Pinned rhs copy with zero image rows (fccd_matrixfree.py:514-516, 370-376)."""
        b = np.asarray(rhs, dtype=np.float64).copy()
        b.ravel()[self.pin] = 0.0
        if self.images is not None:
            b.ravel()[self.images] = 0.0
        return b

    def matrix(self):
        """Dense operator, column by column (tiny grids only)."""
        n = int(np.prod(self.shape))
        A = np.empty((n, n))
        e = np.zeros(n)
        for j in range(n):
            e[j] = 1.0
            A[:, j] = self.apply(e.reshape(self.shape)).ravel()
            e[j] = 0.0
        return A

    def solve(self, rhs, rho, tol=1e-8, maxiter=500, restart=None, p_init=None):
        """SciPy gmres on the pinned system (fccd_matrixfree.py:505-600).
        Returns (p, info dict with info, iterations, relres)."""
        self.prepare_operator(rho)
        b = self.system_rhs(rhs).ravel()
        n = b.size
        A = spla.LinearOperator((n, n), dtype=np.float64,
                                matvec=lambda v: self.apply(np.asarray(v).reshape(self.shape)).ravel())
        M = None
        if self.preconditioner == "jacobi":
            dinv = self.jacobi_inv()

            def prec(v):
                z = np.asarray(v).reshape(self.shape) * dinv
                z = z.copy()
                z.ravel()[self.pin] = 0.0
                return z.ravel()

            M = spla.LinearOperator((n, n), matvec=prec, dtype=np.float64)
        if p_init is None:
            x0 = np.zeros(n)
        else:
            x0 = np.asarray(p_init, dtype=np.float64).copy()
            if self.images is not None:
                x0.ravel()[self.images] = x0.ravel()[self.sources]
            x0 = x0.ravel()
            x0[self.pin] = 0.0
        it = [0]

        def cb(_pr):
            it[0] += 1

        sol, info = spla.gmres(A, b, x0=x0, M=M, restart=restart, maxiter=maxiter,
                               atol=0.0, rtol=tol, callback=cb, callback_type="pr_norm")
        p = sol.reshape(self.shape).copy()
        p.ravel()[self.pin] = 0.0
        if self.images is not None:
            p.ravel()[self.images] = p.ravel()[self.sources]
        res = self.apply(p).ravel() - b
        nb = float(np.linalg.norm(b))
        rel = float(np.linalg.norm(res)) / nb if nb > 0 else float(np.linalg.norm(res))
        return p, {"info": int(info), "iterations": it[0], "relres": rel}


def fccd_symbol(theta, h):
    """Periodic, constant-density symbol of D_f G_f per unit 1/rho on mode
    e^{i theta j}: (2 cos theta - 2)/h^2 * (1 - (h^2/24) * s2), s2 the CCD f''
    symbol (the face gradient's two terms share the factor e^{i theta} - 1)."""
    _, s2 = periodic_symbol(theta, h)
    return (2.0 * np.cos(theta) - 2.0) / (h * h) * (1.0 - (h * h / 24.0) * s2)


class OracleFCCD:
    """CPU restatement of FCCDSolver's face primitives (ccd/fccd.py:165-521,
    uniform weights ccd/fccd_helpers.py:17-35), on the CCD oracle's q."""

    def __init__(self, N, L, bc_type="wall"):
        self.N = tuple(int(n) for n in N)
        self.ndim = len(self.N)
        self.L = tuple(float(x) for x in L)
        self.kinds = axis_kinds(bc_type, self.ndim)
        self.ccd = OracleCCD(self.N, self.L, bc_type)

    def h(self, axis):
        return self.L[axis] / self.N[axis]

    def _pair(self, a, axis):
        """(lo, hi) nodal values of every face (ccd/fccd.py:212-232)."""
        a = _front(np.asarray(a, dtype=np.float64), axis)
        if self.kinds[axis] == "periodic":
            n = self.N[axis]
            return a[:n], np.roll(a[:n], -1, axis=0)
        return a[:-1], a[1:]

    def _q(self, u, axis, q):
        return self.ccd.second_derivative(u, axis) if q is None else q

    def face_gradient(self, u, axis, q=None):
        q = self._q(u, axis, q)
        H = self.h(axis)
        (ul, uh), (ql, qh) = self._pair(u, axis), self._pair(q, axis)
        return np.moveaxis((uh - ul) * (1.0 / H) - (H / 24.0) * (qh - ql), 0, axis)

    def face_value(self, u, axis, q=None):
        """ccd/fccd.py:101-109, 279-299."""
        q = self._q(u, axis, q)
        H = self.h(axis)
        (ul, uh), (ql, qh) = self._pair(u, axis), self._pair(q, axis)
        return np.moveaxis(0.5 * (ul + uh) - (H * H / 16.0) * (ql + qh), 0, axis)

    def face_second_derivative(self, u, axis, q=None):
        """ccd/fccd.py:112-114, 301-317."""
        q = self._q(u, axis, q)
        ql, qh = self._pair(q, axis)
        return np.moveaxis(0.5 * (ql + qh), 0, axis)

    def upwind_face_value(self, u, vel, axis, d=None, q=None):
        """ccd/fccd.py:115-125, 333-390."""
        if d is None:
            d = self.ccd.first_derivative(u, axis)
        q = self._q(u, axis, q)
        H = self.h(axis)
        (ul, uh), (dl, dh), (ql, qh) = self._pair(u, axis), self._pair(d, axis), self._pair(q, axis)
        left = ul + 0.5 * H * dl + (H * H / 8.0) * ql
        right = uh - 0.5 * H * dh + (H * H / 8.0) * qh
        v = _front(np.asarray(vel, dtype=np.float64), axis)
        return np.moveaxis(np.where(v >= 0.0, left, right), 0, axis)

    def node_gradient(self, u, axis, q=None):
        """R4 Hermite, one-sided at wall nodes (ccd/fccd.py:128-131, 392-462)."""
        q = self._q(u, axis, q)
        H = self.h(axis)
        d = _front(self.face_gradient(u, axis, q=q), axis)
        qq = _front(q, axis)
        n = self.N[axis]
        out = np.empty((n + 1,) + d.shape[1:])
        if self.kinds[axis] == "periodic":
            qu = qq[:n]
            g = 0.5 * (np.roll(d, 1, axis=0) + d) - (H / 16.0) * (np.roll(qu, -1, axis=0)
                                                                   - np.roll(qu, 1, axis=0))
            out[:n], out[n] = g, g[0]
        else:
            out[1:n] = 0.5 * (d[:-1] + d[1:]) - (H / 16.0) * (qq[2:] - qq[:-2])
            out[0], out[n] = d[0], d[n - 1]
        return np.moveaxis(out, 0, axis)

    def face_divergence(self, F, axis):
        """ccd/fccd.py:464-521 (uniform)."""
        H = self.h(axis)
        f = _front(np.asarray(F, dtype=np.float64), axis)
        n = self.N[axis]
        out = np.zeros((n + 1,) + f.shape[1:])
        if self.kinds[axis] == "periodic":
            div = (f - np.roll(f, 1, axis=0)) * (1.0 / H)
            out[:n], out[n] = div, div[0]
        else:
            out[1:n] = (f[1:] - f[:-1]) * (1.0 / H)
        return np.moveaxis(out, 0, axis)


def fccd_face_symbol_truncated(omega, H):
    """The reference's truncated face-operator symbol i w (1 - 7 (w H)^4 / 5760)
    (ccd/fccd_helpers.py:65-67)."""
    return 1j * omega * (1.0 - 7.0 * (omega * H) ** 4 / 5760.0)
