"""Oracle: variable-density FVM PPE, FP64, NumPy/SciPy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``PPESolverFVMMatrixFree`` and its assembled twin (paths relative
to ``src/twophase``):

* grid metrics            ``ppe/fvm_matrixfree_helpers.py:23-43``
* line coefficients       ``ppe/fvm_matrixfree_helpers.py:46-73``
* matrix-free apply       ``ppe/fvm_matrixfree.py:281-300``
* Jacobi state            ``ppe/fvm_matrixfree_helpers.py:95-103``,
                          ``ppe/fvm_matrixfree.py:325-331``
* CG system form          ``ppe/fvm_matrixfree.py:163-221, 345-350``
* CG / GMRES calls        ``ppe/gmres_helpers.py:16-80``
* GMRES form, line / Jacobi preconditioners
                          ``ppe/fvm_matrixfree.py:150-260, 302-330``,
                          ``ppe/fvm_matrixfree_helpers.py:76-101``
* Thomas / PCR line solves ``linalg_backend.py:77-115, 118-163, 270-340``
* assembled CSR           ``ppe/ppe_builder_helpers.py:50-122, 264-311``
* pin DOF                 ``core/boundary.py:181-191``
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def pin_dof(N) -> int:
    """Centre node (N//2, ...) for wall BC (core/boundary.py:181-191)."""
    shape = tuple(n + 1 for n in N)
    return int(np.ravel_multi_index(tuple(n // 2 for n in N), shape))


def grid_metrics(N, L):
    """Face widths and node control widths (fvm_matrixfree_helpers.py:23-43).

    Wall nodes carry half a cell (dv[0] = dv[-1] = h/2).
    """
    ndim = len(N)
    d_face, dv_node = [], []
    h_min = min(L[a] / N[a] for a in range(ndim))
    for ax in range(ndim):
        x = np.linspace(0.0, L[ax], N[ax] + 1)
        d = x[1:] - x[:-1]
        dv = np.empty_like(x)
        dv[0], dv[-1] = d[0] / 2.0, d[-1] / 2.0
        dv[1:-1] = (x[2:] - x[:-2]) / 2.0
        fs = [1] * ndim
        ns = [1] * ndim
        fs[ax], ns[ax] = d.size, dv.size
        d_face.append(d.reshape(fs))
        dv_node.append(dv.reshape(ns))
    return h_min, d_face, dv_node


def _sl(axis, start, stop, ndim):
    s = [slice(None)] * ndim
    s[axis] = slice(start, stop)
    return tuple(s)


def line_coeffs(rho, N, d_face, dv_node, axis):
    """lower/main/upper of one axis (fvm_matrixfree_helpers.py:46-73).

    Face coefficient 2/(rho_i+rho_{i+1})/d_face (harmonic mean of 1/rho).
    """
    nd = rho.ndim
    n = N[axis]
    lower, upper = np.zeros_like(rho), np.zeros_like(rho)
    cf = 2.0 / (rho[_sl(axis, 0, n, nd)] + rho[_sl(axis, 1, n + 1, nd)]) / d_face[axis]
    upper[_sl(axis, 0, n, nd)] = cf / dv_node[axis][_sl(axis, 0, n, nd)]
    lower[_sl(axis, 1, n + 1, nd)] = cf / dv_node[axis][_sl(axis, 1, n + 1, nd)]
    main = -(lower + upper)
    return lower, main, upper


def fvm_apply(p, coeffs, N, pin):
    """Pinned matrix-free operator (fvm_matrixfree.py:281-300): out[pin]=p[pin]."""
    nd = p.ndim
    out = np.zeros_like(p)
    for ax, (lo, mn, up) in enumerate(coeffs):
        n = N[ax]
        out += mn * p
        out[_sl(ax, 1, n + 1, nd)] += lo[_sl(ax, 1, n + 1, nd)] * p[_sl(ax, 0, n, nd)]
        out[_sl(ax, 0, n, nd)] += up[_sl(ax, 0, n, nd)] * p[_sl(ax, 1, n + 1, nd)]
    out.ravel()[pin] = p.ravel()[pin]
    return out


def cv_weight(shape, dv_node):
    """Tensor-product control volume W (fvm_matrixfree.py:345-350)."""
    w = np.ones(shape)
    for dv in dv_node:
        w = w * dv
    return w


def jacobi_diag_inv(coeffs, pin, shape):
    """diag = sum main, diag[pin]=1 (fvm_matrixfree_helpers.py:95-103)."""
    diag = np.zeros(shape)
    for _lo, mn, _up in coeffs:
        diag += mn
    diag.ravel()[pin] = 1.0
    return 1.0 / np.where(np.abs(diag) > 1e-30, diag, 1.0)


def _front(a, axis):
    return np.moveaxis(np.asarray(a, dtype=np.float64), axis, 0)


def thomas_lines(lo, di, up, r, axis):
    """Exact Thomas sweep on every line along ``axis``
    (linalg_backend.py:323-339); lo[0] and up[-1] are ignored."""
    a, d, c, b = (_front(x, axis) for x in (lo, di, up, r))
    n = d.shape[0]
    cp = np.empty_like(d)
    dp = np.empty_like(d)
    dp[0] = b[0] / d[0]
    if n > 1:
        cp[0] = c[0] / d[0]
    for i in range(1, n):
        den = d[i] - a[i] * cp[i - 1]
        if i < n - 1:
            cp[i] = c[i] / den
        dp[i] = (b[i] - a[i] * dp[i - 1]) / den
    x = np.empty_like(d)
    x[n - 1] = dp[n - 1]
    for i in range(n - 2, -1, -1):
        x[i] = dp[i] - cp[i] * x[i + 1]
    return np.moveaxis(x, 0, axis)


def pcr_stage_count(n, max_stages=None):
    """ceil(log2 n) stages (at least 1), capped at max(1, max_stages)
    (linalg_backend.py:146-150)."""
    full = max(1, int(np.ceil(np.log2(n)))) if n > 1 else 1
    return full if max_stages is None else min(full, max(1, int(max_stages)))


def pcr_reduce(lo, di, up, r, axis, stages):
    """``stages`` PCR reductions along ``axis`` (linalg_backend.py:77-115).

    Returns the reduced (lo, di, up, rw): row i then couples only unknowns
    i - 2^stages, i, i + 2^stages.  Division guards |d| > 0 as the reference.
    """
    a, d, c, b = (_front(x, axis).copy() for x in (lo, di, up, r))
    n = d.shape[0]
    s = 1
    for _ in range(stages):
        a2, d2, c2, b2 = a * 0.0, d.copy(), c * 0.0, b.copy()
        if s < n:
            dl = d[:-s]
            al = -a[s:] / np.where(np.abs(dl) > 0, dl, 1.0)       # rows i >= s
            a2[s:] = al * a[:-s]
            d2[s:] += al * c[:-s]
            b2[s:] += al * b[:-s]
            dr = d[s:]
            be = -c[:-s] / np.where(np.abs(dr) > 0, dr, 1.0)      # rows i < n - s
            d2[:-s] += be * a[s:]
            c2[:-s] = be * c[s:]
            b2[:-s] += be * b[s:]
        a, d, c, b = a2, d2, c2, b2
        s *= 2
    return tuple(np.moveaxis(x, 0, axis) for x in (a, d, c, b))


def pcr_lines(lo, di, up, r, axis, max_stages=None):
    """PCR solve rw / diag after the capped stage count (exact when
    2^stages >= n; linalg_backend.py:118-163)."""
    n = np.asarray(di).shape[axis]
    _a, d, _c, b = pcr_reduce(lo, di, up, r, axis, pcr_stage_count(n, max_stages))
    return b / d


class OraclePPE:
    """CPU FVM PPE in the reference's CG form (uniform wall grids).

    The CG system is ``(-W L) p = -W rhs`` with the pin row and column set to
    identity (fvm_matrixfree.py:163-202); it is SPD.  ``system_matvec`` and
    ``system_rhs`` expose it so a residual can be recomputed on any p.
    """

    def __init__(self, N, L=None):
        self.N = tuple(int(n) for n in N)
        self.L = tuple(float(x) for x in (L or (1.0,) * len(self.N)))
        self.ndim = len(self.N)
        self.shape = tuple(n + 1 for n in self.N)
        self.pin = pin_dof(self.N)
        self.h_min, self.d_face, self.dv_node = grid_metrics(self.N, self.L)
        self.W = cv_weight(self.shape, self.dv_node)
        self.coeffs = None

    def prepare_operator(self, rho):
        rho = np.asarray(rho, dtype=np.float64)
        self._rho = rho
        self.coeffs = [line_coeffs(rho, self.N, self.d_face, self.dv_node, a)
                       for a in range(self.ndim)]
        self.diag_inv = jacobi_diag_inv(self.coeffs, self.pin, self.shape)

    def apply(self, p):
        return fvm_apply(np.asarray(p, dtype=np.float64), self.coeffs, self.N, self.pin)

    # --- CG system (fvm_matrixfree.py:190-212)
    def system_rhs(self, rhs):
        b = -(np.asarray(rhs, dtype=np.float64) * self.W).ravel().copy()
        b[self.pin] = 0.0
        return b

    def system_matvec(self, p_flat):
        p = np.asarray(p_flat, dtype=np.float64).copy()
        pinned = p[self.pin]
        p[self.pin] = 0.0
        out = -(self.apply(p.reshape(self.shape)) * self.W).ravel()
        out[self.pin] = pinned
        return out

    def system_precond(self, r_flat):
        r = np.asarray(r_flat, dtype=np.float64)
        z = (r.reshape(self.shape) / self.W) * self.diag_inv
        z.ravel()[self.pin] = 0.0
        z = -z.ravel()
        z[self.pin] = r[self.pin]
        return z

    def relres(self, rhs, p):
        """FP64 ||b - A p|| / ||b|| on the CG system."""
        b = self.system_rhs(rhs)
        r = b - self.system_matvec(np.asarray(p).ravel())
        return float(np.linalg.norm(r) / np.linalg.norm(b))

    def solve_cg(self, rhs, rho, tol=1e-8, maxiter=500, p_init=None,
                 preconditioner="jacobi"):
        """Jacobi-PCG through SciPy ``cg`` (gmres_helpers.py:44-80).

        Stops on the recursively updated unpreconditioned residual
        (SciPy 1.15 semantics); raises RuntimeError when it does not converge
        (fvm_matrixfree.py:250-256).  Returns (p, info_dict).
        """
        self.prepare_operator(rho)
        n = int(np.prod(self.shape))
        A = spla.LinearOperator((n, n), matvec=self.system_matvec, dtype=np.float64)
        M = None
        if preconditioner == "jacobi":
            M = spla.LinearOperator((n, n), matvec=self.system_precond, dtype=np.float64)
        x0 = np.zeros(n) if p_init is None else np.asarray(p_init, dtype=np.float64).ravel().copy()
        x0[self.pin] = 0.0
        it = [0]

        def cb(_xk):
            it[0] += 1

        sol, info = spla.cg(A, self.system_rhs(rhs), x0=x0, M=M, maxiter=maxiter,
                            atol=0.0, rtol=tol, callback=cb)
        if info != 0:
            raise RuntimeError(f"OraclePPE CG did not converge (info={info})")
        p = sol.reshape(self.shape).copy()
        p.ravel()[self.pin] = 0.0
        return p, {"iterations": it[0], "relres": self.relres(rhs, p)}

    # --- GMRES form (fvm_matrixfree.py:150-260, 302-330)
    def shifted_line_coeffs(self, c_tau=2.0):
        """(-lower, shift - main, -upper) per axis, shift = 2/(c_tau rho h_min^2)
        (fvm_matrixfree_helpers.py:76-90); needs prepare_operator first."""
        shift = 2.0 / (c_tau * self._rho * self.h_min ** 2)
        return [(-lo, shift - mn, -up) for lo, mn, up in self.coeffs]

    def gmres_precond(self, r, preconditioner="line_pcr", c_tau=2.0, max_stages=None):
        """z = M^-1 r of the GMRES form, z[pin] = 0.  line_pcr: per-axis line
        solves (Thomas when max_stages is None, else capped PCR) averaged
        over axes; jacobi: r * diag_inv."""
        r = np.asarray(r, dtype=np.float64).reshape(self.shape)
        if preconditioner == "jacobi":
            z = r * self.diag_inv
        elif preconditioner == "line_pcr":
            rw = r.copy()
            rw.ravel()[self.pin] = 0.0
            z = np.zeros_like(rw)
            for ax, (lo, di, up) in enumerate(self.shifted_line_coeffs(c_tau)):
                if max_stages is None:
                    z += thomas_lines(lo, di, up, rw, ax)
                else:
                    z += pcr_lines(lo, di, up, rw, ax, max_stages)
            z /= float(self.ndim)
        else:
            raise ValueError(preconditioner)
        z = z.copy()
        z.ravel()[self.pin] = 0.0
        return z

    def solve_gmres(self, rhs, rho, tol=1e-8, maxiter=500, restart=None,
                    preconditioner="line_pcr", c_tau=2.0, max_stages=None, p_init=None):
        """SciPy ``gmres`` on L p = rhs, rhs[pin] = 0 (gmres_helpers.py:16-42).
        Returns (p, {"info", "iterations", "relres"}); p[pin] = 0."""
        self.prepare_operator(rho)
        n = int(np.prod(self.shape))
        b = np.asarray(rhs, dtype=np.float64).ravel().copy()
        b[self.pin] = 0.0
        restart = min(80, max(20, maxiter)) if restart is None else int(restart)
        A = spla.LinearOperator(
            (n, n), matvec=lambda v: self.apply(np.asarray(v).reshape(self.shape)).ravel(),
            dtype=np.float64)
        M = None
        if preconditioner in ("jacobi", "line_pcr"):
            M = spla.LinearOperator(
                (n, n), matvec=lambda v: self.gmres_precond(v, preconditioner, c_tau,
                                                            max_stages).ravel(),
                dtype=np.float64)
        x0 = np.zeros(n) if p_init is None else np.asarray(p_init, dtype=np.float64).ravel().copy()
        x0[self.pin] = 0.0
        it = [0]

        def cb(_pr):
            it[0] += 1

        sol, info = spla.gmres(A, b, x0=x0, M=M, restart=restart, maxiter=maxiter,
                               atol=0.0, rtol=tol, callback=cb, callback_type="pr_norm")
        p = sol.reshape(self.shape).copy()
        p.ravel()[self.pin] = 0.0
        res = float(np.linalg.norm(b - self.apply(sol.reshape(self.shape)).ravel())
                    / np.linalg.norm(b)) if np.linalg.norm(b) > 0 else 0.0
        return p, {"info": int(info), "iterations": it[0], "relres": res}

    # --- assembled twin (ppe_builder_helpers.py:50-122, 264-311)
    def csr(self, rho):
        """Assembled operator: uniform faces a_f/h^2, doubled on the wall node
        side (dv = h/2 there), pin row -> identity."""
        rho = np.asarray(rho, dtype=np.float64)
        shape = self.shape
        n = int(np.prod(shape))
        idx = np.arange(n).reshape(shape)
        rf = rho.ravel()
        data, rows, cols = [], [], []
        for ax in range(self.ndim):
            Nax = self.N[ax]
            h = self.L[ax] / Nax
            iL = idx[_sl(ax, 0, Nax, self.ndim)].ravel()
            iR = idx[_sl(ax, 1, Nax + 1, self.ndim)].ravel()
            pos_L = np.indices(shape)[ax][_sl(ax, 0, Nax, self.ndim)].ravel()
            pos_R = np.indices(shape)[ax][_sl(ax, 1, Nax + 1, self.ndim)].ravel()
            a_f = 2.0 / (rf[iL] + rf[iR]) / (h * h)
            cL = np.where(pos_L == 0, 2.0 * a_f, a_f)
            cR = np.where(pos_R == Nax, 2.0 * a_f, a_f)
            data += [cL, cR, -cL, -cR]
            rows += [iL, iR, iL, iR]
            cols += [iR, iL, iL, iR]
        data = np.concatenate(data)
        rows = np.concatenate(rows)
        cols = np.concatenate(cols)
        keep = rows != self.pin
        data = np.concatenate([data[keep], [1.0]])
        rows = np.concatenate([rows[keep], [self.pin]])
        cols = np.concatenate([cols[keep], [self.pin]])
        return sp.csr_matrix((data, (rows, cols)), shape=(n, n))

    def solve_direct(self, rhs, rho):
        """SuperLU solve of the assembled twin (fvm_spsolve.py:117-161)."""
        A = self.csr(rho)
        b = np.asarray(rhs, dtype=np.float64).ravel().copy()
        b[self.pin] = 0.0
        p = spla.spsolve(A.tocsc(), b).reshape(self.shape)
        p.ravel()[self.pin] = 0.0
        return p
