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

Phase-separated coefficients (``ppe_coefficient_scheme="phase_separated"``)
with the three interface couplings follow
``ppe/fccd_matrixfree_lifecycle.py:44-117`` (density threshold
0.5 (rho_min + rho_max), none with affine_jump),
``ppe/fccd_matrixfree_helpers.py:99-160`` (per-phase control-volume means),
``:207-248`` (face coefficients: zero across the density threshold, or the
affine cut-face resistance of ``coupling/interface_stress_closure.py:225-277``),
``:251-264`` (Jacobi without pins under the mean gauge), ``:267-337``
(jump decomposition J = j (1 - psi)), ``:339-463`` (right-hand-side
compatibility), ``ppe/fccd_matrixfree.py:378-397`` (apply P L P + (I - P)),
``:451-504`` (rhs - L J; affine drive D_f[alpha_f B_Gamma(j)] with the
signed cut-face jump of ``interface_stress_closure.py:285-401``, Young-Laplace
j = -sigma kappa interpolated to the cut point, or the face average of a
given nodal jump), ``:506-611`` (solve, gauge, + J) and ``core/grid.py:
1123-1170`` (control volumes: h per node on wall axes, zero on periodic image
planes).

Pinning: ``fccd_symbol`` (closed form from the CCD symbol and the face
stencils) checks the periodic operator; cubic exactness of the face gradient
at face midpoints and the discrete divergence theorem (sum of w_i out_i over a
wall line is zero) check the wall operator; the GMRES solve is checked
against a dense direct solve of the assembled operator.  The phase-separated
forms are pinned by the reference's own cut-face resistance example
(``tests/test_interface_stress_closure.py``), the closed-form affine drive of
a straight interface, per-phase constants as fixed points of the gauged
operator with zero weighted means of its range, the jump decomposition's
exact recovery of J, and dense direct solves.  The face primitives
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


COEFFICIENT_SCHEMES = ("phase_density", "phase_separated")
INTERFACE_COUPLINGS = ("none", "jump_decomposition", "affine_jump")


class OracleFCCDPPE:
    """CPU FCCD PPE with the reference call surface (uniform grids)."""

    def __init__(self, N, L=None, bc_type="wall", preconditioner="none", pin=None,
                 coefficient_scheme="phase_density", interface_coupling="none"):
        # fccd_matrixfree.py:165-183
        if coefficient_scheme not in COEFFICIENT_SCHEMES:
            raise ValueError("FCCD PPE supports ppe_coefficient_scheme="
                             "'phase_density'|'phase_separated'")
        if interface_coupling not in INTERFACE_COUPLINGS:
            raise ValueError("FCCD PPE supports ppe_interface_coupling_scheme="
                             "'none'|'jump_decomposition'|'affine_jump'")
        if coefficient_scheme == "phase_density" and interface_coupling != "none":
            raise ValueError("phase_density PPE requires interface_coupling='none'")
        self.coefficient_scheme = coefficient_scheme
        self.interface_coupling = interface_coupling
        self.jump_ctx = None     # jump decomposition: psi, nodal jump, sigma
        self.stress_ctx = None   # affine coupling: psi, jump, kappa, sigma, threshold
        self.threshold = None
        self.coeffs = None
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

    # ---- interface data (fccd_matrixfree.py:230-304)
    def set_interface_jump_context(self, psi, kappa=None, sigma=0.0, pressure_jump=None,
                                   phase_threshold=0.5):
        """Young-Laplace j = p_gas - p_liquid = -sigma kappa with cut-face
        curvature (the nodal_cut_face route), or a given nodal jump
        (face-averaged, sigma 0: the jump decomposition stays off)."""
        psi = np.asarray(psi, dtype=np.float64)
        if pressure_jump is None:
            kappa = np.asarray(kappa, dtype=np.float64)
            jump = -float(sigma) * kappa
            self.jump_ctx = {"psi": psi, "jump": jump, "sigma": float(sigma)}
            self.stress_ctx = {"psi": psi, "jump": jump, "kappa": kappa, "sigma": float(sigma),
                               "threshold": float(phase_threshold), "cut": True}
        else:
            jump = np.asarray(pressure_jump, dtype=np.float64)
            self.jump_ctx = {"psi": psi, "jump": jump, "sigma": 0.0}
            self.stress_ctx = {"psi": psi, "jump": jump, "kappa": None, "sigma": 0.0,
                               "threshold": float(phase_threshold), "cut": False}

    def clear_interface_jump_context(self):
        self.jump_ctx = None
        self.stress_ctx = None

    def cell_volume(self):
        """core/grid.py:1146-1170: h per node on wall axes; on periodic axes
        h on nodes 0..N-1 (node 0 takes the wrapped half cell) and 0 on the
        image plane."""
        vol = np.ones(self.shape)
        for axis in range(self.ndim):
            w = np.full(self.N[axis] + 1, self.h(axis))
            if self.kinds[axis] == "periodic":
                w[-1] = 0.0
            shape = [1] * self.ndim
            shape[axis] = -1
            vol = vol * w.reshape(shape)
        return vol

    def prepare_operator(self, rho):
        """fccd_matrixfree_lifecycle.py:44-141: density threshold and phase
        masks (phase_separated without affine coupling), face coefficients."""
        self.rho = np.asarray(rho, dtype=np.float64)
        self.threshold = None
        if self.coefficient_scheme == "phase_separated" and \
                self.interface_coupling != "affine_jump":
            rmin, rmax = float(np.min(self.rho)), float(np.max(self.rho))
            if np.isfinite(rmin + rmax) and abs(rmax - rmin) >= 1.0e-14:
                self.threshold = 0.5 * (rmin + rmax)
        if self.threshold is not None:
            vol = self.cell_volume()
            self.gas = self.rho < self.threshold
            self.weights = (np.where(self.gas, vol, 0.0), np.where(self.gas, 0.0, vol))
            self.weight_sums = tuple(float(np.sum(w)) for w in self.weights)
        self.coeffs = [self._build_face_coeff(a) for a in range(self.ndim)]

    def uses_mean_gauge(self):
        return self.coefficient_scheme == "phase_separated" and self.threshold is not None

    def _build_face_coeff(self, axis):
        """fccd_matrixfree_helpers.py:207-248 on faces j+1/2, j = 0..N-1 (no
        wrap): 2/(rho_lo + rho_hi); zero across the density threshold
        (phase_separated); the affine cut-face resistance
        1/(theta rho_lo + (1 - theta) rho_hi), theta the psi cut point
        (interface_stress_closure.py:225-277)."""
        r = _front(self.rho, axis)
        r_lo, r_hi = r[:-1], r[1:]
        coeff = 2.0 / (r_lo + r_hi)
        if self.coefficient_scheme == "phase_separated" and \
                self.interface_coupling == "affine_jump":
            if self.stress_ctx is None:
                raise RuntimeError("phase_separated affine-jump face coefficient requires an "
                                   "interface-stress context")
            ps = _front(self.stress_ctx["psi"], axis)
            p_lo, p_hi = ps[:-1], ps[1:]
            t = self.stress_ctx["threshold"]
            cut = (p_lo < t) != (p_hi < t)
            den = np.where(cut, p_hi - p_lo, 1.0)
            theta = np.where(cut, (t - p_lo) / den, 0.0)
            with np.errstate(divide="ignore"):
                cut_coeff = 1.0 / (theta * r_lo + (1.0 - theta) * r_hi)
            return np.moveaxis(np.where(cut, cut_coeff, coeff), 0, axis)
        if self.coefficient_scheme != "phase_separated" or self.threshold is None:
            return np.moveaxis(coeff, 0, axis)
        same = (r_lo >= self.threshold) == (r_hi >= self.threshold)
        return np.moveaxis(np.where(same, coeff, 0.0), 0, axis)

    def face_coeff(self, axis):
        """The prepared coefficient of faces j+1/2, j = 0..N-1 along axis."""
        return _front(self.coeffs[axis], axis)

    def project_means(self, arr):
        """p - <p>_k per phase, <p>_k the control-volume mean over the
        phase (fccd_matrixfree_helpers.py:120-160)."""
        arr = np.asarray(arr, dtype=np.float64)
        m = [float(np.sum(w * arr)) / ws for w, ws in zip(self.weights, self.weight_sums)]
        return arr - (self.gas * m[0] + (~self.gas) * m[1])

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

    def divergence(self, faces):
        """Sum over axes of the face-flux divergence with the wall Neumann
        rows (fccd_matrixfree.py:746-818; periodic axes: fccd.py:490-502);
        faces[a] holds the N faces j+1/2 of axis a."""
        out = np.zeros(self.shape)
        for axis in range(self.ndim):
            F = _front(faces[axis], axis)
            o = _front(out, axis)
            n = self.N[axis]
            if self.kinds[axis] == "periodic":
                div = (F - np.roll(F, 1, axis=0)) * (1.0 / self.h(axis))
                o[:n] += div
                o[n] += div[0]
            else:
                inv_w = (1.0 / self.node_width(axis)).reshape((-1,) + (1,) * (self.ndim - 1))
                o[1:n] += (F[1:] - F[:-1]) * inv_w[1:n]
                o[0] += F[0] * inv_w[0]
                o[n] -= F[n - 1] * inv_w[n]
        return out

    def core(self, p):
        """D_f[(1/rho)_f G_f p] without gauge rows (fccd_matrixfree.py:399-417,
        769-818; periodic axes: fccd.py:490-502)."""
        return self.divergence([self.coeffs[a] * self.face_gradient(p, a)
                                for a in range(self.ndim)])

    def _sync_images(self, x):
        x = np.array(x, dtype=np.float64)
        if self.images is not None:
            x.ravel()[self.images] = x.ravel()[self.sources]
        return x

    def apply(self, p):
        """fccd_matrixfree.py:378-397: pin row identity, or under the phase
        mean gauge P L P p + (I - P) p; periodic image rows p_img - p_src."""
        p = np.asarray(p, dtype=np.float64)
        ps = self._sync_images(p)
        if self.uses_mean_gauge():
            pm = self.project_means(ps)
            out = self.project_means(self.core(pm))
            out = out + ps - pm
        else:
            out = self.core(ps)
            out.ravel()[self.pin] = p.ravel()[self.pin]
        if self.images is not None:
            out.ravel()[self.images] = p.ravel()[self.images] - p.ravel()[self.sources]
        return out

    def jacobi_inv(self):
        """fccd_matrixfree_helpers.py:251-264 (h_min per axis = face width;
        no pin under the mean gauge, fccd_matrixfree_lifecycle.py:86-96)."""
        diag = np.zeros(self.shape)
        for axis in range(self.ndim):
            x = np.linspace(0.0, self.L[axis], self.N[axis] + 1)
            h = float(np.min(x[1:] - x[:-1]))
            diag -= 2.0 / (self.rho * h * h)
        if not self.uses_mean_gauge():
            diag.ravel()[self.pin] = 1.0
        return 1.0 / np.where(np.abs(diag) > 1e-30, diag, 1.0)

    def jump_active(self):
        """fccd_matrixfree_helpers.py:281-292."""
        return (self.coefficient_scheme == "phase_separated" and
                self.interface_coupling == "jump_decomposition" and
                self.jump_ctx is not None and self.jump_ctx["sigma"] > 0.0)

    def jump_pressure(self):
        """J = j (1 - psi) of the jump decomposition (helpers.py:295-337)."""
        return self.jump_ctx["jump"] * (1.0 - self.jump_ctx["psi"])

    def signed_jump_gradient(self, axis):
        """B_Gamma(j) on the faces of axis (interface_stress_closure.py:285-401):
        (gas_hi - gas_lo) j_f / d_f on cut faces, j_f = -sigma kappa at the psi
        cut point (Young-Laplace) or the face average of a given nodal jump."""
        c = self.stress_ctx
        ps = _front(c["psi"], axis)
        p_lo, p_hi = ps[:-1], ps[1:]
        t = c["threshold"]
        gas_lo, gas_hi = p_lo < t, p_hi < t
        cut = gas_lo != gas_hi
        if c["cut"]:
            k = _front(c["kappa"], axis)
            den = np.where(cut, p_hi - p_lo, 1.0)
            theta = np.where(cut, (t - p_lo) / den, 0.0)
            jf = -c["sigma"] * ((1.0 - theta) * k[:-1] + theta * k[1:])
        else:
            j = _front(c["jump"], axis)
            jf = 0.5 * (j[:-1] + j[1:])
        orient = gas_hi.astype(np.float64) - gas_lo.astype(np.float64)
        B = np.where(cut, orient * jf, 0.0) / self.h(axis)
        return np.moveaxis(B, 0, axis)

    def affine_rhs(self):
        """D_f[alpha_f B_Gamma(j)] (fccd_matrixfree.py:462-504)."""
        if self.stress_ctx is None or self.stress_ctx.get("psi") is None:
            raise RuntimeError("phase_separated affine-jump FCCD PPE requires an "
                               "interface-stress context before applying affine RHS")
        return self.divergence([self.coeffs[a] * self.signed_jump_gradient(a)
                                for a in range(self.ndim)])

    def system_rhs(self, rhs):
        """The right-hand side GMRES sees (fccd_matrixfree.py:506-517): rhs - L J
        (jump decomposition), + the affine drive, the per-phase mean
        projection (mean gauge) or the pin, zero image rows."""
        b = np.array(rhs, dtype=np.float64)
        if self.jump_active():
            b = b - self.core(self.jump_pressure())
        if self.interface_coupling == "affine_jump":
            b = b + self.affine_rhs()
        if self.uses_mean_gauge():
            b = self.project_means(b)
        else:
            b.ravel()[self.pin] = 0.0
        if self.images is not None:
            b.ravel()[self.images] = 0.0
        return b

    def _gauge(self, x):
        x = np.array(x, dtype=np.float64)
        if self.uses_mean_gauge():
            return self.project_means(x)
        x.ravel()[self.pin] = 0.0
        return x

    def finish(self, p):
        """Gauge, image sync, + J (fccd_matrixfree.py:584-611)."""
        p = self._sync_images(self._gauge(p))
        if self.jump_active():
            p = self._sync_images(p + self.jump_pressure())
        return p

    def solve_direct(self, rhs, rho):
        """Dense direct solve of the gauged system (tiny grids only)."""
        self.prepare_operator(rho)
        b = self.system_rhs(rhs)
        x = np.linalg.solve(self.matrix(), b.ravel()).reshape(self.shape)
        return self.finish(x)

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
                return self._gauge(np.asarray(v).reshape(self.shape) * dinv).ravel()

            M = spla.LinearOperator((n, n), matvec=prec, dtype=np.float64)
        if p_init is None:
            x0 = np.zeros(n)
        else:
            x0 = self._gauge(self._sync_images(np.asarray(p_init).reshape(self.shape))).ravel()
        it = [0]

        def cb(_pr):
            it[0] += 1

        sol, info = spla.gmres(A, b, x0=x0, M=M, restart=restart, maxiter=maxiter,
                               atol=0.0, rtol=tol, callback=cb, callback_type="pr_norm")
        p = self._sync_images(self._gauge(sol.reshape(self.shape)))
        res = self.apply(p).ravel() - b
        nb = float(np.linalg.norm(b))
        rel = float(np.linalg.norm(res)) / nb if nb > 0 else float(np.linalg.norm(res))
        if self.jump_active():
            p = self._sync_images(p + self.jump_pressure())
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
