"""Oracle: 6th-order combined compact difference (CCD), FP64, NumPy/SciPy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates the reference CCD path (paths relative to ``src/twophase``):

* coefficients           ``ccd/ccd_solver.py:70-76``
* wall closures          ``ccd/ccd_solver.py:408-465``
* wall axis system       ``ccd/ccd_solver_helpers.py:227-314``
* wall differentiate     ``ccd/ccd_solver_helpers.py:38-77``
* periodic axis system   ``ccd/ccd_solver_helpers.py:406-423``
* periodic differentiate ``ccd/ccd_solver_helpers.py:426-465``
* block-Thomas template  ``ccd/block_tridiag.py:29-175``
* wall Neumann zeroing   ``ccd/ccd_boundary_helpers.py:52-61``
* grid metric            ``core/metrics.py:22-88`` (non-uniform branch),
  ``ccd/ccd_solver_helpers.py:509-519`` (``apply_ccd_metric``) and the
  computational spacing 1/N of non-uniform grids (``ccd/ccd_solver.py:126-130``)

The dense systems are solved with ``scipy.linalg.lu_factor``/``lu_solve``
exactly as the reference CPU path does (``ccd_solver_helpers.py:17, 270, 422,
449``), so this module doubles as the reference-CPU-path baseline.
"""

from __future__ import annotations

import math

import numpy as np
import scipy.linalg as sla

# Table 1 of the paper, ccd/ccd_solver.py:70-76
ALPHA1 = 7.0 / 16.0
A1 = 15.0 / 16.0
B1 = 1.0 / 16.0
BETA2 = -1.0 / 8.0
A2 = 3.0
B2 = -9.0 / 8.0


def lower_block(h: float) -> np.ndarray:
    """Coupling of row i to unknown block i-1 (ccd_solver_helpers.py:235)."""
    return np.array([[ALPHA1, B1 * h], [B2 / h, BETA2]])


def upper_block(h: float) -> np.ndarray:
    """Coupling of row i to unknown block i+1 (ccd_solver_helpers.py:236)."""
    return np.array([[ALPHA1, -B1 * h], [-B2 / h, BETA2]])


def boundary_coeffs_left(h: float, n_pts: int) -> dict:
    """One-sided closure at node 0 (ccd_solver.py:408-433)."""
    require_wall_points(n_pts)
    return {
        "M": np.array([[-1.5, 1.5 * h], [0.0, 0.0]]),
        "c_I": np.array([-23.0 / 6.0, 21.0 / 4.0, -1.5, 1.0 / 12.0]) / h,
        "c_II": np.array([45.0, -154.0, 214.0, -156.0, 61.0, -10.0]) / (12.0 * h * h),
    }


def boundary_coeffs_right(h: float, n_pts: int) -> dict:
    """Mirror closure at node N, h -> -h (ccd_solver.py:436-456).

    ``c_I``/``c_II`` apply to f[N], f[N-1], ... in that order.
    """
    require_wall_points(n_pts)
    return {
        "M": np.array([[-1.5, -1.5 * h], [0.0, 0.0]]),
        "c_I": np.array([23.0 / 6.0, -21.0 / 4.0, 1.5, -1.0 / 12.0]) / h,
        "c_II": np.array([45.0, -154.0, 214.0, -156.0, 61.0, -10.0]) / (12.0 * h * h),
    }


def require_wall_points(n_pts: int) -> None:
    """Wall closure needs six points (ccd_solver.py:459-465)."""
    if n_pts < 6:
        raise ValueError(
            "CCD wall Eq-II boundary closure requires n_pts >= 6 for the "
            "six-point O(h^4) formula; no lower-order fallback exists."
        )


# --------------------------------------------------------------------------
# Wall axis
# --------------------------------------------------------------------------

def build_wall_axis(n_pts: int, h: float) -> dict:
    """Dense interior system with closure-modified end blocks.

    Restates ccd_solver_helpers.py:227-270: n_int = N-1 unknown 2-blocks;
    D_0 += L @ M_left, D_last += U @ M_right, and the couplings of the first
    and last rows to the boundary nodes move to the right-hand side.
    """
    N = n_pts - 1
    n_int = N - 1
    if n_int < 1:
        raise ValueError(f"Need >= 3 grid points; got {n_pts}")
    bl = boundary_coeffs_left(h, n_pts)
    br = boundary_coeffs_right(h, n_pts)
    Lb, Ub = lower_block(h), upper_block(h)
    A = np.zeros((2 * n_int, 2 * n_int))
    for i in range(n_int):
        D = np.eye(2)
        if i == 0:
            D = D + Lb @ bl["M"]
        if i == n_int - 1:
            D = D + Ub @ br["M"]
        A[2 * i:2 * i + 2, 2 * i:2 * i + 2] = D
        if i >= 1:
            A[2 * i:2 * i + 2, 2 * (i - 1):2 * i] = Lb
        if i <= n_int - 2:
            A[2 * i:2 * i + 2, 2 * (i + 1):2 * (i + 2)] = Ub
    lu, piv = sla.lu_factor(A)
    return {"A": A, "lu": lu, "piv": piv, "h": h, "N": N, "n_int": n_int,
            "L0": Lb, "UN": Ub, "bc_left": bl, "bc_right": br}


def _as_lines(data, axis):
    f = np.moveaxis(np.asarray(data, dtype=np.float64), axis, 0)
    shape = f.shape
    return f.reshape(shape[0], -1), shape


def _from_lines(arr, shape, axis):
    return np.moveaxis(arr.reshape(shape), 0, axis)


def differentiate_wall(data, axis: int, info: dict):
    """Wall-BC CCD along ``axis`` (ccd_solver_helpers.py:38-77)."""
    h, N, n_int = info["h"], info["N"], info["n_int"]
    f, shape = _as_lines(data, axis)
    f_m1, f_0, f_p1 = f[0:n_int], f[1:n_int + 1], f[2:n_int + 2]
    rhs = np.empty((n_int, 2, f.shape[1]))
    rhs[:, 0, :] = (A1 / h) * (f_p1 - f_m1)
    rhs[:, 1, :] = (A2 / (h * h)) * (f_m1 - 2.0 * f_0 + f_p1)
    bl, br = info["bc_left"], info["bc_right"]
    bc_lo = np.stack([bl["c_I"] @ f[:4], bl["c_II"] @ f[:6]])
    f_rev = f[N::-1]
    bc_hi = np.stack([br["c_I"] @ f_rev[:4], br["c_II"] @ f_rev[:6]])
    rhs[0] -= info["L0"] @ bc_lo
    rhs[-1] -= info["UN"] @ bc_hi
    sol = sla.lu_solve((info["lu"], info["piv"]), rhs.reshape(2 * n_int, -1))
    sol = sol.reshape(n_int, 2, -1)
    out = np.empty((2, N + 1, f.shape[1]))
    out[:, 1:-1, :] = sol.transpose(1, 0, 2)
    out[:, 0, :] = bl["M"] @ out[:, 1, :] + bc_lo
    out[:, N, :] = br["M"] @ out[:, N - 1, :] + bc_hi
    return _from_lines(out[0], shape, axis), _from_lines(out[1], shape, axis)


def differentiate_wall_extended(data, axis: int, info: dict):
    """The wall system of ``differentiate_wall`` solved in 80-bit long double
    (block Thomas, right-hand side and closure rows formed in long double from
    the FP64 input): the exact discrete answer to ~1e-18, the yardstick on
    long wall lines where the FP64 dense LU's own error (second-difference
    cancellation, the closure rows) reaches 1e-10 of max|f''| at N = 1024.
    Returns FP64 (d1, d2)."""
    LD = np.longdouble
    N, n_int = info["N"], info["n_int"]
    hL = LD(info["h"])
    a1, a2 = LD(15) / 16, LD(3)
    al, b1, be, b2 = LD(7) / 16, LD(1) / 16, LD(-1) / 8, LD(-9) / 8
    Lb = np.array([[al, b1 * hL], [b2 / hL, be]], dtype=LD)
    Ub = np.array([[al, -b1 * hL], [-b2 / hL, be]], dtype=LD)
    bl, br = info["bc_left"], info["bc_right"]
    Ml, Mr = np.array(bl["M"], dtype=LD), np.array(br["M"], dtype=LD)
    f, shape = _as_lines(data, axis)
    fL = f.astype(LD)
    rhs = np.empty((n_int, 2, fL.shape[1]), dtype=LD)
    rhs[:, 0] = (a1 / hL) * (fL[2:] - fL[:-2])
    rhs[:, 1] = (a2 / (hL * hL)) * (fL[:-2] - 2 * fL[1:-1] + fL[2:])
    bc_lo = np.stack([np.array(bl["c_I"], dtype=LD) @ fL[:4],
                      np.array(bl["c_II"], dtype=LD) @ fL[:6]])
    fr = fL[::-1]
    bc_hi = np.stack([np.array(br["c_I"], dtype=LD) @ fr[:4],
                      np.array(br["c_II"], dtype=LD) @ fr[:6]])
    rhs[0] -= Lb @ bc_lo
    rhs[-1] -= Ub @ bc_hi

    def inv2(a):
        det = a[0, 0] * a[1, 1] - a[0, 1] * a[1, 0]
        return np.array([[a[1, 1], -a[0, 1]], [-a[1, 0], a[0, 0]]], dtype=LD) / det

    D = [np.eye(2, dtype=LD) for _ in range(n_int)]
    D[0] = D[0] + Lb @ Ml
    D[-1] = D[-1] + Ub @ Mr
    Dp, y = [D[0]], np.empty_like(rhs)
    y[0] = rhs[0]
    for i in range(1, n_int):
        Lm = Lb @ inv2(Dp[i - 1])
        Dp.append(D[i] - Lm @ Ub)
        y[i] = rhs[i] - Lm @ y[i - 1]
    x = np.empty_like(rhs)
    x[-1] = inv2(Dp[-1]) @ y[-1]
    for i in range(n_int - 2, -1, -1):
        x[i] = inv2(Dp[i]) @ (y[i] - Ub @ x[i + 1])
    out = np.empty((2, N + 1, fL.shape[1]), dtype=LD)
    out[:, 1:-1] = x.transpose(1, 0, 2)
    out[:, 0] = Ml @ out[:, 1] + bc_lo
    out[:, N] = Mr @ out[:, N - 1] + bc_hi
    out = out.astype(np.float64)
    return _from_lines(out[0], shape, axis), _from_lines(out[1], shape, axis)


# --------------------------------------------------------------------------
# Periodic axis
# --------------------------------------------------------------------------

def build_periodic_axis(n_pts: int, h: float) -> dict:
    """Block-circulant 2N x 2N system over unique nodes 0..N-1
    (ccd_solver_helpers.py:406-423)."""
    N = n_pts - 1
    if N < 3:
        raise ValueError(f"Need >= 3 unique nodes for periodic CCD; got {N}")
    Lb, Ub = lower_block(h), upper_block(h)
    A = np.zeros((2 * N, 2 * N))
    for i in range(N):
        A[2 * i:2 * i + 2, 2 * i:2 * i + 2] += np.eye(2)
        j = (i - 1) % N
        A[2 * i:2 * i + 2, 2 * j:2 * j + 2] += Lb
        j = (i + 1) % N
        A[2 * i:2 * i + 2, 2 * j:2 * j + 2] += Ub
    lu, piv = sla.lu_factor(A)
    return {"A": A, "lu": lu, "piv": piv, "h": h, "N": N}


def differentiate_periodic(data, axis: int, info: dict):
    """Periodic CCD along ``axis`` (ccd_solver_helpers.py:426-465).

    The input image node N is ignored; the output image node N is a copy of
    node 0.
    """
    h, N = info["h"], info["N"]
    f, shape = _as_lines(data, axis)
    fu = f[:N]
    f_m = np.roll(fu, 1, axis=0)
    f_p = np.roll(fu, -1, axis=0)
    rhs = np.empty((2 * N, f.shape[1]))
    rhs[0::2] = (A1 / h) * (f_p - f_m)
    rhs[1::2] = (A2 / (h * h)) * (f_m - 2.0 * fu + f_p)
    x = sla.lu_solve((info["lu"], info["piv"]), rhs)
    d1 = np.empty((N + 1, f.shape[1]))
    d2 = np.empty((N + 1, f.shape[1]))
    d1[:N], d2[:N] = x[0::2], x[1::2]
    d1[N], d2[N] = x[0], x[1]
    return _from_lines(d1, shape, axis), _from_lines(d2, shape, axis)


def periodic_symbol(theta, h: float):
    """Exact discrete CCD response to the Fourier mode exp(i*theta*j).

    Derived from the circulant blocks (ccd_solver_helpers.py:235-240,
    410-419): with c=cos(theta), s=sin(theta),
        Delta = (1+2 a c)(1+2 b c) + 4 B1 B2 s^2      (a=ALPHA1, b=BETA2)
        D1    = (2 i s / h) [A1 (1+2 b c) + 2 A2 B1 (c-1)] / Delta
        D2    = [2 A2 (c-1)(1+2 a c) - 4 A1 B2 s^2] / (h^2 Delta)
    c-1 is evaluated as -2 sin^2(theta/2) to avoid cancellation.
    Returns (D1, D2) with D1 complex.
    """
    theta = np.asarray(theta, dtype=np.float64)
    c, s = np.cos(theta), np.sin(theta)
    cm1 = -2.0 * np.sin(0.5 * theta) ** 2
    delta = (1 + 2 * ALPHA1 * c) * (1 + 2 * BETA2 * c) + 4 * B1 * B2 * s * s
    d1 = 2j * s / h * (A1 * (1 + 2 * BETA2 * c) + 2 * A2 * B1 * cm1) / delta
    d2 = (2 * A2 * cm1 * (1 + 2 * ALPHA1 * c) - 4 * A1 * B2 * s * s) / (h * h * delta)
    return d1, d2


def periodic_spectral(data, axis: int, h: float):
    """Roundoff-free periodic CCD output via the closed-form symbol.

    Applies ``periodic_symbol`` per Fourier mode of the N unique nodes; there
    is no second-difference cancellation, so it is the accuracy yardstick for
    both the dense-LU restatement and the GPU kernels.
    """
    f, shape = _as_lines(data, axis)
    N = f.shape[0] - 1
    F = np.fft.fft(f[:N], axis=0)
    theta = 2.0 * np.pi * np.fft.fftfreq(N)
    s1, s2 = periodic_symbol(theta, h)
    if N % 2 == 0:
        s1 = s1.copy()
        s1[N // 2] = 0.0  # sin(pi)=0 exactly; keeps the output real
    d1 = np.fft.ifft(F * s1[:, None], axis=0).real
    d2 = np.fft.ifft(F * s2[:, None], axis=0).real
    d1 = np.concatenate([d1, d1[:1]], axis=0)
    d2 = np.concatenate([d2, d2[:1]], axis=0)
    return _from_lines(d1, shape, axis), _from_lines(d2, shape, axis)


# --------------------------------------------------------------------------
# Block-Thomas (legacy O(n) template, block_tridiag.py)
# --------------------------------------------------------------------------

def block_thomas_factor(diag, lower, upper):
    """Block LU of a 2x2-block tridiagonal matrix (block_tridiag.py:43-104).

    Raises ``np.linalg.LinAlgError`` on a singular pivot (block_tridiag.py:
    167-174); no pseudo-inverse fallback.
    """
    n = len(diag)
    LU = [None] * n
    Ls = [np.zeros((2, 2))] * n
    LU[0] = np.array(diag[0], dtype=float)
    _require_pivot(LU[0], 0)
    for i in range(1, n):
        try:
            Linv = np.linalg.solve(LU[i - 1].T, np.asarray(lower[i]).T).T
        except np.linalg.LinAlgError as exc:
            raise np.linalg.LinAlgError(
                f"CCD block-tridiagonal factorization: singular pivot at block "
                f"{i - 1}; pseudo-inverse fallback is not permitted") from exc
        Ls[i] = Linv
        LU[i] = np.asarray(diag[i]) - Linv @ np.asarray(upper[i - 1])
        _require_pivot(LU[i], i)
    return {"LU": np.stack(LU), "L": np.stack(Ls),
            "U": np.stack([np.asarray(u, dtype=float) for u in upper])}


def _require_pivot(A, index):
    det = float(A[0, 0] * A[1, 1] - A[0, 1] * A[1, 0])
    if not math.isfinite(det) or abs(det) < 1e-300:
        raise np.linalg.LinAlgError(
            f"CCD block-tridiagonal factorization: singular pivot at block "
            f"{index}; pseudo-inverse fallback is not permitted")


def block_thomas_solve(fac, rhs):
    """Forward/back substitution, rhs shape (n, 2, batch) (block_tridiag.py:108-164)."""
    n = fac["LU"].shape[0]
    y = np.empty_like(rhs)
    y[0] = rhs[0]
    for i in range(1, n):
        y[i] = rhs[i] - fac["L"][i] @ y[i - 1]
    x = np.empty_like(rhs)
    x[n - 1] = np.linalg.solve(fac["LU"][n - 1], y[n - 1])
    for i in range(n - 2, -1, -1):
        x[i] = np.linalg.solve(fac["LU"][i], y[i] - fac["U"][i] @ x[i + 1])
    return x


# --------------------------------------------------------------------------
# Facade with the reference CCDSolver surface (ccd_solver.py:79-281)
# --------------------------------------------------------------------------

def axis_kinds(bc_type, ndim: int):
    """Per-axis 'wall'/'periodic' (core/boundary.py:45-76)."""
    if isinstance(bc_type, (tuple, list)):
        return tuple(str(a).lower() for a in bc_type)
    v = str(bc_type).lower()
    if v == "wall":
        return ("wall",) * ndim
    if v == "periodic":
        return ("periodic",) * ndim
    if v == "periodic_wall" and ndim == 2:
        return ("periodic", "wall")
    if v == "wall_periodic" and ndim == 2:
        return ("wall", "periodic")
    raise ValueError(f"unsupported boundary type {bc_type!r} for ndim={ndim}")


def metric_from_coords(coords):
    """J = dxi/dx and dJ/dxi of one axis of a non-uniform grid from its node
    coordinates (core/metrics.py:55-60): x(xi) differentiated by the wall CCD
    solve with the computational spacing 1/N (ccd_solver.py:174-203), then
    J = 1 / x_xi and dJ/dxi = -x_xixi / x_xi^2."""
    x = np.asarray(coords, dtype=np.float64).reshape(-1)
    N = x.shape[0] - 1
    d1, d2 = differentiate_wall(x, 0, build_wall_axis(N + 1, 1.0 / N))
    return 1.0 / d1, -d2 / (d1 ** 2)


def apply_ccd_metric(d1_xi, d2_xi, J, dJ, axis: int):
    """ccd_solver_helpers.py:509-519: d1_x = J d1_xi and
    d2_x = J J d2_xi + J dJ d1_xi, J and dJ broadcast along ``axis``."""
    shape = [1] * np.ndim(d1_xi)
    shape[axis] = -1
    J = np.asarray(J).reshape(shape)
    dJ = np.asarray(dJ).reshape(shape)
    return J * d1_xi, J * J * d2_xi + J * dJ * d1_xi


class OracleCCD:
    """CPU CCD with the reference ``CCDSolver`` call surface.  With
    ``coords`` (one coordinate vector or None per axis) the grid is
    non-uniform as the reference's ``alpha_grid > 1``: every axis is solved
    in xi with spacing 1/N and mapped by its metric (axes given None keep
    the uniform coordinates, whose metric is 1/L)."""

    def __init__(self, N, L, bc_type="wall", coords=None):
        self.N = tuple(int(n) for n in N)
        self.L = tuple(float(x) for x in L)
        self.ndim = len(self.N)
        self.bc_type = bc_type
        self.kinds = axis_kinds(bc_type, self.ndim)
        self._info = {}
        self.uniform = coords is None
        if not self.uniform:
            self.coords = [np.linspace(0.0, self.L[a], self.N[a] + 1)
                           if coords[a] is None else np.asarray(coords[a], dtype=np.float64)
                           for a in range(self.ndim)]
            m = [metric_from_coords(c) for c in self.coords]
            self.J = [t[0] for t in m]
            self.dJ_dxi = [t[1] for t in m]

    def h(self, axis):
        return self.L[axis] / self.N[axis] if self.uniform else 1.0 / self.N[axis]

    def _axis_info(self, axis):
        if axis not in self._info:
            n_pts = self.N[axis] + 1
            if self.kinds[axis] == "periodic":
                self._info[axis] = build_periodic_axis(n_pts, self.h(axis))
            else:
                self._info[axis] = build_wall_axis(n_pts, self.h(axis))
        return self._info[axis]

    def differentiate(self, data, axis, apply_metric=True):
        info = self._axis_info(axis)
        if self.kinds[axis] == "periodic":
            d1, d2 = differentiate_periodic(data, axis, info)
        else:
            d1, d2 = differentiate_wall(data, axis, info)
        if not self.uniform and apply_metric:
            d1, d2 = apply_ccd_metric(d1, d2, self.J[axis], self.dJ_dxi[axis], axis)
        return d1, d2

    def first_derivative(self, data, axis, apply_metric=True):
        return self.differentiate(data, axis, apply_metric)[0]

    def second_derivative(self, data, axis, apply_metric=True):
        return self.differentiate(data, axis, apply_metric)[1]

    def enforce_wall_neumann(self, grad, ax):
        """ccd_boundary_helpers.py:52-61."""
        if self.kinds[ax] != "wall":
            return
        lo = [slice(None)] * grad.ndim
        hi = [slice(None)] * grad.ndim
        lo[ax], hi[ax] = 0, -1
        grad[tuple(lo)] = 0.0
        grad[tuple(hi)] = 0.0
