"""Pins of the FCCD PPE oracle (oracle/fccd.py) against what the mathematics
fixes, plus the configuration errors of the FCCD PPE plugin (raised before any
device work).

* periodic, constant density: the operator equals the closed-form symbol
  (2 cos t - 2)/h^2 (1 - h^2/24 S2(t)) / rho per Fourier mode, S2 the CCD f''
  symbol -- a dropped H/24 term, a wrong sign or a wrong face pairing fails;
* wall: the FCCD face gradient is exact for cubics at face midpoints, so the
  interior rows equal (c_hi p'(x_hi) - c_lo p'(x_lo)) / h exactly;
* wall: discrete divergence theorem, sum_i w_i (L p)_i = 0 per line (full
  face space: the boundary rows carry +-F/w);
* gauge rows: out[pin] = p[pin], image rows p_img - p_src;
* GMRES solve (SciPy) equals a dense direct solve of the assembled operator.
"""

import numpy as np
import pytest

from oracle.fccd import OracleFCCDPPE, fccd_symbol


def _mode(shape, N, ks):
    idx = np.indices(shape)
    ph = sum(2 * np.pi * k * idx[a] / N[a] for a, k in enumerate(ks))
    return np.cos(ph)


@pytest.mark.parametrize("N,ks", [((32,), (5,)), ((32, 16), (3, 5)), ((16, 12, 8), (2, 3, 1)),
                                  ((24, 16), (11, 7))])
def test_periodic_constant_density_matches_symbol(N, ks):
    o = OracleFCCDPPE(N, (1.0,) * len(N), "periodic")
    rho = 3.0
    o.prepare_operator(np.full(o.shape, rho))
    p = _mode(o.shape, N, ks)
    out = o.core(p)
    lam = sum(fccd_symbol(2 * np.pi * k / n, 1.0 / n) for k, n in zip(ks, N)) / rho
    inner = tuple(slice(0, n) for n in N)
    err = np.abs(out[inner] - lam * p[inner]).max() / abs(lam)
    assert err < 1e-12


def test_symbol_check_catches_a_planted_sign_error(monkeypatch):
    """The symbol test is not vacuous: flipping the H/24 term fails it."""
    N = (32, 16)
    o = OracleFCCDPPE(N, (1.0, 1.0), "periodic")
    o.prepare_operator(np.ones(o.shape))
    good = o.face_gradient

    def bad(p, axis, q=None):
        q = o.ccd.second_derivative(p, axis) if q is None else q
        return good(p, axis, q=-q)

    monkeypatch.setattr(o, "face_gradient", bad)
    p = _mode(o.shape, N, (3, 5))
    out = o.core(p)
    lam = fccd_symbol(2 * np.pi * 3 / 32, 1 / 32) + fccd_symbol(2 * np.pi * 5 / 16, 1 / 16)
    assert np.abs(out[:32, :16] - lam * p[:32, :16]).max() / abs(lam) > 1e-4


def test_wall_face_gradient_exact_for_cubics():
    o = OracleFCCDPPE((20,), (2.0,), "wall")
    x = np.linspace(0.0, 2.0, 21)
    p = 0.3 * x ** 3 - x ** 2 + 2.0 * x - 1.0
    g = o.face_gradient(p, 0)
    xm = 0.5 * (x[1:] + x[:-1])
    assert np.abs(g - (0.9 * xm ** 2 - 2.0 * xm + 2.0)).max() < 1e-12


def test_wall_interior_rows_exact_for_cubics_variable_density():
    rng = np.random.default_rng(3)
    N = (16, 12)
    o = OracleFCCDPPE(N, (1.0, 1.5), "wall")
    rho = 1.0 + 9.0 * rng.random(o.shape)
    o.prepare_operator(rho)
    x = np.linspace(0.0, 1.0, 17)[:, None]
    y = np.linspace(0.0, 1.5, 13)[None, :]
    p = x ** 3 - 2.0 * x * y ** 2 + y ** 3 + 0.5 * x
    out = o.core(p)
    hx, hy = 1.0 / 16, 1.5 / 12
    dpx = lambda xx, yy: 3 * xx ** 2 - 2 * yy ** 2 + 0.5
    dpy = lambda xx, yy: -4 * xx * yy + 3 * yy ** 2
    cx = 2.0 / (rho[:-1, :] + rho[1:, :])
    cy = 2.0 / (rho[:, :-1] + rho[:, 1:])
    Fx = cx * dpx(x[:-1] + hx / 2, y)
    Fy = cy * dpy(x, y[:, :-1] + hy / 2)
    want = (Fx[1:, 1:-1] - Fx[:-1, 1:-1]) / hx + (Fy[1:-1, 1:] - Fy[1:-1, :-1]) / hy
    got = out[1:-1, 1:-1]
    assert np.abs(got - want).max() / np.abs(want).max() < 1e-11


def test_wall_discrete_divergence_theorem():
    rng = np.random.default_rng(4)
    o = OracleFCCDPPE((12, 10, 8), (1.0, 2.0, 0.5), "wall")
    o.prepare_operator(1.0 + 100.0 * rng.random(o.shape))
    p = rng.standard_normal(o.shape)
    out = o.core(p)
    w = [o.node_width(a) for a in range(3)]
    tot = np.einsum("ijk,i,j,k->", out, w[0], w[1], w[2])
    scale = np.einsum("ijk,i,j,k->", np.abs(out), w[0], w[1], w[2])
    assert abs(tot) < 1e-13 * scale


def test_gauge_rows():
    rng = np.random.default_rng(5)
    o = OracleFCCDPPE((8, 6), (1.0, 1.0), "periodic")
    o.prepare_operator(1.0 + rng.random(o.shape))
    p = rng.standard_normal(o.shape)
    out = o.apply(p)
    assert out.ravel()[o.pin] == p.ravel()[o.pin]
    assert o.images is not None and len(o.images) == 7 + 9 - 1  # nodes with x = 8 or y = 6
    np.testing.assert_array_equal(out.ravel()[o.images],
                                  p.ravel()[o.images] - p.ravel()[o.sources])
    # corner (8, 6): first periodic axis (x) decides its source (0, 6)
    corner = np.ravel_multi_index((8, 6), o.shape)
    k = int(np.where(o.images == corner)[0][0])
    assert o.sources[k] == np.ravel_multi_index((0, 6), o.shape)


@pytest.mark.parametrize("bc,pre", [("wall", "jacobi"), ("wall", "none"), ("periodic", "none")])
def test_gmres_solve_matches_dense_direct(bc, pre):
    rng = np.random.default_rng(6)
    o = OracleFCCDPPE((8, 6), (1.0, 1.0), bc, pre)
    rho = 1.0 + 20.0 * rng.random(o.shape)
    if bc == "periodic":
        rho[-1, :], rho[:, -1] = rho[0, :], rho[:, 0]
    rhs = rng.standard_normal(o.shape)
    p, info = o.solve(rhs, rho, tol=1e-12, maxiter=3000)
    assert info["info"] == 0 and info["relres"] < 1e-11
    A = o.matrix()
    x = np.linalg.solve(A, o.system_rhs(rhs).ravel()).reshape(o.shape)
    assert np.abs(p - x).max() <= 1e-9 * np.abs(x).max()
    assert p.ravel()[o.pin] == 0.0


# --- plugin configuration errors (raised before any device work) ------------
class _Cfg:
    def __init__(self, **kw):
        for k, v in kw.items():
            setattr(self, k, v)


class _Grid:
    ndim, N, L, uniform = 2, (8, 8), (1.0, 1.0), True


@pytest.mark.parametrize("kw,exc", [
    ({"ppe_preconditioner": "line_pcr"}, ValueError),
    ({"ppe_coefficient_scheme": "bogus"}, ValueError),
    ({"ppe_interface_coupling_scheme": "affine_jump"}, ValueError),
    ({"ppe_interface_coupling_scheme": "bogus"}, ValueError),
    ({"ppe_coefficient_scheme": "phase_density",
      "ppe_interface_coupling_scheme": "jump_decomposition"}, ValueError),
    ({"scalar_operator_pairing": "variational"}, NotImplementedError),
    ({"boundary_face_space": "impermeable_face"}, NotImplementedError),
    ({"scalar_operator_pairing": "require_certified",
      "pressure_force_contract": "variational_adjoint"}, RuntimeError),
])
def test_fccd_ppe_config_errors(kw, exc):
    from twophaseflow_amd import PPESolverFCCDMatrixFree
    with pytest.raises(exc):
        PPESolverFCCDMatrixFree(None, _Cfg(**kw), _Grid(), None, bc_type="wall")


# --- FCCDSolver face primitives (oracle.fccd.OracleFCCD) ---------------------
from oracle.fccd import OracleFCCD, fccd_face_symbol_truncated  # noqa: E402


def _cubic(x):
    return 0.7 * x ** 3 - 1.3 * x ** 2 + 0.4 * x + 2.0


def _dcubic(x):
    return 2.1 * x ** 2 - 2.6 * x + 0.4


def test_face_value_curvature_node_gradient_exact_for_cubics():
    o = OracleFCCD((24,), (1.5,), "wall")
    x = np.linspace(0.0, 1.5, 25)
    xm = 0.5 * (x[1:] + x[:-1])
    u = _cubic(x)
    assert np.abs(o.face_value(u, 0) - _cubic(xm)).max() < 1e-12
    assert np.abs(o.face_second_derivative(u, 0) - (4.2 * xm - 2.6)).max() < 1e-11
    g = o.node_gradient(u, 0)
    assert np.abs(g[1:-1] - _dcubic(x[1:-1])).max() < 1e-11
    # wall nodes: one-sided, the first / last face gradients (exact at the faces)
    assert abs(g[0] - _dcubic(xm[0])) < 1e-11 and abs(g[-1] - _dcubic(xm[-1])) < 1e-11


def test_upwind_face_value_exact_for_quadratics():
    o = OracleFCCD((16,), (1.0,), "wall")
    x = np.linspace(0.0, 1.0, 17)
    xm = 0.5 * (x[1:] + x[:-1])
    u = 3.0 * x ** 2 - x + 0.5
    vel = np.where(np.arange(16) % 2 == 0, 1.0, -1.0)
    got = o.upwind_face_value(u, vel, 0)
    assert np.abs(got - (3.0 * xm ** 2 - xm + 0.5)).max() < 1e-12


def test_face_divergence_linear_faces_and_periodic_closed_form():
    o = OracleFCCD((20, 8), (1.0, 1.0), "wall")
    xm = (np.arange(20) + 0.5) / 20
    F = np.repeat((2.0 * xm + 1.0)[:, None], 9, axis=1)
    div = o.face_divergence(F, 0)
    assert np.abs(div[1:-1] - 2.0).max() < 1e-12 and np.all(div[[0, -1]] == 0.0)
    op = OracleFCCD((32,), (1.0,), "periodic")
    th = 2 * np.pi * 3 / 32
    F = np.sin(th * (np.arange(32) + 0.5))
    div = op.face_divergence(F, 0)
    j = np.arange(33) % 32
    want = 2.0 * np.cos(th * j) * np.sin(th / 2) * 32
    assert np.abs(div - want).max() < 1e-11


@pytest.mark.parametrize("k", [1, 2, 4])
def test_periodic_face_gradient_meets_the_truncated_symbol(k):
    N, H = 64, 1.0 / 64
    o = OracleFCCD((N,), (1.0,), "periodic")
    w = 2 * np.pi * k
    x = np.arange(N + 1) * H
    g = o.face_gradient(np.cos(w * x), 0) + 1j * o.face_gradient(np.sin(w * x), 0)
    sym = g / np.exp(1j * w * (np.arange(N) + 0.5) * H)
    err = np.abs(sym - fccd_face_symbol_truncated(w, H)).max() / w
    assert err <= (w * H) ** 6
