"""Pins of the Rhie-Chow oracle (oracle/rhie_chow.py) by closed forms
(spatial/rhie_chow.py:55-284)."""

import numpy as np
import pytest

from oracle.rhie_chow import OracleRhieChow


def test_linear_velocity_no_pressure():
    N, L = 16, 2.0
    h = L / N
    o = OracleRhieChow((N,), (L,), "wall")
    x = np.linspace(0, L, N + 1)
    u = 3.0 * x + 0.5
    d = o.face_velocity_divergence([u], np.zeros(N + 1), np.ones(N + 1), 0.1)
    faces = 3.0 * 0.5 * (x[1:] + x[:-1]) + 0.5            # faces 1..N
    F = np.concatenate([[0.0], faces])
    assert np.allclose(d[:N], (F[1:] - F[:-1]) / h, rtol=1e-14, atol=1e-12)
    assert abs(d[N] + F[N] / h) < 1e-12
    assert np.allclose(d[1:N], 3.0, atol=1e-12)


def test_cubic_pressure_bracket():
    """p = x^3, u = 0, rho = 2: inside, dp_face - dp_bar = -h^2/2 exactly
    (CCD is exact for cubics); face 1 uses the one-sided node-0 gradient
    (bracket -h^2), face N the one-sided node-N gradient ((3 x h + h^2)/2,
    x = x_{N-1})."""
    N, L, dt, rho = 20, 1.0, 0.05, 2.0
    h = L / N
    o = OracleRhieChow((N,), (L,), "wall")
    x = np.linspace(0, L, N + 1)
    d = o.face_velocity_divergence([np.zeros(N + 1)], x ** 3, np.full(N + 1, rho), dt)
    br = np.full(N + 1, -0.5 * h * h)
    br[0] = 0.0
    br[1] = -h * h
    xm = x[N - 1]
    br[N] = 0.5 * (3 * xm * h + h * h)
    F = -dt / rho * br
    F[0] = 0.0
    want = np.empty(N + 1)
    want[:N] = (F[1:] - F[:-1]) / h
    want[N] = -F[N] / h
    assert np.abs(d - want).max() < 1e-10 * np.abs(want).max()


@pytest.mark.parametrize("bc", ["wall", "periodic"])
def test_balanced_force_equilibrium(bc):
    """p = kappa psi / We, constant kappa: the balanced-force bracket cancels
    the pressure bracket, leaving the p = 0 divergence."""
    N = (24, 16)
    o = OracleRhieChow(N, (1.0, 1.0), bc)
    rng = np.random.default_rng(3)
    x = np.linspace(0, 1, 25)[:, None]
    y = np.linspace(0, 1, 17)[None, :]
    psi = 0.5 + 0.4 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
    kappa = np.full(psi.shape, 3.0)
    we = 2.5
    p = kappa * psi / we
    vel = [rng.standard_normal(psi.shape), rng.standard_normal(psi.shape)]
    rho = 1.0 + rng.random(psi.shape)
    d_bf = o.face_velocity_divergence(vel, p, rho, 0.1, kappa=kappa, psi=psi, we=we)
    d_0 = o.face_velocity_divergence(vel, np.zeros_like(p), rho, 0.1)
    assert np.abs(d_bf - d_0).max() < 1e-11 * np.abs(d_0).max()
    d_nobf = o.face_velocity_divergence(vel, p, rho, 0.1)
    assert np.abs(d_nobf - d_0).max() > 1e-6  # the bracket itself is not zero
