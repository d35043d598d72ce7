"""Pins of the DCCD advection oracle (oracle/dccd.py) by closed forms
(levelset/advection_dccd.py:84-152, time_integration/tvd_rk3.py:26-48)."""

import numpy as np

from oracle.ccd import periodic_symbol
from oracle.dccd import OracleDCCD


def test_tvd_rk3_amplification_of_a_mode():
    N, u, dt, k = 64, 0.7, 0.004, 5
    h = 1.0 / N
    o = OracleDCCD((N,), (1.0,), "periodic", "periodic", eps_d=0.0)
    th = 2 * np.pi * k / N
    j = np.arange(N + 1)
    q0 = 0.5 + 0.1 * np.cos(th * j)
    q = o.advance(q0, [np.full(N + 1, u)], dt, clip_bounds=None)
    s1, _ = periodic_symbol(th, h)
    z = -u * s1 * dt
    G = 1 + z + z * z / 2 + z ** 3 / 6
    want = 0.5 + 0.1 * np.real(G * np.exp(1j * th * j))
    assert np.abs(q - want).max() < 1e-13


def test_filter_transfer_with_constant_psi():
    N, k, eps, c = 48, 7, 0.05, 0.8
    o = OracleDCCD((N,), (1.0,), "periodic", "periodic", eps_d=eps)
    th = 2 * np.pi * k / N
    j = np.arange(N + 1)
    vel = np.cos(th * j)
    r = o.rhs(np.full(N + 1, c), [vel])
    s1, _ = periodic_symbol(th, 1.0 / N)
    fac = 1.0 - 4.0 * eps * (2 * c - 1) ** 2 * np.sin(th / 2) ** 2
    want = -np.real(fac * c * s1 * np.exp(1j * th * j))
    assert np.abs(r - want).max() < 1e-11 * np.abs(want).max()


def test_wall_end_nodes_unfiltered():
    N = 20
    rng = np.random.default_rng(1)
    o = OracleDCCD((N,), (1.0,), "wall", "neumann", eps_d=0.05)
    o0 = OracleDCCD((N,), (1.0,), "wall", "neumann", eps_d=0.0)
    psi, vel = rng.random(N + 1), rng.standard_normal(N + 1)
    r, r0 = o.rhs(psi, [vel]), o0.rhs(psi, [vel])
    assert r[0] == r0[0] and r[-1] == r0[-1]
    assert np.abs(r[1:-1] - r0[1:-1]).max() > 1e-6


def test_mass_correction_restores_mass():
    N = (32, 24)
    x = np.linspace(0, 1, 33)[:, None]
    y = np.linspace(0, 1, 25)[None, :]
    psi = 1.0 / (1.0 + np.exp(np.clip((np.sqrt((x - .5) ** 2 + (y - .5) ** 2) - .25) / 0.05,
                                      -500, 500)))
    vel = [np.full(psi.shape, 0.3), np.full(psi.shape, -0.2)]
    o = OracleDCCD(N, (1.0, 1.0), "periodic", "periodic", mass_correction=True)
    q = o.advance(psi, vel, 0.01)
    dV = o.dV()
    assert abs(np.sum(q * dV) - np.sum(o.sync(psi.copy()) * dV)) < 1e-13
