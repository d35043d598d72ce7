"""Pins of the split-reinitialisation oracle (oracle/reinit.py) by closed forms
(levelset/reinit_split.py, levelset/reinit_ops.py)."""

import numpy as np
import pytest

from oracle.reinit import BETA2, OracleSplitReinit, cn_factors, cn_sweep, compute_dtau


@pytest.mark.parametrize("k", [1, 3, 7])
def test_cn_sweep_on_cosine_modes(k):
    N, mu = 32, 0.37
    th = np.pi * k / N
    i = np.arange(N + 1)
    psi = 0.5 + 0.2 * np.cos(th * i)
    out = cn_sweep(psi, 0, mu, cn_factors(N + 1, mu))
    c = np.cos(th)
    G = (1 + 2 * BETA2 * c - 6 * mu * (1 - c)) / (1 + 2 * BETA2 * c + 6 * mu * (1 - c))
    assert np.abs(out - (0.5 + 0.2 * G * np.cos(th * i))).max() < 1e-14


def test_dtau():
    assert compute_dtau((64, 32), (1.0, 1.0), 0.02) == min(0.5 / 64 ** 2 / (4 * 0.02), 0.5 / 64)
    assert compute_dtau((16,), (1.0,), 1.0) == 0.5 / 16 ** 2 / 2.0


def test_tanh_profile_is_nearly_fixed_and_mass_is_restored():
    N, eps = 128, 0.02
    x = np.linspace(0.0, 1.0, N + 1)
    psi = 1.0 / (1.0 + np.exp(-(x - 0.5) / eps))
    o = OracleSplitReinit((N,), (1.0,), "wall", eps, n_steps=4, bc="zero")
    q = o.reinitialize(psi)
    dV = o.dV()
    assert abs(np.sum(q * dV) - np.sum(psi * dV)) < 1e-14
    assert np.abs(q - psi).max() < 5e-2          # near the equilibrium (2.6 cells per eps)
    bumpy = np.clip(psi + 0.05 * np.sin(40 * np.pi * x), 0, 1)
    assert np.abs(o.reinitialize(bumpy) - psi).max() < np.abs(bumpy - psi).max()
