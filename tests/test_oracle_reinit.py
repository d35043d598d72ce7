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


@pytest.mark.parametrize("k", [1, 5, 11])
def test_compression_filter_transfer_factor(k):
    """Periodic cosine flux: the filtered divergence is the CCD derivative
    (pinned in test_oracle_ccd.py) times 1 - 4 eps_d sin^2(theta/2)."""
    N, eps_d = 48, 0.05
    o = OracleSplitReinit((N,), (1.0,), "periodic", 0.03, bc="periodic", eps_d_comp=eps_d)
    th = 2.0 * np.pi * k / N
    flux = np.cos(th * np.arange(N + 1) + 0.3)
    got = o.filtered_divergence(flux, 0)
    want = (1.0 - 4.0 * eps_d * np.sin(th / 2) ** 2) * o.ccd.first_derivative(flux, 0)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("bc,edge", [("zero", 1.0 - 0.05), ("neumann", 1.0), ("outflow", 1.0)])
def test_compression_filter_linear_flux_and_ghosts(bc, edge):
    """Linear flux 0.3 + 2x on a wall axis: CCD is exact (f' = 2), the filter
    leaves interior nodes at 2; at the two wall nodes the ghost decides: zero
    ghost gives 2 (1 - eps_d), mirrored / repeated ghosts give 2."""
    N = 20
    o = OracleSplitReinit((N, 6), (1.0, 1.0), "wall", 0.03, bc=bc, eps_d_comp=0.05)
    x = np.linspace(0.0, 1.0, N + 1)
    flux = np.broadcast_to((0.3 + 2.0 * x)[:, None], (N + 1, 7)).copy()
    got = o.filtered_divergence(flux, 0)
    assert np.abs(got[1:-1] - 2.0).max() < 1e-11
    assert np.abs(got[0] - 2.0 * edge).max() < 1e-11
    assert np.abs(got[-1] - 2.0 * edge).max() < 1e-11


def test_compression_div_on_monotone_tanh():
    """psi = 1/(1 + exp(-(x - 1/2)/w)) is increasing, so n = +1 and the
    compression divergence is d/dx[psi (1 - psi)] = (1 - 2 psi) psi'
    = (1 - 2 psi) psi (1 - psi) / w; the filter moves it by about
    eps_d h^2 g''' (2e-3 of its maximum at w = 8h; the CCD error is far
    smaller), so 1e-2 separates it from any change of the assembly's scale."""
    N, w = 256, 8.0 / 256
    x = np.linspace(0.0, 1.0, N + 1)
    psi = 1.0 / (1.0 + np.exp(-(x - 0.5) / w))
    o = OracleSplitReinit((N,), (1.0,), "wall", 0.03, bc="neumann", eps_d_comp=0.05)
    got = o.compression_div(psi)
    want = (1.0 - 2.0 * psi) * psi * (1.0 - psi) / w
    assert np.abs(got - want).max() < 1e-2 * np.abs(want).max()


class _SignFlipped(OracleSplitReinit):
    def filtered_divergence(self, flux, a):
        g = self.ccd.first_derivative(flux, a)
        return 2.0 * g - super().filtered_divergence(flux, a)


class _FilterDropped(OracleSplitReinit):
    def filtered_divergence(self, flux, a):
        return self.ccd.first_derivative(flux, a)


class _Halved(OracleSplitReinit):
    def compression_div(self, q):
        return 0.5 * super().compression_div(q)


@pytest.mark.parametrize("planted", [_SignFlipped, _FilterDropped, _Halved])
def test_planted_compression_mistakes_fail_a_pin(planted, monkeypatch):
    """Each planted mistake in the compression step (filter sign flipped,
    filter dropped, compression halved) fails at least one of the pins
    above, run against the planted oracle."""
    import sys
    mod = sys.modules[__name__]
    monkeypatch.setattr(mod, "OracleSplitReinit", planted)
    checks = [lambda: test_compression_filter_transfer_factor(5),
              lambda: test_compression_filter_linear_flux_and_ghosts("zero", 1.0 - 0.05),
              test_compression_div_on_monotone_tanh]
    failed = 0
    for check in checks:
        try:
            check()
        except AssertionError:
            failed += 1
    assert failed >= 1
