"""Pins for the convolution oracle: direct-loop definition and finite
differences of the two gradients.  CPU only."""

import numpy as np

from oracle import unet as ou


def test_conv3d_matches_direct_loops():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 4, 5, 2))
    w = rng.standard_normal((27, 3, 2))
    np.testing.assert_allclose(ou.conv3d(x, w), ou.conv3d_direct(x, w), rtol=1e-12, atol=1e-12)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((3, 3, 4, 2))
    w = rng.standard_normal((27, 2, 2))
    g = rng.standard_normal((3, 3, 4, 2))
    f = lambda xx, ww: float(np.sum(ou.conv3d(xx, ww) * g))
    gx, gw = ou.conv3d_dgrad(g, w), ou.conv3d_wgrad(x, g)
    eps = 1e-6
    for idx in [(0, 0, 0, 0), (1, 2, 3, 1), (2, 1, 0, 1)]:
        d = np.zeros_like(x); d[idx] = eps
        assert abs((f(x + d, w) - f(x - d, w)) / (2 * eps) - gx[idx]) < 1e-6
    for idx in [(0, 0, 0), (13, 1, 1), (26, 0, 1), (5, 1, 0)]:
        d = np.zeros_like(w); d[idx] = eps
        assert abs((f(x, w + d) - f(x, w - d)) / (2 * eps) - gw[idx]) < 1e-6


def test_planted_flip_is_caught():
    """Using the unflipped shift in dgrad (a plausible bug) breaks the check."""
    rng = np.random.default_rng(2)
    x = rng.standard_normal((3, 3, 3, 1)); w = rng.standard_normal((27, 1, 1))
    g = rng.standard_normal((3, 3, 3, 1))
    bad = np.zeros_like(x)
    for t, (dz, dy, dx) in enumerate(ou.taps()):
        bad += ou._shift(g @ w[t], dz, dy, dx)
    good = ou.conv3d_dgrad(g, w)
    assert np.max(np.abs(bad - good)) > 1e-3


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -3.14159])
    r = ou.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and abs(r[3] + 3.140625) < 1e-12
