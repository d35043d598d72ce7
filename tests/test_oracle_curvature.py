"""Pins of the psi-curvature oracle (oracle/curvature.py) by closed forms
(levelset/curvature_psi.py:79-168, levelset/curvature.py:60-128)."""

import numpy as np
import pytest

from oracle.curvature import OracleCurvaturePsi, dccd_filter_nd


def _axes(N, L):
    return [np.linspace(0, l, n + 1) for n, l in zip(N, L)]


def test_circle_level_sets_2d():
    """psi = 0.2 + 0.3 ((x-a)^2 + (y-b)^2): level sets are circles, CCD exact
    for quadratics, so kappa = -1/r at every node inside the mask."""
    N, L = (24, 20), (1.0, 1.2)
    x, y = np.meshgrid(*_axes(N, L), indexing="ij")
    a, b = -0.3, 0.25
    r2 = (x - a) ** 2 + (y - b) ** 2
    psi = 0.2 + 0.3 * r2
    k = OracleCurvaturePsi(N, L, "wall", dccd_eps=0.0).compute(psi)
    m = (psi > 0.01) & (psi < 0.99)
    assert m.sum() > 100
    assert np.abs(k[m] + 1.0 / np.sqrt(r2[m])).max() < 1e-9


def test_sphere_level_sets_3d():
    N, L = (12, 10, 8), (1.0, 1.0, 1.0)
    x, y, z = np.meshgrid(*_axes(N, L), indexing="ij")
    r2 = (x + 0.4) ** 2 + (y - 0.3) ** 2 + (z + 0.2) ** 2
    psi = 0.1 + 0.25 * r2
    k = OracleCurvaturePsi(N, L, "wall", dccd_eps=0.0).compute(psi)
    m = (psi > 0.01) & (psi < 0.99)
    assert m.sum() > 100
    assert np.abs(k[m] + 2.0 / np.sqrt(r2[m])).max() < 1e-9


def test_rotated_ellipse_mixed_term():
    """A rotated ellipse (psi_xy != 0): kappa equals the axis-aligned
    ellipse's closed form -(Q_Y^2 Q_XX + Q_X^2 Q_YY)/|grad Q|^3 at the rotated
    point (curvature is rotation-invariant)."""
    N, L = (24, 24), (1.0, 1.0)
    x, y = np.meshgrid(*_axes(N, L), indexing="ij")
    th = 0.6
    X = np.cos(th) * (x + 0.2) + np.sin(th) * (y + 0.1)
    Y = -np.sin(th) * (x + 0.2) + np.cos(th) * (y + 0.1)
    A, C = 0.5, 0.2
    psi = 0.05 + A * X ** 2 + C * Y ** 2
    k = OracleCurvaturePsi(N, L, "wall", dccd_eps=0.0).compute(psi)
    QX, QY = 2 * A * X, 2 * C * Y
    want = -(QY ** 2 * 2 * A + QX ** 2 * 2 * C) / (QX ** 2 + QY ** 2) ** 1.5
    m = (psi > 0.01) & (psi < 0.99)
    assert m.sum() > 100
    assert np.abs(k[m] - want[m]).max() < 1e-8 * np.abs(want[m]).max()
    wrong = -(QY ** 2 * 2 * A + QX ** 2 * 2 * C) / (QX ** 2 + QY ** 2) ** 1.5 * 0.0
    assert np.abs(k[m] - wrong[m]).max() > 1e-3


@pytest.mark.parametrize("bc", ["wall", "periodic"])
def test_filter_keeps_linear_fields_and_wall_planes(bc):
    N = (16, 12)
    rng = np.random.default_rng(2)
    psi = rng.random((17, 13))
    psi[-1], psi[:, -1] = psi[0], psi[:, 0]  # a periodic field's images
    psi[-1, -1] = psi[0, 0]
    x, y = np.meshgrid(*_axes(N, (1.0, 1.0)), indexing="ij")
    f = 2.0 * x - 3.0 * y + 1.0 if bc == "wall" else np.ones((17, 13)) * 0.7
    kinds = ("wall", "wall") if bc == "wall" else ("periodic", "periodic")
    out = dccd_filter_nd(f, N, kinds, 0.05, psi)
    assert np.abs(out - f).max() < 1e-13
    g = rng.standard_normal((17, 13))
    out = dccd_filter_nd(g, N, kinds, 0.05, psi)
    if bc == "wall":
        assert np.array_equal(out[0], g[0]) and np.array_equal(out[:, -1], g[:, -1])
    else:
        assert np.array_equal(out[-1], out[0]) and np.array_equal(out[:, -1], out[:, 0])


@pytest.mark.parametrize("psi0", [0.8, 0.15, 0.5])
def test_filter_transfer_factor_constant_psi(psi0):
    """Constant psi: the filter coefficient eps (2 psi - 1)^2 is a constant s,
    so a periodic cosine mode cos(t1 i) cos(t2 j) is multiplied by
    (1 - 4 s sin^2(t1/2)) (1 - 4 s sin^2(t2/2)); psi = 0.5 switches it off."""
    N = (32, 24)
    eps = 0.05
    i, j = np.meshgrid(np.arange(33), np.arange(25), indexing="ij")
    t1, t2 = 2 * np.pi * 3 / 32, 2 * np.pi * 5 / 24
    f = np.cos(t1 * i) * np.cos(t2 * j)
    psi = np.full(f.shape, psi0)
    out = dccd_filter_nd(f, N, ("periodic", "periodic"), eps, psi)
    s = eps * (2 * psi0 - 1) ** 2
    G = (1 - 4 * s * np.sin(t1 / 2) ** 2) * (1 - 4 * s * np.sin(t2 / 2) ** 2)
    assert np.abs(out - G * f).max() < 1e-14
    # wall axes: interior nodes of a mirror-symmetric mode carry the same
    # factor, the wall planes keep the input
    k1, k2 = np.pi * 3 / 32, np.pi * 4 / 24
    g = np.cos(k1 * i) * np.cos(k2 * j)
    out = dccd_filter_nd(g, N, ("wall", "wall"), eps, psi)
    Gw = (1 - 4 * s * np.sin(k1 / 2) ** 2) * (1 - 4 * s * np.sin(k2 / 2) ** 2)
    assert np.abs(out[1:-1, 1:-1] - Gw * g[1:-1, 1:-1]).max() < 1e-14
    assert np.array_equal(out[0], g[0]) and np.array_equal(out[:, -1], g[:, -1])


def test_filter_coefficient_is_pointwise():
    """psi varying along axis 0 only and f = x^2 along axis 0: the filter
    adds eps (2 psi_i - 1)^2 * 2 h^2 at interior node i (the axis-1 pass
    sees a field constant along axis 1)."""
    N = (20, 6)
    eps, h = 0.05, 1.0 / 20
    x = np.linspace(0, 1, 21)
    psi = np.broadcast_to((0.5 + 0.45 * np.sin(7 * x))[:, None], (21, 7)).copy()
    f = np.broadcast_to((x ** 2)[:, None], (21, 7)).copy()
    out = dccd_filter_nd(f, N, ("wall", "wall"), eps, psi)
    want = f[1:-1] + eps * (2 * psi[1:-1] - 1) ** 2 * 2 * h * h
    assert np.abs(out[1:-1, 1:-1] - want[:, 1:-1]).max() < 1e-15
    assert np.array_equal(out[:, 0], f[:, 0]) and np.array_equal(out[:, -1], f[:, -1])
