"""Oracle pins of the non-uniform CCD metric (core/metrics.py:22-88,
ccd/ccd_solver_helpers.py:509-519) and the host side of non-uniform grids.

The metric is pinned by closed forms that a wrong formula fails: on a
quadratic map x(xi) the CCD solves in xi are exact, so f = x^2 must give
f' = 2x and f'' = 2 to roundoff (a dropped J dJ term, a wrong sign, J in
place of J^2 or a missing 1/N spacing all break it); a linear map gives
J = 1/L and dJ = 0; on a smooth map J converges at the wall closure's order
(the reference's test_grid MMS check, order >= 3.5)."""

import numpy as np
import pytest

from oracle import ccd as occd
from twophaseflow_amd import Grid, GridConfig, SimulationConfig
from twophaseflow_amd.grid import node_spacing


def _quad_map(N, L, c=0.4):
    xi = np.linspace(0.0, 1.0, N + 1)
    return L * (xi + c * xi * (1.0 - xi))


@pytest.mark.parametrize("N,L", [(16, 1.0), (32, 2.0), (40, 0.02)])
def test_quadratic_map_closed_form(N, L):
    x = _quad_map(N, L)
    o = occd.OracleCCD((N,), (L,), "wall", coords=[x])
    d1, d2 = o.differentiate(x ** 2, 0)
    assert np.abs(d1 - 2.0 * x).max() <= 1e-12 * 2.0 * L
    assert np.abs(d2 - 2.0).max() <= 1e-10
    J, dJ = o.J[0], o.dJ_dxi[0]
    xx = L * (1.0 + 0.4 * (1.0 - 2.0 * np.linspace(0.0, 1.0, N + 1)))  # x_xi
    assert np.allclose(J, 1.0 / xx, rtol=1e-12, atol=0)
    assert np.allclose(dJ, 0.8 * L / xx ** 2, rtol=1e-10, atol=0)


def test_linear_map_metric_is_inverse_length():
    J, dJ = occd.metric_from_coords(np.linspace(0.0, 3.0, 33))
    assert np.allclose(J, 1.0 / 3.0, rtol=1e-14, atol=0)
    assert np.abs(dJ).max() <= 1e-13


def test_metric_converges_on_smooth_map():
    """test_grid.py's MMS check on J: sinusoidal map, order >= 3.5."""
    L, A, errs, Ns = 1.0, 0.05, [], (16, 32, 64, 128)
    for n in Ns:
        xi = np.linspace(0.0, L, n + 1)
        x = xi + A * np.sin(2.0 * np.pi * xi / L)
        J, _ = occd.metric_from_coords(x)
        J_exact = 1.0 / (1.0 + A * (2.0 * np.pi / L) * np.cos(2.0 * np.pi * xi / L))
        errs.append(np.abs(J - J_exact).max())
    order = np.polyfit(np.log([L / n for n in Ns]), np.log(errs), 1)[0]
    assert order >= 3.5, (order, errs)


def test_periodic_axis_on_nonuniform_map_converges():
    """A periodic axis keeps its cyclic solve in xi; the metric comes from
    the wall solve of x(xi).  f = sin(2 pi x) on a stretched map: f' error
    falls at >= 4th order."""
    errs, Ns = [], (32, 64, 128)
    for n in Ns:
        xi = np.linspace(0.0, 1.0, n + 1)
        x = xi + 0.03 * np.sin(2.0 * np.pi * xi)
        o = occd.OracleCCD((n, n), (1.0, 1.0), "periodic", coords=[x, None])
        X = x[:, None] * np.ones((1, n + 1))
        d1, d2 = o.differentiate(np.sin(2.0 * np.pi * X), 0)
        errs.append(max(np.abs(d1 - 2.0 * np.pi * np.cos(2.0 * np.pi * X)).max(),
                        np.abs(d2 + (2.0 * np.pi) ** 2 * np.sin(2.0 * np.pi * X)).max()
                        / (2.0 * np.pi)))
    slopes = np.log(np.array(errs[:-1]) / np.array(errs[1:])) / np.log(2.0)
    assert slopes.min() >= 4.0, (errs, slopes)


def test_planted_metric_mistakes_fail_the_closed_form(monkeypatch):
    """Each plausible mistake in apply_ccd_metric breaks the quadratic-map pin."""
    N, L = 32, 2.0
    x = _quad_map(N, L)
    good = occd.apply_ccd_metric
    mistakes = [
        lambda d1, d2, J, dJ, ax: (J * d1, J * J * d2),                 # dropped J dJ term
        lambda d1, d2, J, dJ, ax: (J * d1, J * J * d2 - J * dJ * d1),   # wrong sign
        lambda d1, d2, J, dJ, ax: (J * d1, J * d2 + J * dJ * d1),       # J instead of J^2
        lambda d1, d2, J, dJ, ax: good(d1, d2, J, 0.5 * dJ, ax),        # wrong dJ scale
    ]
    for bad in mistakes:
        monkeypatch.setattr(occd, "apply_ccd_metric", bad)
        o = occd.OracleCCD((N,), (L,), "wall", coords=[x])
        d1, d2 = o.differentiate(x ** 2, 0)
        assert np.abs(d2 - 2.0).max() > 1e-3
    monkeypatch.setattr(occd, "apply_ccd_metric", good)


# ---- host side of non-uniform grids (no GPU) --------------------------------

def _grid(N=(4, 4), L=(1.0, 1.0), alpha=2.0):
    return Grid(SimulationConfig(grid=GridConfig(ndim=len(N), N=N, L=L,
                                                 alpha_grid=alpha)).grid)


def test_nonuniform_metric_requires_ccd_solver():
    """test_grid.py:196-212: non-uniform coordinates need the CCD metric."""
    g = _grid()
    c = np.asarray([0.0, 0.1, 0.4, 0.7, 1.0])
    g.coords[0] = c
    g.h[0] = node_spacing(c)
    with pytest.raises(ValueError, match="non-uniform grid metrics require a CCDSolver"):
        g._build_metrics()


def test_nonuniform_grid_starts_with_the_uniform_coordinate_metric():
    g = _grid(N=(8, 16), L=(0.02, 3.0))
    assert not g.uniform and _grid(alpha=1.0).uniform
    for a, L in enumerate((0.02, 3.0)):
        assert np.allclose(g.J[a], 1.0 / L, rtol=1e-14, atol=0)
        assert np.abs(g.dJ_dxi[a]).max() == 0.0


def test_set_coordinates_validates():
    g = _grid()
    with pytest.raises(ValueError, match="needs 5"):
        g.set_coordinates(0, [0.0, 0.5, 1.0])
    with pytest.raises(ValueError, match="increase"):
        g.set_coordinates(0, [0.0, 0.5, 0.4, 0.8, 1.0])
    with pytest.raises(ValueError, match="from 0 to L"):
        g.set_coordinates(0, [0.0, 0.2, 0.4, 0.8, 1.1])
    with pytest.raises(ValueError, match="require a CCDSolver"):
        g.set_coordinates(0, [0.0, 0.1, 0.4, 0.7, 1.0])


def test_other_components_refuse_nonuniform_grids():
    from twophaseflow_amd import PPESolverFVMMatrixFree
    from twophaseflow_amd.rhie_chow import RhieChowInterpolator
    cfg = SimulationConfig(grid=GridConfig(ndim=2, N=(8, 8), L=(1.0, 1.0), alpha_grid=2.0))
    g = Grid(cfg.grid)
    with pytest.raises(NotImplementedError, match="CCDSolver only"):
        PPESolverFVMMatrixFree(None, cfg, g)
    with pytest.raises(NotImplementedError, match="CCDSolver only"):
        RhieChowInterpolator(None, g, None)
