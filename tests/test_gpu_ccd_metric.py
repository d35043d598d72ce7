"""Non-uniform grids on the GPU: CCDSolver solves every line in the
computational coordinate xi (spacing 1/N) and maps the result with the
grid metric J = dxi/dx, dJ/dxi (ccd_solver.py:126-130, 169-170;
ccd_solver_helpers.py:509-519).  Checked against the oracle, which builds
its own metric from the same coordinates in FP64 dense LU: max-norm error
relative to max|reference| <= 1e-12 for f' and f'' on every axis (wall,
periodic, strided and contiguous, 2-D and 3-D), the metric itself against
the long-double solve of the coordinate line (bound in _check_metric), and
apply_metric=False, the d1-only / d2-only paths (bit-identical to the full
call), the fused all-axes and per-axis-field launches, and the closed form
f = x^2 on a quadratic map (f' = 2x, f'' = 2)."""

import sys

import numpy as np
import pytest
import torch

from oracle import ccd as occd
from twophaseflow_amd import CCDBatch, CCDSolver, Grid, GridConfig

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _quad(xi):
    return xi + 0.4 * xi * (1.0 - xi)


def _tanh(xi):  # nodes clustered around xi = 0.5 (an interface-fitted layout)
    return 0.5 + np.tanh(2.0 * (xi - 0.5)) / (2.0 * np.tanh(1.0))


def _sine(xi):  # periodic-compatible stretching
    return xi + 0.04 * np.sin(2.0 * np.pi * xi)


def _setup(N, L, bc, maps, precision="fp64"):
    g = Grid(GridConfig(ndim=len(N), N=tuple(N), L=tuple(L), alpha_grid=2.0))
    ccd = CCDSolver(g, bc_type=bc, precision=precision)
    for a, m in enumerate(maps):
        if m is not None:
            g.set_coordinates(a, L[a] * m(np.linspace(0.0, 1.0, N[a] + 1)), ccd)
    o = occd.OracleCCD(N, L, bc, coords=[g.coords[a] if maps[a] is not None else None
                                         for a in range(len(N))])
    return g, ccd, o


def _check_metric(g, a):
    """The grid metric is the CCD of the coordinate line x(xi) (wall
    closure, spacing 1/N): its x_xi = 1/J and x_xixi = -dJ/J^2 are checked
    against the long-double solve of the same system.  x_xi to 1e-12 of
    max|x_xi|; x_xixi to 1e-12 of max|x_xixi| or 1e-14 of max|x| N^2, the
    scale of the second differences it cancels (a smooth line's x_xixi is
    far below that scale: both the kernel and the dense LU sit ~1e-15 of it
    off the exact discrete answer)."""
    x = np.asarray(g.coords[a])
    N = x.shape[0] - 1
    e1, e2 = occd.differentiate_wall_extended(x, 0, occd.build_wall_axis(N + 1, 1.0 / N))
    J, dJ = np.asarray(g.J[a]), np.asarray(g.dJ_dxi[a])
    x1, x2 = 1.0 / J, -dJ / J ** 2
    assert np.abs(x1 - e1).max() <= 1e-12 * np.abs(e1).max(), a
    tol2 = max(1e-12 * np.abs(e2).max(), 1e-14 * np.abs(x).max() * N * N)
    assert np.abs(x2 - e2).max() <= tol2, (a, np.abs(x2 - e2).max(), tol2)


CASES = [((48, 40), (1.0, 2.0), "wall", (_quad, _tanh)),
         ((40, 24, 20), (1.0, 0.5, 2.0), ("wall", "periodic", "wall"), (_tanh, _sine, None)),
         ((160, 40, 160), (1.0, 1.0, 1.0), "periodic", (_sine, None, _sine)),
         ((64, 64), (0.02, 0.02), "wall", (None, _quad))]


@pytest.mark.parametrize("N,L,bc,maps", CASES)
def test_metric_matches_oracle(N, L, bc, maps):
    _announce(f"ccd metric N={N} bc={bc}")
    g, ccd, o = _setup(N, L, bc, maps)
    for a in range(len(N)):
        _check_metric(g, a)
    f = np.random.default_rng(7).standard_normal(g.shape)
    ft = torch.from_numpy(f).cuda()
    for a in range(len(N)):
        r1, r2 = o.differentiate(f, a)
        d1, d2 = ccd.differentiate(f, a)
        assert _rel(d1, r1) <= 1e-12, (a, _rel(d1, r1))
        assert _rel(d2, r2) <= 1e-12, (a, _rel(d2, r2))
        assert np.array_equal(ccd.first_derivative(f, a), d1), a
        assert np.array_equal(ccd.second_derivative(f, a), d2), a
        q1, q2 = o.differentiate(f, a, apply_metric=False)
        w1, w2 = ccd.differentiate(f, a, apply_metric=False)
        assert _rel(w1, q1) <= 1e-12 and _rel(w2, q2) <= 1e-12, a
    # fused launch of all axes, with and without d1
    nd = len(N)
    D1 = torch.empty((nd,) + g.shape, dtype=torch.float64, device="cuda")
    D2 = torch.empty_like(D1)
    ccd.apply_into(ft, D1, D2, (1 << nd) - 1)
    E2 = torch.empty_like(D1)
    ccd.apply_into(ft, None, E2, (1 << nd) - 1)
    torch.cuda.synchronize()
    for a in range(nd):
        r1, r2 = o.differentiate(f, a)
        assert _rel(D1[a].cpu().numpy(), r1) <= 1e-12, a
        assert _rel(D2[a].cpu().numpy(), r2) <= 1e-12, a
        assert torch.equal(E2[a], D2[a]), a
    # a different field per axis (apply_fields_into)
    fs = [torch.from_numpy(np.random.default_rng(20 + a).standard_normal(g.shape)).cuda()
          for a in range(nd)]
    o1 = [torch.empty_like(fs[0]) for _ in range(nd)]
    o2 = [torch.empty_like(fs[0]) for _ in range(nd)]
    ccd.apply_fields_into(fs, None, o2, (1 << nd) - 1)
    ccd.apply_fields_into(fs, o1, None, (1 << nd) - 1)
    torch.cuda.synchronize()
    for a in range(nd):
        r1, r2 = o.differentiate(fs[a].cpu().numpy(), a)
        assert _rel(o1[a].cpu().numpy(), r1) <= 1e-12, a
        assert _rel(o2[a].cpu().numpy(), r2) <= 1e-12, a


def test_quadratic_map_closed_form():
    _announce("ccd metric closed form x^2 on a quadratic map")
    N, L = (32, 48), (2.0, 0.5)
    g, ccd, _ = _setup(N, L, "wall", (_quad, _quad))
    X, Y = g.meshgrid()
    for a, C in ((0, X), (1, Y)):
        d1, d2 = ccd.differentiate(X ** 2 + Y ** 2, a)
        assert np.abs(d1 - 2.0 * C).max() <= 1e-11 * L[a], a
        assert np.abs(d2 - 2.0).max() <= 1e-9, a


def test_short_domain_uses_dimensionless_xi():
    """ccd test_nonuniform_metric_uses_dimensionless_xi_for_short_domain
    (test_ccd.py:355-382): a non-uniform grid of L = 0.02 with its uniform
    starting coordinates must not scale the derivatives by 1/L twice."""
    _announce("ccd metric short domain")
    length = 0.02
    g = Grid(GridConfig(ndim=2, N=(32, 32), L=(length, length), alpha_grid=2.0))
    ccd = CCDSolver(g, bc_type="wall")
    X, _ = g.meshgrid()
    d1, d2 = ccd.differentiate(np.sin(np.pi * X / length), 0)
    k = np.pi / length
    s = (slice(3, -3), slice(None))
    np.testing.assert_allclose(d1[s], k * np.cos(k * X)[s], rtol=2e-4, atol=2e-2)
    np.testing.assert_allclose(d2[s], -k * k * np.sin(k * X)[s], rtol=2e-3, atol=3.0)


def test_metric_refresh_fp32_refined_and_batch_rejection():
    _announce("ccd metric: refresh after new coordinates, fp32_refined, batch")
    N, L = (40, 36), (1.0, 1.0)
    g, ccd, _ = _setup(N, L, "wall", (_quad, None))
    f = np.random.default_rng(9).standard_normal(g.shape)
    ccd.differentiate(f, 0)
    g.set_coordinates(0, _tanh(np.linspace(0.0, 1.0, N[0] + 1)), ccd)  # new metric arrays
    o = occd.OracleCCD(N, L, "wall", coords=[g.coords[0], None])
    r1, r2 = o.differentiate(f, 0)
    d1, d2 = ccd.differentiate(f, 0)
    assert _rel(d1, r1) <= 1e-12 and _rel(d2, r2) <= 1e-12
    ref = CCDSolver(g, bc_type="wall", precision="fp32_refined")
    e1, e2 = ref.differentiate(f, 0)
    assert _rel(e1, r1) <= 1e-11 and _rel(e2, r2) <= 1e-11
    with pytest.raises(ValueError, match="uniform-grid plans"):
        CCDBatch([ccd])
