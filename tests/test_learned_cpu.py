"""CPU checks of the learned preconditioner's fixed grid transfers and of
the FCG bookkeeping on the oracle (no device code)."""

import numpy as np
import torch

from twophaseflow_amd.learned import prolong3, prolong_axis, restrict3, restrict_axis


def test_restriction_preserves_linear_fields_and_constants():
    x = torch.linspace(0, 1, 17, dtype=torch.float64)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    f = 1.0 + 2 * X - Y + 0.5 * Z
    c = restrict3(f)
    xc = torch.linspace(0, 1, 9, dtype=torch.float64)
    Xc, Yc, Zc = torch.meshgrid(xc, xc, xc, indexing="ij")
    ref = 1.0 + 2 * Xc - Yc + 0.5 * Zc
    # interior nodes are exact for linear fields; walls use even reflection
    assert torch.allclose(c[1:-1, 1:-1, 1:-1], ref[1:-1, 1:-1, 1:-1], atol=1e-12)
    assert torch.allclose(restrict3(torch.ones(17, 17, 17)), torch.ones(9, 9, 9))


def test_prolongation_is_linear_interpolation():
    xc = torch.linspace(0, 1, 9, dtype=torch.float64)
    f = prolong_axis(3 * xc - 1, 0)
    assert torch.allclose(f, 3 * torch.linspace(0, 1, 17, dtype=torch.float64) - 1)
    c = torch.rand(5, 5, 5, dtype=torch.float64)
    p = prolong3(c)
    assert p.shape == (9, 9, 9) and torch.equal(p[::2, ::2, ::2], c)


def test_restriction_is_scaled_transpose_of_prolongation_in_the_interior():
    """<R u, v> = <u, P v> / 2 per axis away from the walls (full weighting)."""
    rng = np.random.default_rng(0)
    u = torch.zeros(17, dtype=torch.float64)
    u[4:13] = torch.from_numpy(rng.standard_normal(9))
    v = torch.zeros(9, dtype=torch.float64)
    v[3:6] = torch.from_numpy(rng.standard_normal(3))
    lhs = float((restrict_axis(u, 0) * v).sum())
    rhs = float((u * prolong_axis(v, 0)).sum()) / 2
    assert abs(lhs - rhs) < 1e-12

