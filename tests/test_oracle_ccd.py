"""Pins for the CCD oracle: known answers the reference holds, closed forms,
and mathematical invariants.  CPU only."""

import json
import os

import numpy as np
import pytest

from oracle import ccd as occd

HERE = os.path.dirname(os.path.abspath(__file__))
KNOWN = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))


def _u1a_errors(N, mod=occd):
    x = np.linspace(0.0, 1.0, N + 1)
    X, Y = np.meshgrid(x, x, indexing="ij")
    f = np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)
    info = mod.build_periodic_axis(N + 1, 1.0 / N)
    d1, d2 = mod.differentiate_periodic(f, 0, info)
    e1 = np.max(np.abs(d1 - 2 * np.pi * np.cos(2 * np.pi * X) * np.sin(2 * np.pi * Y)))
    e2 = np.max(np.abs(d2 + (2 * np.pi) ** 2 * f))
    return e1, e2


def test_periodic_matches_paper_table_U1a():
    """paper/sections/12u1_ccd_operator.tex:46-50 (3 significant digits)."""
    t = KNOWN["ccd_periodic_U1a"]
    for N, r1, r2 in zip(t["N"], t["linf_d1"], t["linf_d2"]):
        e1, e2 = _u1a_errors(N)
        assert abs(e1 - r1) <= 0.01 * r1 + 5e-15, (N, e1, r1)
        if N < 256:
            assert abs(e2 - r2) <= 0.01 * r2, (N, e2, r2)
        else:
            # round-off floor entry (dagger in the table): same order only
            assert 0.2 * r2 < e2 < 5.0 * r2, (N, e2, r2)


def test_periodic_dense_lu_matches_closed_form_symbol():
    rng = np.random.default_rng(7)
    for N in (8, 17, 64, 256):
        h = 0.37 / N
        f = rng.standard_normal((N + 1, 3))
        info = occd.build_periodic_axis(N + 1, h)
        d1, d2 = occd.differentiate_periodic(f, 0, info)
        s1, s2 = occd.periodic_spectral(f, 0, h)
        assert np.max(np.abs(d1 - s1)) <= 1e-12 * np.max(np.abs(s1))
        assert np.max(np.abs(d2 - s2)) <= 1e-12 * np.max(np.abs(s2))


def test_symbol_low_frequency_limit():
    """D1 -> i*k, D2 -> -k^2 with 6th-order error."""
    h = 1e-2
    th = np.array([1e-3, 1e-2])
    s1, s2 = occd.periodic_symbol(th, h)
    k = th / h
    assert np.allclose(s1.imag, k, rtol=1e-12)
    assert np.allclose(s2, -k * k, rtol=1e-10)


def test_periodic_unique_nodes_and_image():
    """tests/test_ccd.py:65-81."""
    t = KNOWN["ccd_periodic_unique_nodes"]
    N = t["N"]
    x = np.linspace(0.0, 1.0, N + 1)
    f = np.broadcast_to(np.sin(2 * np.pi * x)[:, None], (N + 1, 5)).copy()
    solver = occd.OracleCCD((N, 4), (1.0, 1.0), "periodic")
    d1, d2 = solver.differentiate(f, 0)
    np.testing.assert_allclose(d1[-1], d1[0], rtol=0, atol=t["atol_image"])
    np.testing.assert_allclose(d2[-1], d2[0], rtol=0, atol=t["atol_image"])
    np.testing.assert_allclose(d1[:, 2], 2 * np.pi * np.cos(2 * np.pi * x), atol=t["atol_d1"])
    np.testing.assert_allclose(d2[:, 2], -(2 * np.pi) ** 2 * f[:, 2], atol=t["atol_d2"])


def test_wall_cubic_exact():
    """tests/test_ccd.py:201-221."""
    t = KNOWN["ccd_wall_cubic_exact"]
    N = t["N"]
    x = np.linspace(0.0, 1.0, N + 1)
    f = np.broadcast_to((x ** 3)[:, None], (N + 1, 5)).copy()
    d1, d2 = occd.OracleCCD((N, 4), (1.0, 1.0)).differentiate(f, 0)
    assert np.max(np.abs(d1[:, 2] - 3 * x ** 2)) < t["tol_d1"]
    assert np.max(np.abs(d2[:, 2] - 6 * x)) < t["tol_d2"]


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_wall_polynomials_exact_all_degrees(deg):
    N, L = 24, 1.7
    x = np.linspace(0.0, L, N + 1)
    d1, d2 = occd.OracleCCD((N,), (L,)).differentiate(x ** deg, 0)
    e1 = deg * x ** max(deg - 1, 0)
    e2 = deg * (deg - 1) * x ** max(deg - 2, 0)
    assert np.max(np.abs(d1 - e1)) < 1e-10
    assert np.max(np.abs(d2 - e2)) < 1e-8


def test_wall_axis_independence():
    """tests/test_ccd.py:226-252."""
    t = KNOWN["ccd_wall_axis_independence"]
    N = t["N"]
    g = np.linspace(0, 1, N + 1)
    X, Y = np.meshgrid(g, g, indexing="ij")
    f = X ** 3 * Y ** 2
    s = occd.OracleCCD((N, N), (1.0, 1.0))
    assert np.max(np.abs(s.differentiate(f, 0)[0] - 3 * X ** 2 * Y ** 2)) < t["tol"]
    assert np.max(np.abs(s.differentiate(f, 1)[0] - 2 * X ** 3 * Y)) < t["tol"]


def test_wall_d2_nullspace_dim_two():
    """tests/test_ccd.py:257-299."""
    N = 8
    _, D2 = occd.OracleCCD((N, 4), (1.0, 1.0)).differentiate(np.eye(N + 1), 0)
    assert N + 1 - np.linalg.matrix_rank(D2, tol=1e-8) == 2
    assert np.max(np.abs(D2 @ np.ones(N + 1))) < 1e-10
    assert np.max(np.abs(D2 @ (np.arange(N + 1) / N))) < 1e-10


def test_wall_requires_six_points():
    with pytest.raises(ValueError, match="n_pts >= 6"):
        occd.OracleCCD((4, 4), (1.0, 1.0)).differentiate(np.zeros((5, 5)), 0)


def test_wall_convergence_slopes():
    """tests/test_ccd.py:109-196: wall d1 slope >= 3.5, d2 slope >= 2.5."""
    e1s, e2s, Ns = [], [], [8, 16, 32, 64]
    for N in Ns:
        x = np.linspace(0.0, 1.0, N + 1)
        d1, d2 = occd.OracleCCD((N,), (1.0,)).differentiate(np.sin(2 * np.pi * x), 0)
        e1s.append(np.max(np.abs(d1 - 2 * np.pi * np.cos(2 * np.pi * x))))
        e2s.append(np.max(np.abs(d2 + (2 * np.pi) ** 2 * np.sin(2 * np.pi * x))))
    s1 = [np.log2(e1s[i] / e1s[i + 1]) for i in range(3)]
    s2 = [np.log2(e2s[i] / e2s[i + 1]) for i in range(3)]
    assert min(s1) >= 3.5 and min(s2) >= 2.5, (s1, s2)


@pytest.mark.parametrize("n_pts", [7, 65, 256, 1025])
def test_block_thomas_matches_dense_lu(n_pts):
    """Unpivoted block-Thomas (block_tridiag.py) vs LAPACK LU on the wall system."""
    h = 1.0 / (n_pts - 1)
    info = occd.build_wall_axis(n_pts, h)
    n = info["n_int"]
    A = info["A"]
    diag = [A[2 * i:2 * i + 2, 2 * i:2 * i + 2] for i in range(n)]
    lower = [np.zeros((2, 2))] + [A[2 * i:2 * i + 2, 2 * i - 2:2 * i] for i in range(1, n)]
    upper = [A[2 * i:2 * i + 2, 2 * i + 2:2 * i + 4] for i in range(n - 1)] + [np.zeros((2, 2))]
    fac = occd.block_thomas_factor(diag, lower, upper)
    rhs = np.random.default_rng(1).standard_normal((n, 2, 3))
    x = occd.block_thomas_solve(fac, rhs)
    ref = np.linalg.solve(A, rhs.reshape(2 * n, -1)).reshape(n, 2, -1)
    assert np.max(np.abs(x - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_block_thomas_rejects_singular_pivot():
    """tests/test_ccd.py:51-62."""
    diag = [np.array([[1.0, 2.0], [2.0, 4.0]]), np.eye(2)]
    with pytest.raises(np.linalg.LinAlgError, match="pseudo-inverse"):
        occd.block_thomas_factor(diag, [np.zeros((2, 2)), np.eye(2)],
                                 [np.eye(2), np.zeros((2, 2))])


def test_pins_catch_planted_coefficient_errors(monkeypatch):
    """A wrong sign of B2 breaks the U1-a table; a wrong closure weight breaks
    cubic exactness.  Mistakes are planted in the oracle, on the CPU."""
    monkeypatch.setattr(occd, "B2", +9.0 / 8.0)
    e1, _ = _u1a_errors(64)
    assert abs(e1 - 5.97e-10) > 0.01 * 5.97e-10
    monkeypatch.setattr(occd, "B2", -9.0 / 8.0)
    orig = occd.boundary_coeffs_left

    def bad_left(h, n_pts):
        d = orig(h, n_pts)
        d["c_II"] = d["c_II"].copy()
        d["c_II"][5] *= 1.001
        return d

    monkeypatch.setattr(occd, "boundary_coeffs_left", bad_left)
    N = 20
    x = np.linspace(0.0, 1.0, N + 1)
    _, d2 = occd.OracleCCD((N,), (1.0,)).differentiate(x ** 3, 0)
    assert np.max(np.abs(d2 - 6 * x)) > 1e-9


def test_wall_extended_precision_solve():
    """differentiate_wall_extended: the same wall system in long double is
    exact on cubics (CCD is) to rounding, agrees with the FP64 dense LU on
    random data to 1e-13, and on a smooth 1024-node line it moves the dense
    LU's f'' by at most a few 1e-10 of max|f''| (the dense LU's own error)."""
    N = 40
    x = np.linspace(0.0, 1.0, N + 1)
    info = occd.build_wall_axis(N + 1, 1.0 / N)
    f = np.stack([x ** 3, 0.5 * x ** 2 - x, np.ones_like(x)], axis=1)
    d1, d2 = occd.differentiate_wall_extended(f, 0, info)
    assert np.abs(d1[:, 0] - 3 * x ** 2).max() < 1e-13
    assert np.abs(d2[:, 0] - 6 * x).max() < 1e-11
    assert np.abs(d2[:, 1] - 1.0).max() < 1e-11 and np.abs(d1[:, 2]).max() < 1e-13
    g = np.random.default_rng(3).standard_normal((N + 1, 5))
    e1, e2 = occd.differentiate_wall_extended(g, 0, info)
    r1, r2 = occd.differentiate_wall(g, 0, info)
    assert np.abs(e1 - r1).max() <= 1e-13 * np.abs(r1).max()
    assert np.abs(e2 - r2).max() <= 1e-13 * np.abs(r2).max()
    n = 1024
    xs = np.linspace(0.0, 1.0, n + 1)
    big = occd.build_wall_axis(n + 1, 1.0 / n)
    s = np.exp(np.sin(np.pi * xs))[:, None]
    e1, e2 = occd.differentiate_wall_extended(s, 0, big)
    r1, r2 = occd.differentiate_wall(s, 0, big)
    err = np.abs(e2 - r2).max() / np.abs(e2).max()
    assert 1e-14 < err < 1e-9, err
