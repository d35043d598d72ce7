"""Pins for the FVM PPE oracle.  CPU only."""

import json
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import ppe as oppe
from oracle import fields as ofields
from oracle import ccd as occd

HERE = os.path.dirname(os.path.abspath(__file__))
KNOWN = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))


def test_matrix_free_matches_assembled_csr():
    """tests/test_fvm_matrixfree.py:33-56 (2-D) plus a 3-D case."""
    t = KNOWN["fvm_mf_vs_csr"]
    for N in [(t["N"], t["N"]), (5, 6, 7)]:
        rng = np.random.default_rng(t["seed"])
        o = oppe.OraclePPE(N)
        rho = 1.0 + rng.uniform(0.0, 1.0, o.shape)
        p = rng.standard_normal(o.shape)
        o.prepare_operator(rho)
        out = o.apply(p).ravel()
        ref = o.csr(rho) @ p.ravel()
        np.testing.assert_allclose(out, ref, rtol=t["rtol"], atol=t["atol"])


def test_interior_row_sums_vanish():
    """tests/test_pressure.py:51-69: constant p is in the (unpinned) null space."""
    o = oppe.OraclePPE((9, 10, 11))
    rho = 1.0 + np.random.default_rng(3).uniform(0, 1, o.shape)
    o.prepare_operator(rho)
    out = o.apply(np.ones(o.shape)).ravel()
    out[o.pin] = 0.0
    assert np.max(np.abs(out)) < 1e-10


def test_cg_system_is_symmetric_positive_definite():
    o = oppe.OraclePPE((6, 5, 4))
    rho = 1e-3 + np.random.default_rng(4).uniform(0, 1, o.shape)
    o.prepare_operator(rho)
    n = int(np.prod(o.shape))
    A = np.stack([o.system_matvec(e) for e in np.eye(n)], axis=1)
    assert np.max(np.abs(A - A.T)) <= 1e-12 * np.max(np.abs(A))
    assert np.min(np.linalg.eigvalsh(0.5 * (A + A.T))) > 0.0


def test_one_dimensional_manufactured_answer():
    """1-D wall Laplacian with rho=1 on a quadratic: interior rows give the
    exact second difference, wall rows the half-cell flux balance."""
    N = 10
    o = oppe.OraclePPE((N,), (2.0,))
    o.prepare_operator(np.ones(o.shape))
    x = np.linspace(0, 2.0, N + 1)
    h = 0.2
    out = o.apply(x ** 2)
    assert np.allclose(out[1:-1][np.arange(N - 1) + 1 != o.pin], 2.0)
    assert np.isclose(out[0], (x[1] ** 2 - x[0] ** 2) / h / (h / 2))
    assert np.isclose(out[-1], (x[-2] ** 2 - x[-1] ** 2) / h / (h / 2))


def test_jacobi_cg_matches_direct():
    """tests/test_fvm_matrixfree.py:153-191."""
    t = KNOWN["fvm_solve_vs_direct"]
    rng = np.random.default_rng(t["seed"])
    o = oppe.OraclePPE((t["N"], t["N"]))
    rho = 1.0 + rng.uniform(0.0, 1.0, o.shape)
    rhs = rng.standard_normal(o.shape)
    p_cg, info = o.solve_cg(rhs, rho, tol=1e-10, maxiter=200)
    p_ref = o.solve_direct(rhs, rho)
    np.testing.assert_allclose(p_cg, p_ref, rtol=t["rtol"], atol=t["atol"])
    assert info["relres"] < 1e-9


def test_jacobi_cg_density_ratio_1000_3d():
    o = oppe.OraclePPE((12, 12, 12))
    psi = ofields.gas_bubbles_psi(o.N, o.L, [(0.3, 0.5, 0.5), (0.7, 0.6, 0.4)],
                                  [0.2, 0.15], 1.5 / 12)
    rho = ofields.density(psi, 1.0, 1e-3)
    rhs = np.random.default_rng(5).standard_normal(o.shape)
    p, info = o.solve_cg(rhs, rho, tol=1e-10, maxiter=5000)
    p_ref = o.solve_direct(rhs, rho)
    assert np.max(np.abs(p - p_ref)) <= 1e-7 * np.max(np.abs(p_ref))
    assert info["relres"] < 1e-9


def test_paper_table_U2b():
    """paper/sections/12u2_ccd_poisson_ppe_bc.tex:71-75."""
    t = KNOWN["fvm_ppe_U2b"]
    for N, ref in zip(t["N"][:4], t["linf"][:4]):
        o = oppe.OraclePPE((N, N))
        x = np.linspace(0, 1, N + 1)
        X, Y = np.meshgrid(x, x, indexing="ij")
        pe = np.cos(np.pi * X) * np.cos(np.pi * Y)
        p = o.solve_direct(-2 * np.pi ** 2 * pe, np.ones(o.shape))
        p = p - (p.ravel()[o.pin] - pe.ravel()[o.pin])
        err = np.max(np.abs(p - pe))
        assert abs(err - ref) <= 0.01 * ref, (N, err, ref)


def test_mass_and_volumes():
    N, L = (32, 32, 32), (1.0, 1.0, 1.0)
    dV = ofields.cell_volumes(N, L, ("wall",) * 3)
    assert np.isclose(dV.sum(), (33 / 32) ** 3)  # full h at wall nodes (grid.py:66-69)
    dVp = ofields.cell_volumes(N, L, ("periodic",) * 3)
    assert np.isclose(dVp.sum(), 1.0)
    psi = ofields.gas_bubbles_psi((64, 64, 64), L, [(0.5, 0.5, 0.5)], [0.25], 1.5 / 64)
    gas = ofields.mass(1.0 - psi, ofields.cell_volumes((64,) * 3, L, ("periodic",) * 3))
    # logistic smoothing of width eps adds pi^2 eps^2 r to r^3 (its 2nd moment)
    eps, r = 1.5 / 64, 0.25
    assert abs(gas - 4 / 3 * np.pi * (r ** 3 + np.pi ** 2 * eps ** 2 * r)) < 2e-4


def test_mass_correction_restores_target():
    N, L = (32, 32), (1.0, 1.0)
    psi = ofields.gas_bubbles_psi(N, L, [(0.5, 0.5)], [0.25], 1.5 / 32)
    dV = ofields.cell_volumes(N, L, ("wall", "wall"))
    m0 = ofields.mass(psi, dV)
    q = ofields.mass_correction(psi * 0.999, dV, m0)
    assert abs(ofields.mass(q, dV) - m0) < 1e-10


def test_projection_reduces_divergence():
    """Projection of a wall-compatible smooth field (zero normal velocity):
    the interior CCD divergence after correction falls by > 100x and keeps
    falling under refinement (FVM/CCD mismatch is O(h^2))."""
    N = 24
    o = oppe.OraclePPE((N, N))
    ccd = occd.OracleCCD((N, N), (1.0, 1.0), "wall")
    x = np.linspace(0, 1, N + 1)
    X, Y = np.meshgrid(x, x, indexing="ij")
    u = [np.sin(np.pi * X) * Y * (1 - Y), np.sin(np.pi * Y) * X * (1 - X) * 0.3]
    rho = np.ones(o.shape)
    div0 = ofields.divergence(ccd, u)
    p = o.solve_direct(div0, rho)
    u1 = ofields.correct_velocity(ccd, u, p, rho, 1.0)
    div1 = ofields.divergence(ccd, u1)
    assert np.max(np.abs(div1[2:-2, 2:-2])) < 0.01 * np.max(np.abs(div0[2:-2, 2:-2]))


# --- GMRES form: line solves, preconditioners, SciPy gmres --------------------
def _tridiag_dense(a, d, c):
    n = d.size
    A = np.diag(d)
    A[np.arange(1, n), np.arange(n - 1)] = a[1:]
    A[np.arange(n - 1), np.arange(1, n)] = c[:-1]
    return A


def _random_lines(shape, axis, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1.0, 0.0, shape)
    c = rng.uniform(-1.0, 0.0, shape)
    d = 2.5 + rng.uniform(0.0, 1.0, shape)   # diagonally dominant
    r = rng.standard_normal(shape)
    return a, d, c, r


@pytest.mark.parametrize("shape,axis", [((37,), 0), ((9, 13), 0), ((9, 13), 1), ((5, 6, 33), 2)])
def test_thomas_and_full_pcr_match_dense_lines(shape, axis):
    """Both line solvers against numpy.linalg.solve line by line."""
    a, d, c, r = _random_lines(shape, axis, 11)
    xt = oppe.thomas_lines(a, d, c, r, axis)
    xp_ = oppe.pcr_lines(a, d, c, r, axis, max_stages=None)
    A_, D_, C_, R_ = (np.moveaxis(x, axis, -1).reshape(-1, shape[axis]) for x in (a, d, c, r))
    ref = np.stack([np.linalg.solve(_tridiag_dense(A_[k], D_[k], C_[k]), R_[k])
                    for k in range(A_.shape[0])])
    ref = np.moveaxis(ref.reshape(np.moveaxis(a, axis, -1).shape), -1, axis)
    np.testing.assert_allclose(xt, ref, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(xp_, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("stages", [1, 2, 3, 4])
def test_pcr_reduction_keeps_the_exact_solution(stages):
    """After k stages row i couples x[i -+ 2^k] only, and the exact solution
    still satisfies every reduced row (a wrong sign or neighbour fails)."""
    n = 37
    a, d, c, r = _random_lines((n,), 0, 5)
    a[0], c[-1] = 0.0, 0.0   # no couplings past the line ends
    x = np.linalg.solve(_tridiag_dense(a, d, c), r)
    a2, d2, c2, b2 = oppe.pcr_reduce(a, d, c, r, 0, stages)
    s = 2 ** stages
    lhs = d2 * x
    lhs[s:] += a2[s:] * x[:-s]
    lhs[:-s] += c2[:-s] * x[s:]
    np.testing.assert_allclose(lhs, b2, rtol=1e-12, atol=1e-12)
    assert np.all(a2[:s] == 0.0) and np.all(c2[n - s:] == 0.0)


def test_truncated_pcr_is_exact_on_short_lines():
    a, d, c, r = _random_lines((16, 7), 0, 9)
    np.testing.assert_allclose(oppe.pcr_lines(a, d, c, r, 0, max_stages=4),
                               oppe.thomas_lines(a, d, c, r, 0), rtol=1e-13, atol=1e-13)
    assert oppe.pcr_stage_count(257, 4) == 4 and oppe.pcr_stage_count(257, None) == 9
    assert oppe.pcr_stage_count(16, 0) == 1


def test_line_preconditioner_uniform_density_textbook_rows():
    """rho = 1: per-axis rows (shift + 2/h^2) z_i - z_{i+-1}/h^2 with the wall
    face doubled (dv = h/2); z = mean over axes of the line solves, z[pin] = 0."""
    N, c_tau = (12, 8), 2.0
    o = oppe.OraclePPE(N)
    rho = np.ones(o.shape)
    o.prepare_operator(rho)
    r = np.random.default_rng(3).standard_normal(o.shape)
    rw = r.copy()
    rw.ravel()[o.pin] = 0.0
    h_min = 1.0 / max(N)
    shift = 2.0 / (c_tau * h_min ** 2)
    z = np.zeros_like(r)
    for ax in range(2):
        n, h = o.shape[ax], 1.0 / N[ax]
        A = np.zeros((n, n))
        for i in range(n):
            if i > 0:
                A[i, i - 1] = -(2.0 if i == n - 1 else 1.0) / h ** 2
            if i < n - 1:
                A[i, i + 1] = -(2.0 if i == 0 else 1.0) / h ** 2
            A[i, i] = shift - A[i].sum()
        z += np.moveaxis(np.linalg.solve(A, np.moveaxis(rw, ax, 0).reshape(n, -1))
                         .reshape(np.moveaxis(rw, ax, 0).shape), 0, ax)
    z /= 2.0
    z.ravel()[o.pin] = 0.0
    for stages in (None, 10):
        np.testing.assert_allclose(o.gmres_precond(r, "line_pcr", c_tau, stages), z,
                                   rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("pre,seed", [("line_pcr", 456), ("jacobi", 789), ("none", 456)])
def test_gmres_matches_direct(pre, seed):
    """test_fvm_matrixfree.py:88-103, 153-169: N=8, tol 1e-10, maxiter 200."""
    o = oppe.OraclePPE((8, 8))
    rng = np.random.default_rng(seed)
    rho = 1.0 + rng.uniform(0.0, 1.0, o.shape)
    rhs = rng.standard_normal(o.shape)
    p, info = o.solve_gmres(rhs, rho, tol=1e-10, maxiter=200, preconditioner=pre)
    assert info["info"] == 0 and info["relres"] <= 1e-10
    np.testing.assert_allclose(p, o.solve_direct(rhs, rho), rtol=1e-9, atol=1e-10)
