"""Pins for the multigrid oracle (CPU only).

The fine operator is checked against the reference CG system of the PPE
oracle; coarse conductances against closed forms (2h rediscretisation for
constant density, exact Schur complement in 1-D); restriction against the
adjoint of prolongation; the cycles for symmetry and contraction.
"""

import numpy as np
import pytest

from oracle import mg as omg
from oracle import ppe as oppe
from oracle import fields as ofields


def _slots(a):
    a = np.asarray(a)
    return a.reshape((1,) * (3 - a.ndim) + a.shape)


def _bubbles(N, ratio=1e-3):
    nd = len(N)
    c = [(0.3, 0.5, 0.5)[:nd], (0.7, 0.55, 0.4)[:nd], (0.5, 0.25, 0.7)[:nd]]
    psi = ofields.gas_bubbles_psi(N, (1.0,) * nd, c, [0.18, 0.15, 0.12], 1.5 / max(N))
    return ofields.density(psi, 1.0, ratio)


def _h3(N):
    h = [1.0 / n for n in N]
    return [1.0] * (3 - len(N)) + h


@pytest.mark.parametrize("N", [(8, 12), (6, 8, 10)])
def test_fine_operator_is_the_reference_cg_system(N):
    o = oppe.OraclePPE(N)
    rho = _bubbles(N)
    o.prepare_operator(rho)
    T = omg.fine_conductances(_slots(rho), _h3(N))
    x = np.random.default_rng(1).standard_normal(o.shape)
    x.ravel()[o.pin] = 0.0
    ref = o.system_matvec(x.ravel()).reshape(o.shape)
    ref.ravel()[o.pin] = 0.0
    got = omg.apply(T, _slots(x), o.pin).reshape(o.shape)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_constant_density_coarse_is_2h_rediscretisation():
    N = (16, 8, 8)
    rho = np.full(tuple(n + 1 for n in N), 0.37)
    T = omg.fine_conductances(rho, [1.0 / n for n in N])
    Tc = omg.coarse_conductances(T, tuple(n // 2 + 1 for n in N))
    ref = omg.fine_conductances(np.full(Tc[0].shape, 0.37), [2.0 / n for n in N])
    for a in range(3):
        np.testing.assert_allclose(Tc[a], ref[a], rtol=1e-13, atol=1e-15)


def test_one_dimensional_coarse_operator_is_the_schur_complement():
    """Eliminating the odd nodes of a 1-D variable-density operator leaves the
    coarse operator exactly (series conductances)."""
    n = 17
    rho = 1.0 + 999.0 * (np.random.default_rng(3).uniform(size=n) > 0.5)
    T = omg.fine_conductances(_slots(rho), [1.0, 1.0, 1.0 / (n - 1)])
    A = np.zeros((n, n))
    for i in range(n - 1):
        t = T[2].ravel()[i]
        A[i, i] += t
        A[i + 1, i + 1] += t
        A[i, i + 1] -= t
        A[i + 1, i] -= t
    ev, od = np.arange(0, n, 2), np.arange(1, n, 2)
    S = A[np.ix_(ev, ev)] - A[np.ix_(ev, od)] @ np.linalg.solve(A[np.ix_(od, od)],
                                                                 A[np.ix_(od, ev)])
    Tc = omg.coarse_conductances(T, (1, 1, (n - 1) // 2 + 1))[2].ravel()
    np.testing.assert_allclose(-np.diag(S, 1), Tc[:-1], rtol=1e-12)


def test_restriction_is_the_adjoint_of_prolongation():
    fshape, fpin = (9, 17, 13), 5
    cshape, cpin = (5, 9, 7), 3
    rng = np.random.default_rng(4)
    r = rng.standard_normal(fshape)
    r.ravel()[fpin] = 0.0                           # pinned entries carry no data
    xc = rng.standard_normal(cshape)
    xc.ravel()[cpin] = 0.0
    lhs = np.vdot(omg.restrict(r, cshape, cpin), xc)
    rhs = np.vdot(r, omg.prolong(xc, fshape, fpin))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # trilinear interpolation reproduces linear functions away from the pin
    g = np.indices(cshape).astype(float)
    lin = 1.0 + 2.0 * g[0] - 0.5 * g[1] + 0.25 * g[2]
    gf = np.indices(fshape).astype(float) / 2.0
    want = 1.0 + 2.0 * gf[0] - 0.5 * gf[1] + 0.25 * gf[2]
    got = omg.prolong(lin, fshape, fpin)
    mask = np.ones(fshape, bool)
    mask.ravel()[fpin] = False
    np.testing.assert_allclose(got[mask], want[mask], rtol=1e-14)


def test_level_shapes_follow_the_pin():
    shapes, pins = omg.level_shapes((1, 65, 65), np.ravel_multi_index((0, 32, 32), (1, 65, 65)))
    assert shapes == [(1, 65, 65), (1, 33, 33), (1, 17, 17)]
    assert pins[-1] == np.ravel_multi_index((0, 8, 8), (1, 17, 17))
    s2, _ = omg.level_shapes((25, 25, 25), np.ravel_multi_index((12, 12, 12), (25, 25, 25)))
    assert s2 == [(25, 25, 25), (13, 13, 13), (7, 7, 7)]


@pytest.mark.parametrize("gamma", [1, 2])
def test_cycle_is_symmetric_and_contracts(gamma):
    N = (32, 32, 32)
    rho = _bubbles(N)
    o = oppe.OraclePPE(N)
    M = omg.OracleMG(rho, [1.0 / 32] * 3, o.pin)
    assert len(M.shapes) == 3
    rng = np.random.default_rng(7)
    a, b = rng.standard_normal(o.shape), rng.standard_normal(o.shape)
    for v in (a, b):
        v.ravel()[o.pin] = 0.0
    Ma, Mb = M.cycle(a, gamma=gamma), M.cycle(b, gamma=gamma)
    assert abs(np.vdot(Ma, b) - np.vdot(a, Mb)) <= 1e-10 * abs(np.vdot(Ma, b))
    # stationary iteration x <- x + M(b - A x) reduces the A-norm error
    x_true = rng.standard_normal(o.shape)
    x_true.ravel()[o.pin] = 0.0
    rhs = M.apply(0, x_true)
    x = np.zeros(o.shape)
    errs = []
    for _ in range(4):
        e = x - x_true
        errs.append(np.sqrt(np.vdot(e, M.apply(0, e))))
        x = x + M.cycle(rhs - M.apply(0, x), gamma=gamma)
    assert all(e1 < 0.7 * e0 for e0, e1 in zip(errs, errs[1:]))


def test_symmetric_gauss_seidel_converges_to_the_direct_solve():
    N = (8, 8)
    rho = _slots(_bubbles(N, 1e-2))
    o = oppe.OraclePPE(N)
    T = omg.fine_conductances(rho, _h3(N))
    di = omg.diag_inv(T, o.pin)
    b = _slots(np.random.default_rng(2).standard_normal(o.shape))
    b.ravel()[o.pin] = 0.0
    x = np.zeros_like(b)
    omg.smooth(T, di, b, x, 3000, False)
    np.testing.assert_allclose(omg.apply(T, x, o.pin), b, atol=1e-9 * np.abs(b).max())
