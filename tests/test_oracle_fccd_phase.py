"""Oracle pins of the FCCD PPE's phase-separated forms (no GPU).

The reference's own closed-form examples (tests/test_interface_stress_closure.py:
cut-face orientation, the cut-face resistance, the neutral context, the
affine drive staying nonzero with one pin) are restated on the oracle; a
straight interface gives the affine drive in closed form; per-phase
constants are fixed points of the gauged operator and its range has zero
weighted phase means; the jump decomposition recovers J exactly; each
planted mistake (orientation sign, theta <-> 1 - theta, unweighted means,
J with psi in place of 1 - psi) fails one of them.
"""

import numpy as np
import pytest

from oracle.fccd import OracleFCCDPPE


def _two_cell(cls=OracleFCCDPPE):
    # unit faces, as the reference's two-cell grid (N = (2, 2), L = (2, 2))
    return cls((8, 8), (8.0, 8.0), "wall", coefficient_scheme="phase_separated",
               interface_coupling="affine_jump")


def test_signed_jump_orientation_reference_example(cls=OracleFCCDPPE):
    """test_signed_pressure_jump_gradient_orientation: a given nodal jump
    j = -3 * 2 gives -6 on the liquid -> gas face, +6 reversed."""
    o = _two_cell(cls)
    psi = np.ones(o.shape)
    psi[1:, :] = 0.0
    kappa = np.full(o.shape, 2.0)
    o.set_interface_jump_context(psi, pressure_jump=-3.0 * kappa)
    B = o.signed_jump_gradient(0)
    np.testing.assert_allclose(B[0, :], -6.0)
    np.testing.assert_allclose(B[1:, :], 0.0)
    o.set_interface_jump_context(1.0 - psi, pressure_jump=-3.0 * kappa)
    np.testing.assert_allclose(o.signed_jump_gradient(0)[0, :], 6.0)


def test_cut_face_resistance_reference_example(cls=OracleFCCDPPE):
    """test_affine_jump_cut_face_coefficient_uses_phase_resistance."""
    o = _two_cell(cls)
    psi = np.ones(o.shape)
    psi[0, :] = 0.25
    kappa = np.full(o.shape, 2.0)
    rho = np.full(o.shape, 1000.0)
    rho[psi < 0.5] = 1.2
    o.set_interface_jump_context(psi, kappa, 3.0)
    o.prepare_operator(rho)
    theta = (0.5 - 0.25) / (1.0 - 0.25)
    want = 1.0 / (theta * 1.2 + (1.0 - theta) * 1000.0)
    np.testing.assert_allclose(o.face_coeff(0)[0, :], want, rtol=1e-15)
    np.testing.assert_allclose(o.face_coeff(0)[1:, :], 2.0 / 2000.0, rtol=1e-15)
    assert not np.allclose(want, 2.0 / (1.2 + 1000.0))
    assert o.threshold is None and not o.uses_mean_gauge()


def test_neutral_context_and_missing_context():
    o = _two_cell()
    psi = np.ones(o.shape)
    psi[0, :] = 0.25
    rho = np.where(psi < 0.5, 1.2, 1000.0)
    with pytest.raises(RuntimeError, match="interface-stress context"):
        o.prepare_operator(rho)
    o.set_interface_jump_context(psi, np.zeros(o.shape), 0.0)
    o.prepare_operator(rho)
    rhs = np.ones(o.shape)
    np.testing.assert_allclose(o.affine_rhs(), 0.0)
    b = o.system_rhs(rhs)
    assert b.ravel()[o.pin] == 0.0
    np.testing.assert_allclose(np.delete(b.ravel(), o.pin), 1.0)


def test_affine_drive_of_a_straight_interface(cls=OracleFCCDPPE):
    """psi depends on x only and crosses 0.5 between nodes i and i+1: the
    drive is alpha_c B / w at node i and -alpha_c B / w at node i+1 of every
    row, B = -sigma kappa_f / H (liquid below, gas above), kappa_f the
    linear interpolant of kappa at the cut point, alpha_c the cut-face
    resistance, w = H on interior nodes."""
    N, L = (12, 10), (1.2, 1.0)
    o = cls(N, L, "wall", coefficient_scheme="phase_separated", interface_coupling="affine_jump")
    x = np.linspace(0.0, L[0], N[0] + 1)
    psi1 = 1.0 / (1.0 + np.exp((x - 0.53) / 0.05))  # liquid (psi ~ 1) at small x
    psi = np.repeat(psi1[:, None], N[1] + 1, axis=1)
    y = np.linspace(0.0, L[1], N[1] + 1)
    kappa = 1.5 + 0.3 * np.add.outer(x, y)
    rho = 1.0 + 999.0 * psi
    sigma = 0.7
    o.set_interface_jump_context(psi, kappa, sigma)
    o.prepare_operator(rho)
    i = int(np.nonzero((psi1[:-1] >= 0.5) & (psi1[1:] < 0.5))[0][0])
    theta = (0.5 - psi1[i]) / (psi1[i + 1] - psi1[i])
    alpha = 1.0 / (theta * rho[i, 0] + (1.0 - theta) * rho[i + 1, 0])
    H = L[0] / N[0]
    kf = (1.0 - theta) * kappa[i, :] + theta * kappa[i + 1, :]
    B = -sigma * kf / H  # orientation +1: gas above the face
    want = np.zeros(o.shape)
    want[i, :] = alpha * B / H
    want[i + 1, :] = -alpha * B / H
    np.testing.assert_allclose(o.affine_rhs(), want, rtol=1e-13, atol=1e-12)


def _phase_case(bc, N=(10, 8), cls=OracleFCCDPPE):
    o = cls(N, (1.0, 0.8), bc, coefficient_scheme="phase_separated",
            interface_coupling="none", preconditioner="jacobi")
    X, Y = np.meshgrid(np.linspace(0, 1.0, N[0] + 1), np.linspace(0, 0.8, N[1] + 1),
                       indexing="ij")
    r = np.sqrt((X - 0.45) ** 2 + (Y - 0.4) ** 2)
    psi = 1.0 / (1.0 + np.exp(-(r - 0.25) / 0.04))
    if bc == "periodic":
        psi[-1, :] = psi[0, :]
        psi[:, -1] = psi[:, 0]
    rho = 1e-3 + (1.0 - 1e-3) * psi
    o.prepare_operator(rho)
    return o, psi, rho


@pytest.mark.parametrize("bc", ["wall", "periodic"])
def test_mean_gauge_fixed_points_and_range(bc, cls=OracleFCCDPPE):
    o, psi, rho = _phase_case(bc, cls=cls)
    assert o.uses_mean_gauge()
    assert o.threshold == pytest.approx(0.5 * (rho.min() + rho.max()))
    gas = (rho < o.threshold).astype(float)
    for c in (gas, 1.0 - gas, 2.0 * gas - 3.0 * (1.0 - gas)):
        want = c.copy()
        if bc == "periodic":  # image rows are p_img - p_src
            c = o._sync_images(c)
            want = c.copy()
            want.ravel()[o.images] = 0.0
        np.testing.assert_allclose(o.apply(c), want, atol=1e-12)
    p = np.random.default_rng(1).standard_normal(o.shape)
    p = o._sync_images(o.project_means(p))
    out = o.apply(p)
    vol = o.cell_volume()
    for w in (gas * vol, (1.0 - gas) * vol):
        assert abs(np.sum(w * out)) <= 1e-9 * np.sum(np.abs(w * out))
    # no face couples the phases
    for a in range(2):
        c = o.face_coeff(a)
        lo = np.moveaxis(rho, a, 0)[:-1] >= o.threshold
        hi = np.moveaxis(rho, a, 0)[1:] >= o.threshold
        assert np.all(c[lo != hi] == 0.0) and np.all(c[lo == hi] > 0.0)


def test_periodic_image_plane_carries_no_measure():
    """Planted: equal weights on the image plane change the means."""
    o, psi, rho = _phase_case("periodic")
    f = np.random.default_rng(3).standard_normal(o.shape)
    good = o.project_means(f)
    w_bad = tuple(np.where(w > 0, 1.0, 0.0) + 0.0 for w in o.weights)
    gas = o.gas
    m_bad = [np.sum(np.where(m, 1.0, 0.0) * f) / np.sum(m) for m in (gas, ~gas)]
    bad = f - (gas * m_bad[0] + (~gas) * m_bad[1])
    assert np.abs(good - bad).max() > 1e-6
    vol = o.cell_volume()
    assert np.all(vol[-1, :] == 0.0) and np.all(vol[:, -1] == 0.0)
    assert np.allclose(vol[:-1, :-1], (1.0 / 10) * (0.8 / 8))


def test_uniform_density_falls_back_to_the_pin():
    o = OracleFCCDPPE((8, 8), (1.0, 1.0), "wall", coefficient_scheme="phase_separated")
    o.prepare_operator(np.full(o.shape, 3.0))
    assert o.threshold is None and not o.uses_mean_gauge()
    np.testing.assert_allclose(o.face_coeff(0), 2.0 / 6.0)


@pytest.mark.parametrize("bc", ["wall", "periodic"])
def test_jump_decomposition_recovers_J(bc, cls=OracleFCCDPPE):
    """rhs = L(J): the solve's rhs - L(J) is zero, so p = gauge(0) + J = J."""
    N = (10, 8)
    o = cls(N, (1.0, 0.8), bc, coefficient_scheme="phase_separated",
            interface_coupling="jump_decomposition")
    _, psi, rho = _phase_case(bc, N)
    kappa = 1.0 + 0.5 * psi
    o.set_interface_jump_context(psi, kappa, 0.3)
    o.prepare_operator(rho)
    J = -0.3 * kappa * (1.0 - psi)
    np.testing.assert_allclose(o.jump_pressure(), J)
    rhs = o.core(J)
    p = o.solve_direct(rhs, rho)
    want = o._sync_images(J) if bc == "periodic" else J
    np.testing.assert_allclose(p, want, atol=1e-12)


@pytest.mark.parametrize("coupling", ["none", "jump_decomposition", "affine_jump"])
@pytest.mark.parametrize("bc", ["wall", "periodic"])
def test_gmres_matches_dense_direct(coupling, bc):
    N = (10, 8)
    o = OracleFCCDPPE(N, (1.0, 0.8), bc, coefficient_scheme="phase_separated",
                      interface_coupling=coupling, preconditioner="jacobi")
    _, psi, rho = _phase_case(bc, N)
    o.set_interface_jump_context(psi, 2.0 + psi, 0.05)
    rhs = np.random.default_rng(7).standard_normal(o.shape)
    want = o.solve_direct(rhs, rho)
    p, info = o.solve(rhs, rho, tol=1e-13, maxiter=400, restart=80)
    assert info["info"] == 0
    np.testing.assert_allclose(p, want, atol=1e-8 * np.abs(want).max())


# ---- planted mistakes: each must fail a check above
class _FlippedOrientation(OracleFCCDPPE):
    def signed_jump_gradient(self, axis):
        return -super().signed_jump_gradient(axis)


class _ThetaSwapped(OracleFCCDPPE):
    def _build_face_coeff(self, axis):  # theta <-> 1 - theta in the resistance
        r = np.moveaxis(self.rho, axis, 0)
        ps = np.moveaxis(self.stress_ctx["psi"], axis, 0)
        t = self.stress_ctx["threshold"]
        cut = (ps[:-1] < t) != (ps[1:] < t)
        theta = np.where(cut, (t - ps[:-1]) / np.where(cut, ps[1:] - ps[:-1], 1.0), 0.0)
        c = np.where(cut, 1.0 / ((1.0 - theta) * r[:-1] + theta * r[1:]),
                     2.0 / (r[:-1] + r[1:]))
        return np.moveaxis(c, 0, axis)


class _UnweightedMeans(OracleFCCDPPE):
    def project_means(self, arr):
        arr = np.asarray(arr, dtype=np.float64)
        m = [np.sum(np.where(g, arr, 0.0)) / np.sum(g) for g in (self.gas, ~self.gas)]
        return arr - (self.gas * m[0] + (~self.gas) * m[1])


class _WrongJ(OracleFCCDPPE):
    def jump_pressure(self):
        return self.jump_ctx["jump"] * self.jump_ctx["psi"]


def test_planted_orientation_fails():
    with pytest.raises(AssertionError):
        test_signed_jump_orientation_reference_example(_FlippedOrientation)
    with pytest.raises(AssertionError):
        test_affine_drive_of_a_straight_interface(_FlippedOrientation)


def test_planted_theta_fails():
    with pytest.raises(AssertionError):
        test_cut_face_resistance_reference_example(_ThetaSwapped)


def test_planted_unweighted_means_fail():
    with pytest.raises(AssertionError):
        test_mean_gauge_fixed_points_and_range("periodic", _UnweightedMeans)


def test_planted_jump_fails():
    with pytest.raises(AssertionError):
        test_jump_decomposition_recovers_J("wall", _WrongJ)
