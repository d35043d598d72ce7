"""Host logic of PPESolverDefectCorrection (ppe/defect_correction.py:28-439)
driven by oracle components on the CPU: L_H = the FCCD PPE oracle, L_L = the
FVM oracle's direct solve.  The loop must reach the dense direct solution of
L_H (the defect-correction fixed point), its residual history must fall, and
the reference's argument checks, same-operator refusal and
fail_on_nonconvergence must hold."""

import numpy as np
import pytest

from oracle import ppe as oppe
from oracle.fccd import OracleFCCDPPE
from twophaseflow_amd.defect_correction import PPESolverDefectCorrection


class _Op:
    def __init__(self, N):
        self.o = OracleFCCDPPE(N)
        self._pin_dof = self.o.pin

    def prepare_operator(self, rho):
        self.o.prepare_operator(rho)

    def apply(self, p):
        return self.o.apply(p)


class _Base:
    def __init__(self, N):
        self.o = oppe.OraclePPE(N)
        self._pin_dof = self.o.pin

    def prepare_operator(self, rho):
        pass

    def solve(self, rhs, rho, dt=0.0, p_init=None):
        return self.o.solve_direct(rhs, rho)


def _case(N, ratio, seed):
    x = np.linspace(0, 1, N[0] + 1)[:, None]
    y = np.linspace(0, 1, N[1] + 1)[None, :]
    r = np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2)
    psi = 1.0 / (1.0 + np.exp(np.clip((0.25 - r) / (1.5 / N[0]), -500, 500)))
    rho = 1.0 / ratio + (1.0 - 1.0 / ratio) * psi
    rhs = np.random.default_rng(seed).standard_normal(rho.shape)
    return rho, rhs


@pytest.mark.parametrize("ratio", [10.0, 1000.0])
def test_defect_correction_reaches_the_high_order_solution(ratio):
    N = (24, 20)
    rho, rhs = _case(N, ratio, 1)
    dc = PPESolverDefectCorrection(None, None, _Base(N), _Op(N), max_corrections=60,
                                   tolerance=1e-10)
    p = dc.solve(rhs, rho)
    h = dc.last_residual_history
    assert all(b < a for a, b in zip(h, h[1:]))
    assert dc.last_diagnostics["ppe_dc_converged"] == 1.0 and not dc.last_stalled
    o = OracleFCCDPPE(N)
    o.prepare_operator(rho)
    x = np.linalg.solve(o.matrix(), o.system_rhs(rhs).ravel()).reshape(o.shape)
    assert np.abs(p - x).max() <= 1e-9 * np.abs(x).max()
    assert p.ravel()[o.pin] == 0.0


def test_defect_correction_checks_and_stall():
    N = (12, 10)
    rho, rhs = _case(N, 100.0, 2)
    with pytest.raises(ValueError):
        PPESolverDefectCorrection(None, None, _Base(N), _Op(N), max_corrections=0)
    with pytest.raises(ValueError):
        PPESolverDefectCorrection(None, None, _Base(N), _Op(N), tolerance=0.0)
    with pytest.raises(TypeError):
        PPESolverDefectCorrection(None, None, _Base(N), object())
    class _Solving(_Op):
        def solve(self, rhs, rho, dt=0.0, p_init=None):
            return rhs

    with pytest.raises(ValueError):  # the same operator as its own base solve
        PPESolverDefectCorrection(None, None, _Solving(N), _Solving(N))
    dc = PPESolverDefectCorrection(None, None, _Base(N), _Op(N), max_corrections=1,
                                   tolerance=1e-12)
    dc.solve(rhs, rho)
    assert dc.last_stalled and dc.last_diagnostics["ppe_dc_iterations"] == 1.0
    assert len(dc.last_residual_history) == 2
    dc = PPESolverDefectCorrection(None, None, _Base(N), _Op(N), max_corrections=1,
                                   tolerance=1e-12, fail_on_nonconvergence=True)
    with pytest.raises(RuntimeError, match="did not converge"):
        dc.solve(rhs, rho)
