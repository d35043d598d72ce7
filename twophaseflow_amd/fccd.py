"""FCCDSolver: face-centred compact differences on the device.

Mirrors ``FCCDSolver`` (ccd/fccd.py:165-624) on uniform grids: the face
gradient, face value, face second derivative and face jet from the nodal CCD
second derivative, the R4 Hermite node gradient, the upwind Taylor face
state, the face-flux divergence, the wall-option hooks and the periodic
symbol.  The nodal CCD second derivative comes from this package's
``CCDSolver``; the face stencils run in ``tpf_fccd_face_op`` (libtpf).
``advection_rhs`` belongs to the physics step and raises
``NotImplementedError``.

Inputs are NumPy arrays (copied) or CUDA torch tensors (no copy); results
come back in the same form.  Face arrays have ``grid.N[axis]`` entries along
``axis`` (face j between nodes j and j+1; periodic axes pair node N-1 with
node 0).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _native as nat
from .boundary import axis_kinds
from .ccd import CCDSolver
from .device import Handoff, ptr, require_gpu, stream_handle
from .grid import require_uniform


@dataclass(frozen=True)
class FCCDFaceJet:
    """Face-located compact jet (value, gradient, curvature), ccd/fccd.py:71-84."""

    value: object
    gradient: object
    curvature: object


_OPS = {"gradient": 0, "value": 1, "curvature": 2, "node_gradient": 3, "divergence": 4,
        "upwind": 5}


class FCCDSolver:
    """Face-centred compact-difference differentiator (ccd/fccd.py:138-195)."""

    def __init__(self, grid, backend=None, bc_type: str = "wall", ccd_solver=None):
        require_uniform(grid, "FCCDSolver")
        self.grid = grid
        self.backend = backend
        self.ndim = grid.ndim
        self.bc_type = bc_type
        if ccd_solver is not None:
            if axis_kinds(getattr(ccd_solver, "bc_type", bc_type), self.ndim) != \
                    axis_kinds(bc_type, self.ndim):
                raise ValueError("FCCD and CCD solvers must share the same boundary topology")
            self._ccd = ccd_solver
        else:
            self._ccd = CCDSolver(grid, backend, bc_type)
        if not getattr(grid, "uniform", True):
            raise ValueError("this build's FCCDSolver needs a uniform grid")
        self._kinds = axis_kinds(bc_type, self.ndim)
        self.shape = tuple(int(n) + 1 for n in grid.N)
        self._lib = nat.lib()

    # -- helpers -------------------------------------------------------------------
    def _axis_periodic(self, axis: int) -> bool:
        return self._kinds[axis] == "periodic"

    def _face_shape(self, axis):
        s = list(self.shape)
        s[axis] = int(self.grid.N[axis])
        return tuple(s)

    def _op(self, op, axis, u, q=None, d=None, vel=None, out_shape=None):
        if not 0 <= axis < self.ndim:
            raise ValueError(f"axis {axis} out of range for ndim={self.ndim}")
        require_gpu()
        out = torch.empty(out_shape, dtype=torch.float64, device="cuda")
        bc = [nat.BC_PERIODIC if k == "periodic" else nat.BC_WALL for k in self._kinds]
        nat.check(self._lib.tpf_fccd_face_op(
            _OPS[op], self.ndim, nat.i64arr(self.shape), nat.darr(self.grid.L), nat.iarr(bc),
            axis, ptr(u), ptr(q), ptr(d), ptr(vel), ptr(out), stream_handle()))
        return out

    def _nodal(self, io, x, name):
        t = io.to_device(x)
        if tuple(t.shape) != self.shape:
            raise ValueError(f"{name} shape {tuple(t.shape)} != grid.shape {self.shape}")
        return t

    def _q(self, io, u, axis, q):
        if q is None:
            return self._ccd.second_derivative(u, axis)
        return self._nodal(io, q, "q")

    # -- public primitives ---------------------------------------------------------
    def face_gradient(self, u, axis: int, q=None):
        """(u_hi - u_lo)/H - (H/24)(q_hi - q_lo) (ccd/fccd.py:243-277)."""
        io = Handoff(u)
        t = self._nodal(io, u, "u")
        return io.back(self._op("gradient", axis, t, self._q(io, t, axis, q),
                                out_shape=self._face_shape(axis)))

    def face_value(self, u, axis: int, q=None):
        """(u_lo + u_hi)/2 - (H^2/16)(q_lo + q_hi) (ccd/fccd.py:279-299)."""
        io = Handoff(u)
        t = self._nodal(io, u, "u")
        return io.back(self._op("value", axis, t, self._q(io, t, axis, q),
                                out_shape=self._face_shape(axis)))

    def face_second_derivative(self, u, axis: int, q=None):
        """(q_lo + q_hi)/2 (ccd/fccd.py:301-317)."""
        io = Handoff(u)
        t = self._nodal(io, u, "u")
        return io.back(self._op("curvature", axis, t, self._q(io, t, axis, q),
                                out_shape=self._face_shape(axis)))

    def face_jet(self, u, axis: int, q=None) -> FCCDFaceJet:
        """The three face members sharing one q (ccd/fccd.py:319-331)."""
        io = Handoff(u)
        t = self._nodal(io, u, "u")
        qq = self._q(io, t, axis, q)
        fs = self._face_shape(axis)
        return FCCDFaceJet(value=io.back(self._op("value", axis, t, qq, out_shape=fs)),
                           gradient=io.back(self._op("gradient", axis, t, qq, out_shape=fs)),
                           curvature=io.back(self._op("curvature", axis, t, qq, out_shape=fs)))

    def upwind_face_value(self, u, advective_velocity_face, axis: int, nodal_gradient=None,
                          q=None):
        """Taylor state from the upwind node: u_lo + (H/2)u'_lo + (H^2/8)u''_lo
        for velocity >= 0, else u_hi - (H/2)u'_hi + (H^2/8)u''_hi
        (ccd/fccd.py:333-390)."""
        io = Handoff(u)
        t = self._nodal(io, u, "u")
        if nodal_gradient is None and q is None:
            d1, qq = self._ccd.differentiate(t, axis)
        else:
            d1 = (self._ccd.first_derivative(t, axis) if nodal_gradient is None
                  else self._nodal(io, nodal_gradient, "nodal_gradient"))
            qq = self._q(io, t, axis, q)
        fs = self._face_shape(axis)
        vel = io.to_device(advective_velocity_face)
        if tuple(vel.shape) != fs:
            raise ValueError(f"face velocity shape {tuple(vel.shape)} != {fs}")
        return io.back(self._op("upwind", axis, t, qq, d1, vel, out_shape=fs))

    def node_gradient(self, u, axis: int, q=None):
        """R4 Hermite node gradient 0.5(d_L + d_R) - (H/16)(q_{i+1} - q_{i-1});
        one-sided face gradients at wall nodes (ccd/fccd.py:392-462)."""
        io = Handoff(u)
        t = self._nodal(io, u, "u")
        return io.back(self._op("node_gradient", axis, t, self._q(io, t, axis, q),
                                out_shape=self.shape))

    def face_divergence(self, F_face, axis: int):
        """(F[i] - F[i-1]) / H at nodes; periodic wraps and repeats node 0 at
        node N; wall end nodes are 0 (ccd/fccd.py:464-521)."""
        io = Handoff(F_face)
        f = io.to_device(F_face)
        if tuple(f.shape) != self._face_shape(axis):
            raise ValueError(f"face array shape {tuple(f.shape)} != {self._face_shape(axis)}")
        return io.back(self._op("divergence", axis, f, out_shape=self.shape))

    def advection_rhs(self, *args, **kwargs):
        raise NotImplementedError("FCCD advection belongs to the physics step, which this "
                                  "build does not provide")

    def enforce_wall_option_iii(self, face_array, axis: int):
        """No-op on interior-only face arrays (ccd/fccd.py:579-592)."""
        return face_array

    def enforce_wall_option_iv(self, face_array, axis: int, wall_value: float = 0.0):
        """No-op hook, as in the reference (ccd/fccd.py:594-612)."""
        return face_array

    def periodic_symbol(self, omega: float, axis: int = 0) -> complex:
        """i w (1 - 7 (w H)^4 / 5760), the truncated face-operator symbol
        (ccd/fccd.py:616-624, fccd_helpers.py:65-67)."""
        H = float(self.grid.L[axis]) / int(self.grid.N[axis])
        return 1j * omega * (1.0 - 7.0 * (omega * H) ** 4 / 5760.0)
