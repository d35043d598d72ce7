"""RhieChowInterpolator: the Rhie-Chow face-velocity divergence on the device.

Mirrors ``RhieChowInterpolator`` (spatial/rhie_chow.py:40-284): the PPE
right-hand side of the projection step, with the optional balanced-force
bracket.  The cell gradients of p (and psi) come from this package's
``CCDSolver`` in one all-axes launch; the face fluxes and their divergence
run in ``tpf_rhie_chow_divergence`` (libtpf) without materialising faces.
Inputs are NumPy arrays (copied) or CUDA torch tensors; the result comes back
in the same form.
"""

from __future__ import annotations

import ctypes as C

import torch

from . import _native as nat
from .boundary import axis_kinds
from .device import Handoff, ptr, require_gpu, stream_handle
from .grid import require_uniform


class RhieChowInterpolator:
    """Face velocities and Rhie-Chow divergence (spatial/rhie_chow.py:40-60)."""

    def __init__(self, backend, grid, ccd, bc_type: str = "wall"):
        require_uniform(grid, "RhieChowInterpolator")
        self.backend = backend
        self.grid = grid
        self.ndim = grid.ndim
        self.ccd = ccd
        self.bc_type = bc_type
        self._kinds = axis_kinds(bc_type, self.ndim)
        self.shape = tuple(int(n) + 1 for n in grid.N)
        if not hasattr(ccd, "apply_into"):
            raise TypeError("RhieChowInterpolator needs this package's CCDSolver")
        self._lib = nat.lib()

    def _grad(self, f: torch.Tensor) -> torch.Tensor:
        d = torch.empty((self.ndim,) + self.shape, dtype=torch.float64, device="cuda")
        self.ccd.apply_into(f, d, None, (1 << self.ndim) - 1)
        return d

    def face_velocity_divergence(self, vel_star, p, rho, dt: float, kappa=None, psi=None,
                                 we: float = None):
        """div u_RC (rhie_chow.py:55-107); the balanced-force bracket is used
        when kappa, psi and we are all given."""
        require_gpu()
        if len(vel_star) != self.ndim:
            raise ValueError(f"vel_star needs {self.ndim} components")
        io = Handoff(vel_star[0])
        us = [io.to_device(u) for u in vel_star]
        pt, rt = io.to_device(p), io.to_device(rho)
        for t in us + [pt, rt]:
            if tuple(t.shape) != self.shape:
                raise ValueError(f"field shape {tuple(t.shape)} != grid.shape {self.shape}")
        dp = self._grad(pt)
        bf = kappa is not None and psi is not None and we is not None
        kt = st = dpsi = None
        if bf:
            kt, st = io.to_device(kappa), io.to_device(psi)
            dpsi = self._grad(st)
        out = torch.empty_like(pt)
        bc = [nat.BC_PERIODIC if k == "periodic" else nat.BC_WALL for k in self._kinds]
        vel = (C.c_void_p * self.ndim)(*[ptr(u) for u in us])
        nat.check(self._lib.tpf_rhie_chow_divergence(
            self.ndim, nat.i64arr(self.shape), nat.darr(self.grid.L), nat.iarr(bc), vel,
            ptr(pt), ptr(rt), ptr(dp), ptr(kt), ptr(st), ptr(dpsi),
            float(we) if bf else 0.0, float(dt), ptr(out), stream_handle()))
        return io.back(out)
