"""SplitReinitializer: operator-split CLS reinitialisation on the device.

Mirrors ``SplitReinitializer`` (levelset/reinit_split.py:22-102): n_steps
pseudo-time steps of DCCD-filtered compression d/dtau psi = -div[psi (1-psi)
n] and CN-ADI diffusion eps laplacian psi, clipped to [0, 1] after each, then
the interface-weighted mass correction.  The whole loop runs natively in
``tpf_reinit_split`` (libtpf) on this package's CCD plan.  NumPy in, NumPy
out; CUDA tensors in, CUDA tensors out.
"""

from __future__ import annotations

import torch

from . import _native as nat
from .device import Handoff, ptr, require_gpu, stream_handle
from .grid import require_uniform

_BC = {"zero": 0, "periodic": 1, "neumann": 2, "outflow": 3}


def compute_dtau(grid, eps: float) -> float:
    """Pseudo-time step min(dx^2 / (4 ndim eps), dx / 2) (reinit_ops.py:36-51)."""
    dx = min(float(grid.L[a]) / int(grid.N[a]) for a in range(grid.ndim))
    return min(0.5 * dx ** 2 / (2.0 * grid.ndim * float(eps)), 0.5 * dx)


class SplitReinitializer:
    def __init__(self, backend, grid, ccd, eps: float, n_steps: int = 4, bc="zero",
                 eps_d_comp: float = 0.05, mass_correction: bool = True):
        require_uniform(grid, "SplitReinitializer")
        if not hasattr(ccd, "_plan"):
            raise TypeError("SplitReinitializer needs this package's CCDSolver")
        if not getattr(grid, "uniform", True):
            raise ValueError("this build's SplitReinitializer needs a uniform grid")
        self.backend = backend
        self.grid = grid
        self.ccd = ccd
        self.eps = float(eps)
        self.n_steps = int(n_steps)
        self._bc = bc
        names = [str(b).strip().lower() for b in bc] if isinstance(bc, (tuple, list)) \
            else [str(bc).strip().lower()] * grid.ndim
        bad = [b for b in names if b not in _BC]
        if bad:
            raise ValueError(f"unsupported level-set boundary {bad[0]!r}")
        self._bc_codes = [_BC[b] for b in names]
        self._eps_d_comp = float(eps_d_comp)
        self._mass_correction = bool(mass_correction)
        self.dtau = compute_dtau(grid, self.eps)
        self.shape = tuple(int(n) + 1 for n in grid.N)
        self._lib = nat.lib()

    def reinitialize(self, psi):
        require_gpu()
        io = Handoff(psi)
        q = io.to_device(psi)
        if tuple(q.shape) != self.shape:
            raise ValueError(f"psi shape {tuple(q.shape)} != grid.shape {self.shape}")
        out = torch.empty_like(q)
        nat.check(self._lib.tpf_reinit_split(
            self.ccd._plan, self.grid.ndim, nat.i64arr(self.shape), nat.darr(self.grid.L),
            nat.iarr(self._bc_codes), self.eps, self.n_steps, self._eps_d_comp,
            1 if self._mass_correction else 0, ptr(q), ptr(out), stream_handle()))
        return io.back(out)
