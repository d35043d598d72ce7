"""DissipativeCCDAdvection: level-set advection by dissipative CCD + TVD-RK3.

Mirrors ``DissipativeCCDAdvection`` (levelset/advection_dccd.py:20-152) on
uniform grids: per TVD-RK3 stage (time_integration/tvd_rk3.py:26-48) the
fluxes psi u_a (``tpf_dccd_flux``), their CCD first derivatives (this
package's ``CCDSolver``, one launch per axis), then one fused kernel
(``tpf_dccd_stage``) for the adaptive filter fp + eps_d (2 psi - 1)^2 (fp_{+1}
- 2 fp + fp_{-1}) with unfiltered non-periodic end nodes, the sum over axes,
the stage combination, the clip and the periodic image copies; optionally
the interface-weighted mass correction (levelset/heaviside.py:118-149,
``tpf_dccd_mass_correct``).  As in the reference, ``advance`` first
synchronises the periodic images of the psi it is given, in place.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as nat
from .device import Handoff, ptr, require_gpu, stream_handle
from .grid import require_uniform

EPS_D_ADV = 0.05  # levelset/advection_kernels.py:9


class DissipativeCCDAdvection:
    """Advects psi using dissipative CCD + TVD-RK3."""

    scheme_names = ("dissipative_ccd",)
    _scheme_aliases = {"dccd": "dissipative_ccd"}

    def __init__(self, backend, grid, ccd, bc="periodic", eps_d: float = EPS_D_ADV,
                 mass_correction: bool = False):
        require_uniform(grid, "DissipativeCCDAdvection")
        if not getattr(grid, "uniform", True):
            raise ValueError("this build's DissipativeCCDAdvection needs a uniform grid")
        if not hasattr(ccd, "differentiate_into"):
            raise TypeError("DissipativeCCDAdvection needs this package's CCDSolver")
        self.backend = backend
        self._grid = grid
        self._ccd = ccd
        self.ndim = grid.ndim
        if isinstance(bc, (tuple, list)):
            self._bc = tuple(str(v).strip().lower() for v in bc)
        else:
            self._bc = str(bc).strip().lower()
        self._periodic = [self._axis_bc(a) == "periodic" for a in range(self.ndim)]
        self._eps_d = float(eps_d)
        self._mass_correction = bool(mass_correction)
        self.shape = tuple(int(n) + 1 for n in grid.N)
        self._lib = nat.lib()

    def _axis_bc(self, axis: int) -> str:
        return self._bc[axis] if isinstance(self._bc, tuple) else self._bc

    def _sync_host(self, arr: np.ndarray) -> None:
        """core/boundary.py:116-133 on a NumPy array, in place."""
        for a, per in enumerate(self._periodic):
            if per:
                src = [slice(None)] * arr.ndim
                img = [slice(None)] * arr.ndim
                src[a], img[a] = 0, -1
                arr[tuple(img)] = arr[tuple(src)]

    def advance(self, psi, velocity_components, dt: float, clip_bounds=(0.0, 1.0)):
        require_gpu()
        if len(velocity_components) != self.ndim:
            raise ValueError(f"velocity_components needs {self.ndim} arrays")
        if isinstance(psi, np.ndarray) and psi.dtype == np.float64:
            self._sync_host(psi)  # the reference synchronises its input in place
        io = Handoff(psi)
        q0 = io.to_device(psi)
        if tuple(q0.shape) != self.shape:
            raise ValueError(f"psi shape {tuple(q0.shape)} != grid.shape {self.shape}")
        us = [io.to_device(u) for u in velocity_components]
        for u in us:
            if tuple(u.shape) != self.shape:
                raise ValueError(f"velocity shape {tuple(u.shape)} != grid.shape {self.shape}")
        dims = nat.i64arr(self.shape)
        per = nat.iarr([1 if p else 0 for p in self._periodic])
        st = stream_handle()
        lib = self._lib
        nat.check(lib.tpf_sync_periodic_images(self.ndim, dims, per, ptr(q0), ptr(q0), st))
        m_old = None
        if self._mass_correction:
            m = C.c_double()
            bcs = nat.iarr([nat.BC_PERIODIC if p else nat.BC_WALL for p in self._periodic])
            nat.check(lib.tpf_fields_mass(self.ndim, dims, nat.darr(self._grid.L), bcs, ptr(q0),
                                          C.byref(m), st))
            m_old = m.value
        f = torch.empty((self.ndim,) + self.shape, dtype=torch.float64, device="cuda")
        fp = torch.empty_like(f)
        vel = (C.c_void_p * self.ndim)(*[ptr(u) for u in us])
        clip = clip_bounds is not None
        lo, hi = (float(clip_bounds[0]), float(clip_bounds[1])) if clip else (0.0, 0.0)

        def stage(q, ca, cb):
            nat.check(lib.tpf_dccd_flux(self.ndim, dims, ptr(q), vel, ptr(f), st))
            if hasattr(self._ccd, "apply_fields_into"):  # one launch for every axis
                self._ccd.apply_fields_into(list(f), list(fp), None, (1 << self.ndim) - 1)
            else:
                for a in range(self.ndim):
                    self._ccd.differentiate_into(f[a], fp[a], None, a)
            out = torch.empty_like(q0)
            nat.check(lib.tpf_dccd_stage(self.ndim, dims, per, ptr(q0), ptr(q), ptr(fp),
                                         self._eps_d, float(dt), ca, cb, 1 if clip else 0,
                                         lo, hi, ptr(out), st))
            return out

        # tvd_rk3.py:43-48
        q1 = stage(q0, 0.0, 1.0)
        q2 = stage(q1, 0.75, 0.25)
        qn = stage(q2, 1.0 / 3.0, 2.0 / 3.0)
        if self._mass_correction:
            out = torch.empty_like(qn)
            nat.check(lib.tpf_dccd_mass_correct(self.ndim, dims, nat.darr(self._grid.L), per,
                                                ptr(qn), m_old, ptr(out), None, st))
            qn = out
        return io.back(qn)
