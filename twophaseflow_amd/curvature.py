"""CurvatureCalculatorPsi: interface curvature straight from psi, on the device.

Mirrors ``CurvatureCalculatorPsi`` (levelset/curvature_psi.py:45-196): the CCD
first and second derivatives of psi (one all-axes launch), the optional DCCD
filter with switch (2 psi - 1)^2 (``tpf_dccd_filter_nd``), the mixed
derivatives by sequential CCD (2-D: psi_xy from the filtered psi_x, then
filtered; 3-D: psi_ij from the filtered psi_i), the 2-D / 3-D curvature with
the far-field mask (``tpf_curvature_psi``) and the optional 3x3 Gaussian
(2-D, ``tpf_gaussian_3x3``).  NumPy in, NumPy out; CUDA tensors in, CUDA
tensors out.
"""

from __future__ import annotations

import ctypes as C

import torch

from . import _native as nat
from .boundary import axis_kinds
from .device import Handoff, ptr, require_gpu, stream_handle
from .grid import require_uniform

DCCD_EPS_D = 0.05  # levelset/curvature.py:43


class CurvatureCalculatorPsi:
    def __init__(self, backend, ccd, psi_min: float = 0.01, gaussian_filter: bool = False,
                 dccd_eps: float = DCCD_EPS_D):
        require_uniform(getattr(ccd, "grid", None), "CurvatureCalculatorPsi")
        if not hasattr(ccd, "apply_into"):
            raise TypeError("CurvatureCalculatorPsi needs this package's CCDSolver")
        self.backend = backend
        self.ccd = ccd
        self.psi_min = float(psi_min)
        self.gaussian_filter = bool(gaussian_filter)
        self.dccd_eps = float(dccd_eps)
        self._lib = nat.lib()

    def compute(self, psi):
        require_gpu()
        ccd = self.ccd
        ndim = ccd.ndim
        if ndim not in (2, 3):
            raise ValueError("curvature from psi is defined for 2-D and 3-D grids")
        shape = tuple(int(n) + 1 for n in ccd.grid.N)
        io = Handoff(psi)
        p = io.to_device(psi)
        if tuple(p.shape) != shape:
            raise ValueError(f"psi shape {tuple(p.shape)} != grid.shape {shape}")
        dims = nat.i64arr(shape)
        per = nat.iarr([1 if k == "periodic" else 0 for k in axis_kinds(ccd.bc_type, ndim)])
        st = stream_handle()
        d1 = torch.empty((ndim,) + shape, dtype=torch.float64, device="cuda")
        d2 = torch.empty_like(d1)
        ccd.apply_into(p, d1, d2, (1 << ndim) - 1)
        scratch = torch.empty_like(p) if self.dccd_eps > 0 else None

        def filt(f, out=None):
            if self.dccd_eps <= 0:
                return f
            if out is None:
                out = torch.empty_like(f)
            nat.check(self._lib.tpf_dccd_filter_nd(ndim, dims, per, ptr(p), self.dccd_eps,
                                                   ptr(f), ptr(out), ptr(scratch), st))
            return out

        if self.dccd_eps > 0:
            # every first and second derivative through one batched filter
            # (one launch per pass), straight into stacked outputs
            d1f, d2f = torch.empty_like(d1), torch.empty_like(d2)
            scr = torch.empty((2 * ndim,) + shape, dtype=torch.float64, device="cuda")
            arr = C.c_void_p * (2 * ndim)
            fin = arr(*([ptr(d1[a]) for a in range(ndim)] + [ptr(d2[a]) for a in range(ndim)]))
            fout = arr(*([ptr(d1f[a]) for a in range(ndim)] + [ptr(d2f[a]) for a in range(ndim)]))
            fscr = arr(*[ptr(scr[k]) for k in range(2 * ndim)])
            nat.check(self._lib.tpf_dccd_filter_nd_multi(ndim, dims, per, ptr(p), self.dccd_eps,
                                                         2 * ndim, fin, fout, fscr, st))
            d1, d2 = d1f, d2f
        pairs = [(0, 1)] if ndim == 2 else [(0, 1), (0, 2), (1, 2)]
        mixed = torch.empty((len(pairs),) + shape, dtype=torch.float64, device="cuda")
        for k, (i, j) in enumerate(pairs):
            ccd.differentiate_into(d1[i].contiguous(), mixed[k], None, j)
        if ndim == 2:
            mixed[0] = filt(mixed[0].clone())
        kappa = torch.empty_like(p)
        nat.check(self._lib.tpf_curvature_psi(ndim, dims, ptr(p), ptr(d1), ptr(d2), ptr(mixed),
                                              self.psi_min, ptr(kappa), st))
        if self.gaussian_filter and ndim == 2:
            sm = torch.empty_like(kappa)
            nat.check(self._lib.tpf_gaussian_3x3(dims, ptr(kappa), ptr(sm), st))
            kappa = sm
        return io.back(kappa)
