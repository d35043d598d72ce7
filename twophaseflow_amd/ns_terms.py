"""Momentum terms, predictor, CFL, boundary handler and field extension of the
legacy two-phase step, on the GPU.

Mirrors of the reference classes the legacy builder wires
(``simulation/legacy_builder_runtime.py:34-199``), with the same names,
arguments and results:

* ``UCCD6ConvectionTerm`` (``ns_terms/uccd6_convection.py:37-128``) ->
  ``tpf_ns_uccd6_convection``;
* ``ViscousTerm`` with the default ``ccd_bulk`` body and the Picard CN
  advance (``ns_terms/viscous.py:49-204``, ``ns_terms/viscous_spatial.py``,
  ``time_integration/cn_advance/picard_cn.py``) -> ``tpf_ns_viscous``;
* ``GravityTerm`` / ``SurfaceTensionTerm`` (``ns_terms/gravity.py``,
  ``ns_terms/surface_tension.py``) and the IPC gradient, fused in
  ``Predictor.compute`` (``time_integration/ab2_predictor.py:57-204``) ->
  ``tpf_ns_predictor_rhs`` / ``tpf_ns_predictor_final``;
* ``CFLCalculator`` (``time_integration/cfl.py:40-137``),
  ``BoundaryConditionHandler`` (``simulation/boundary_condition.py``),
  ``ClosestPointExtender`` / ``NullFieldExtender``
  (``levelset/closest_point_extender.py``, ``levelset/field_extender.py:
  205-224``) and ``VelocityCorrector`` (``coupling/velocity_corrector.py``).

Fields are CUDA float64 tensors of ``grid.shape``; NumPy inputs are copied to
the device and results come back as NumPy (the reference's drop-in path).
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Any, List, Optional

import numpy as np
import torch

from . import _native as nat
from .boundary import axis_kinds
from .device import Handoff, ptr, require_gpu, stream_handle
from .grid import require_uniform


def _empty_like_list(t: torch.Tensor, k: int) -> List[torch.Tensor]:
    return [torch.empty_like(t) for _ in range(k)]


class _Work:
    """Grow-only device workspace (doubles)."""

    def __init__(self):
        self.t: Optional[torch.Tensor] = None

    def get(self, n: int) -> torch.Tensor:
        if self.t is None or self.t.numel() < n:
            self.t = torch.empty(int(n), dtype=torch.float64, device="cuda")
        return self.t


@dataclass
class NSComputeContext:
    """ns_terms/context.py:5-20."""
    velocity: List[Any]
    ccd: Any
    rho: Any
    mu: Any
    kappa: Any = None
    psi: Any = None
    fccd: Any = None


@dataclass
class FlowState:
    """core/flow_state.py: the fields one predictor call reads."""
    velocity: List[Any]
    psi: Any
    rho: Any
    mu: Any
    kappa: Any
    pressure: Any


def _geom(ccd):
    shape = tuple(int(n) + 1 for n in ccd.grid.N)
    return ccd.grid.ndim, shape, nat.i64arr(shape), nat.darr(ccd.grid.L)


class UCCD6ConvectionTerm:
    """Skew CCD advection minus sigma max|u| h^7 D2^4 u
    (ns_terms/uccd6_convection.py:37-128)."""

    scheme_names = ("uccd6",)

    def __init__(self, backend, grid, ccd, sigma: float = 1.0e-3):
        require_uniform(grid, "UCCD6ConvectionTerm")
        if sigma <= 0.0:
            raise ValueError(f"sigma must be > 0, got {sigma}")
        self._ccd, self._grid, self._sigma = ccd, grid, float(sigma)
        self._work = _Work()

    def compute_into(self, vel: List[torch.Tensor], out: List[torch.Tensor]) -> None:
        ndim, shape, dims, L = _geom(self._ccd)
        lib = nat.lib()
        w = self._work.get(lib.tpf_ns_uccd6_work(ndim, dims))
        nat.check(lib.tpf_ns_uccd6_convection(
            self._ccd._plan, ndim, dims, L, nat.parr([ptr(u) for u in vel]), self._sigma,
            nat.parr([ptr(o) for o in out]), ptr(w), stream_handle()))

    def compute(self, ctx_or_velocity, ccd=None) -> List:
        vel = ctx_or_velocity if isinstance(ctx_or_velocity, list) else ctx_or_velocity.velocity
        require_gpu()
        io = Handoff(vel[0])
        vt = [io.to_device(u) for u in vel]
        out = _empty_like_list(vt[0], len(vt))
        self.compute_into(vt, out)
        return [io.back(o) for o in out]


class ViscousTerm:
    """(1/Re) div[mu (grad u + grad u^T)] / rho with the ``ccd_bulk`` body
    (ns_terms/viscous.py:49-204, viscous_spatial.py:156-169); the Picard CN
    advance lives in Predictor (picard_cn.py:44-96)."""

    def __init__(self, backend, Re: float, cn_viscous: bool = True,
                 spatial_scheme: str = "ccd_bulk", cn_advance=None):
        name = str(spatial_scheme).strip().lower()
        name = {"ccd": "ccd_bulk"}.get(name, name)
        if name != "ccd_bulk":
            raise NotImplementedError(
                f"viscous spatial scheme {spatial_scheme!r} is not built; this build has "
                "'ccd_bulk' (the reference default)")
        if cn_advance is not None:
            raise NotImplementedError("only the default Picard CN advance is built")
        self.Re = float(Re)
        self.cn_viscous = bool(cn_viscous)
        self.spatial_scheme = name
        self._work = _Work()

    def evaluate_into(self, vel, mu, rho, ccd, out) -> None:
        ndim, shape, dims, L = _geom(ccd)
        lib = nat.lib()
        w = self._work.get(lib.tpf_ns_viscous_work(ndim, dims))
        nat.check(lib.tpf_ns_viscous(ccd._plan, ndim, dims, L, nat.parr([ptr(u) for u in vel]),
                                     ptr(mu), ptr(rho), self.Re,
                                     nat.parr([ptr(o) for o in out]), ptr(w), stream_handle()))

    def _evaluate(self, vel, mu, rho, ccd, psi=None):
        if psi is not None:
            raise NotImplementedError("the interface-band ccd_bulk variant is not built")
        io = Handoff(vel[0])
        vt = [io.to_device(u) for u in vel]
        out = _empty_like_list(vt[0], len(vt))
        self.evaluate_into(vt, io.to_device(mu), io.to_device(rho), ccd, out)
        return [io.back(o) for o in out]

    def compute_explicit(self, velocity_components, mu, rho, ccd, psi=None):
        return self._evaluate(velocity_components, mu, rho, ccd, psi=psi)


class GravityTerm:
    """-rho / Fr^2 on the last axis (ns_terms/gravity.py:43-86); fused into
    the predictor's right-hand side kernel."""

    def __init__(self, backend, Fr: float, ndim: int):
        self.Fr, self.ndim, self.grav_axis = float(Fr), int(ndim), int(ndim) - 1


class SurfaceTensionTerm:
    """CSF kappa D1 psi / We (ns_terms/surface_tension.py:37-96); fused into
    the predictor's right-hand side kernel."""

    def __init__(self, backend, We: float):
        self.We = float(We)


class Predictor:
    """AB2 + IPC + Picard CN predictor (time_integration/ab2_predictor.py:
    57-204).  The first call is forward Euler; each call buffers its
    convection term for the next (the buffer is not part of the run loop's
    retry snapshot, as in the reference)."""

    def __init__(self, backend, config, ccd, convection=None, viscous=None, gravity=None,
                 surface_tension=None, use_gfm: bool = False):
        require_uniform(getattr(ccd, "grid", None), "Predictor")
        if use_gfm:
            raise NotImplementedError("surface_tension_model='gfm' is not built")
        num, fl = config.numerics, config.fluid
        self.config = config
        self.ccd = ccd
        self.convection = convection or UCCD6ConvectionTerm(
            backend, ccd.grid, ccd, sigma=getattr(num, "uccd6_sigma", 1.0e-3))
        if not isinstance(self.convection, UCCD6ConvectionTerm):
            raise NotImplementedError("this build's predictor runs UCCD6ConvectionTerm")
        if getattr(num, "cn_mode", "picard") != "picard":
            raise ValueError(f"Unknown cn_mode={num.cn_mode!r}; supported in this build: "
                             "'picard'")
        self.viscous = viscous or ViscousTerm(backend, fl.Re, num.cn_viscous)
        self.gravity = gravity or GravityTerm(backend, fl.Fr, config.grid.ndim)
        self.surface_tens = surface_tension or SurfaceTensionTerm(backend, fl.We)
        self.use_gfm = False
        self._conv_prev: Optional[List[torch.Tensor]] = None
        self._ab2_ready = False
        self._buf: dict = {}

    def _bufs(self, key, like, k):
        b = self._buf.get(key)
        if b is None or b[0].shape != like.shape:
            b = _empty_like_list(like, k)
            self._buf[key] = b
        return b

    def compute(self, state: FlowState, dt: float) -> List:
        ccd = self.ccd
        ndim, shape, dims, L = _geom(ccd)
        io = Handoff(state.velocity[0])
        vel = [io.to_device(u) for u in state.velocity]
        rho, mu = io.to_device(state.rho), io.to_device(state.mu)
        kappa, psi = io.to_device(state.kappa), io.to_device(state.psi)
        p = io.to_device(state.pressure)
        like = vel[0]
        conv = _empty_like_list(like, ndim)
        self.convection.compute_into(vel, conv)
        # CCD gradients of psi (CSF) and of p (IPC), stacked by axis
        dpsi = torch.empty((ndim,) + shape, dtype=torch.float64, device="cuda")
        dp = torch.empty_like(dpsi)
        mask = (1 << ndim) - 1
        ccd.apply_into(psi, dpsi, None, mask)
        ccd.apply_into(p, dp, None, mask)
        vn = self._bufs("vn", like, ndim)
        self.viscous.evaluate_into(vel, mu, rho, ccd, vn)
        rhs = self._bufs("rhs", like, ndim)
        up = self._bufs("up", like, ndim)
        per = nat.iarr([1 if k == "periodic" else 0 for k in axis_kinds(ccd.bc_type, ndim)])
        prev = (nat.parr([ptr(c) for c in self._conv_prev])
                if self._ab2_ready and self._conv_prev is not None else None)
        lib = nat.lib()
        st = stream_handle()
        nat.check(lib.tpf_ns_predictor_rhs(
            ndim, dims, per, nat.parr([ptr(u) for u in vel]), nat.parr([ptr(c) for c in conv]),
            prev, ptr(rho), ptr(kappa), ptr(dpsi), ptr(dp), self.gravity.Fr,
            self.surface_tens.We, float(dt), nat.parr([ptr(v) for v in vn]),
            nat.parr([ptr(r) for r in rhs]), nat.parr([ptr(u) for u in up]), st))
        if self.viscous.cn_viscous:
            vs = self._bufs("vs", like, ndim)
            self.viscous.evaluate_into(up, mu, rho, ccd, vs)
            ustar = _empty_like_list(like, ndim)
            nat.check(lib.tpf_ns_predictor_final(
                ndim, dims, per, nat.parr([ptr(u) for u in vel]),
                nat.parr([ptr(r) for r in rhs]), ptr(rho), nat.parr([ptr(v) for v in vn]),
                nat.parr([ptr(v) for v in vs]), float(dt), nat.parr([ptr(u) for u in ustar]),
                st))
        else:
            raise NotImplementedError("cn_viscous=False predictor is not built")
        self._conv_prev = conv
        self._ab2_ready = True
        return [io.back(u) for u in ustar]


class CFLCalculator:
    """time_integration/cfl.py:40-137: min of CFL h / sum_i max|u_i|, the
    capillary bound CFL sqrt((1 + rho_ratio) We h^3 / (2 pi)) and (explicit
    viscous only) CFL h^2 / (4 max(mu/rho)), floored at 1e-8."""

    def __init__(self, backend, grid, cfl_number: float, We: Optional[float] = None,
                 rho_ratio: float = 0.0, cn_viscous: bool = True):
        require_uniform(grid, "CFLCalculator")
        self.grid = grid
        self.cfl = float(cfl_number)
        self._h_min = min(float(grid.L[a] / grid.N[a]) for a in range(grid.ndim))
        if We is not None and We > 0.0:
            self._dt_sigma = math.sqrt((1.0 + rho_ratio) * We * self._h_min ** 3 / (2.0 * math.pi))
        else:
            self._dt_sigma = None
        self._cn_viscous = bool(cn_viscous)
        self._work = _Work()

    def compute(self, velocity_components, mu, rho) -> float:
        io = Handoff(velocity_components[0])
        vel = [io.to_device(u) for u in velocity_components]
        lib = nat.lib()
        w = self._work.get(3 * 1024 + 8)
        n = vel[0].numel()
        out = (C.c_double * 3)()
        nat.check(lib.tpf_ns_absmax(len(vel), nat.parr([ptr(u) for u in vel]), n, ptr(w), out,
                                    stream_handle()))
        u_sum = 0.0
        for k in range(len(vel)):  # xp.sum(xp.stack(maxima)) in axis order
            u_sum += out[k]
        h, cfl = self._h_min, self.cfl
        cands = [cfl * h / max(u_sum, 1.0e-14)]
        if not self._cn_viscous:
            r = C.c_double()
            nat.check(lib.tpf_ns_ratio_max(ptr(io.to_device(mu)), ptr(io.to_device(rho)), n,
                                           ptr(w), C.byref(r), stream_handle()))
            cands.append(cfl * h * h / (4.0 * max(r.value, 1.0e-14)))
        if self._dt_sigma is not None:
            cands.append(cfl * self._dt_sigma)
        return float(max(min(cands), 1.0e-8))


class BoundaryConditionHandler:
    """simulation/boundary_condition.py:21-87 (in place on each component)."""

    def __init__(self, config):
        self.bc_type = config.numerics.bc_type
        self.ndim = config.grid.ndim

    def apply(self, velocity) -> None:
        comps = [velocity[a] for a in range(self.ndim)]
        shape = tuple(comps[0].shape)
        if not isinstance(comps[0], torch.Tensor):
            for a, k in enumerate(axis_kinds(self.bc_type, self.ndim)):
                lo = tuple(0 if j == a else slice(None) for j in range(self.ndim))
                hi = tuple(-1 if j == a else slice(None) for j in range(self.ndim))
                for u in comps:
                    if k == "wall":
                        u[lo] = 0.0
                        u[hi] = 0.0
                    else:
                        u[hi] = u[lo]
            return
        per = nat.iarr([1 if k == "periodic" else 0 for k in axis_kinds(self.bc_type, self.ndim)])
        nat.check(nat.lib().tpf_ns_wall_bc(self.ndim, nat.i64arr(shape), per, len(comps),
                                           nat.parr([ptr(u) for u in comps]), stream_handle()))


class NullFieldExtender:
    """levelset/field_extender.py:205-224: no extension."""

    is_null_extender = True

    def compute_normal(self, phi):
        return None

    def extend(self, field_data, phi, n_hat=None):
        return field_data

    def extend_both(self, field_data, phi, n_hat=None):
        return field_data


class ClosestPointExtender:
    """levelset/closest_point_extender.py:67-228 (2-D, as the reference)."""

    is_null_extender = False

    def __init__(self, backend, grid, ccd):
        require_uniform(grid, "ClosestPointExtender")
        self.grid, self.ccd, self.ndim = grid, ccd, grid.ndim
        self._coords = None
        self._work = _Work()

    def compute_normal(self, phi):
        io = Handoff(phi)
        t = io.to_device(phi)
        nd = self.ndim
        g = torch.empty((nd,) + tuple(t.shape), dtype=torch.float64, device="cuda")
        self.ccd.apply_into(t, g, None, (1 << nd) - 1)
        out = _empty_like_list(t, nd)
        nat.check(nat.lib().tpf_ns_normal(nd, t.numel(), ptr(g), nat.parr([ptr(o) for o in out]),
                                          stream_handle()))
        return [io.back(o) for o in out]

    def extend(self, q, phi, n_hat=None):
        if self.ndim != 2:
            raise NotImplementedError("ClosestPointExtender: only ndim=2 implemented")
        if n_hat is None:
            n_hat = self.compute_normal(phi)
        io = Handoff(q)
        qt, pt = io.to_device(q), io.to_device(phi)
        nx, ny = (io.to_device(n) for n in n_hat)
        if self._coords is None:
            self._coords = [torch.from_numpy(np.ascontiguousarray(c, dtype=np.float64)).cuda()
                            for c in self.grid.coords]
        out = torch.empty_like(qt)
        w = self._work.get(8 * qt.numel())
        nat.check(nat.lib().tpf_ns_extend_2d(
            self.ccd._plan, nat.i64arr(qt.shape), ptr(self._coords[0]), ptr(self._coords[1]),
            ptr(qt), ptr(pt), ptr(nx), ptr(ny), ptr(out), ptr(w), stream_handle()))
        return io.back(out)

    def extend_both(self, q, phi, n_hat=None):
        raise NotImplementedError("extend_both is not used by the legacy step")


class VelocityCorrector:
    """u^{n+1} = u* - (dt / rho) D1 p with wall-normal zeroing
    (coupling/velocity_corrector.py:52-89); returns new arrays."""

    def __init__(self, backend, grid, ccd):
        require_uniform(grid, "VelocityCorrector")
        self.grid, self.ccd = grid, ccd

    def correct(self, vel_star, p, rho, dt):
        from .fields import correct_velocity
        io = Handoff(vel_star[0])
        out = [io.to_device(u).clone() for u in vel_star]
        correct_velocity(self.ccd, out, io.to_device(p), io.to_device(rho), dt)
        return [io.back(u) for u in out]
