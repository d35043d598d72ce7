"""``CCDSolver`` on MI355X: the reference call surface over the HIP line kernels.

Mirrors ``ccd/ccd_solver.py:79-403``:
``differentiate(data, axis) -> (d1, d2)``, ``first_derivative``,
``second_derivative`` (each with ``apply_metric``), ``differentiate_raw``,
``enforce_wall_neumann`` and the attributes ``bc_type, grid, backend, xp,
ndim``.  On a non-uniform grid (``grid.uniform`` False) the lines are solved
in xi with spacing 1/N and mapped with the grid's metric ``J``,
``dJ_dxi`` (``ccd_solver_helpers.py:509-519``).  Inputs may be NumPy
arrays (copied to and from the device) or CUDA torch tensors (zero copy);
outputs are new arrays of the input's kind.  Errors follow the reference:
``ValueError`` for a wall axis with fewer than 6 nodes
(ccd_solver.py:459-465), ``np.linalg.LinAlgError`` for a singular pivot.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as nat
from .boundary import axis_kinds
from .device import Handoff, ptr, require_gpu, stream_handle


class CCDSolver:
    def __init__(self, grid, backend=None, bc_type: str = "wall", precision: str = "fp64"):
        self.grid = grid
        self.backend = backend
        self.xp = np
        self.ndim = grid.ndim
        self.bc_type = bc_type
        if hasattr(grid, "set_boundary_type"):
            grid.set_boundary_type(bc_type)
        self._kinds = axis_kinds(bc_type, self.ndim)
        if precision not in ("fp64", "fp32_refined"):
            raise ValueError("precision must be 'fp64' or 'fp32_refined'")
        self.precision = precision
        self.shape = tuple(int(n) + 1 for n in grid.N)
        self._lib = nat.lib()
        # compact-stencil spacing (ccd_solver.py:126-130): L/N on uniform
        # grids, the computational 1/N on non-uniform ones (metric applied)
        self._uniform = bool(getattr(grid, "uniform", True))
        self._h = [float(grid.L[a]) / float(grid.N[a]) if self._uniform else 1.0 / float(grid.N[a])
                   for a in range(self.ndim)]
        self._bcs = [nat.BC_PERIODIC if k == "periodic" else nat.BC_WALL for k in self._kinds]
        self._plan = self._new_plan()
        self._raw = None          # plan without the metric (apply_metric=False)
        self._metric_key = [None] * self.ndim
        # periodic axes are factorised eagerly in the reference (ccd_solver.py:113-120)
        for a, k in enumerate(self._kinds):
            if k == "periodic":
                self.axis_info(a)

    def __del__(self):
        for name in ("_plan", "_raw"):
            plan = getattr(self, name, None)
            if plan is not None and plan.value:
                self._lib.tpf_ccd_plan_destroy(plan)
                setattr(self, name, None)

    def _new_plan(self, shape=None, bcs=None):
        plan = C.c_void_p()
        nat.check(self._lib.tpf_ccd_plan_create(
            C.byref(plan), self.ndim, nat.i64arr(shape or self.shape), nat.darr(self._h),
            nat.iarr(bcs or self._bcs),
            nat.PREC_FP64 if self.precision == "fp64" else nat.PREC_FP32_REFINED))
        return plan

    def _sync_metric(self) -> None:
        """Non-uniform grids: uploads grid.J / grid.dJ_dxi of every axis whose
        arrays changed since the last call (the reference reads them on
        every call, ccd_solver_helpers.py:509-540)."""
        if self._uniform:
            return
        for a in range(self.ndim):
            J, dJ = self.grid.J[a], self.grid.dJ_dxi[a]
            key = (id(J), id(dJ))
            if key == self._metric_key[a]:
                continue
            host = [np.ascontiguousarray(np.asarray(t.cpu() if hasattr(t, "cpu") else t,
                                                    dtype=np.float64).reshape(-1))
                    for t in (J, dJ)]
            if any(h.shape[0] != self.shape[a] for h in host):
                raise ValueError(f"grid metric of axis {a} must have {self.shape[a]} values")
            dp = C.POINTER(C.c_double)
            nat.check(self._lib.tpf_ccd_plan_set_metric(
                self._plan, a, host[0].ctypes.data_as(dp), host[1].ctypes.data_as(dp)))
            self._metric_key[a] = key

    def _raw_plan(self):
        if self._uniform:
            return self._plan
        if self._raw is None:
            self._raw = self._new_plan()
        return self._raw

    # -- introspection -------------------------------------------------------
    def axis_info(self, axis: int) -> dict:
        P, M, W = C.c_int(), C.c_int(), C.c_int()
        t = C.c_double()
        nat.check(self._lib.tpf_ccd_axis_info(self._plan, axis, C.byref(P), C.byref(M),
                                               C.byref(W), C.byref(t)))
        k = C.c_int()
        nat.check(self._lib.tpf_ccd_axis_kernel(self._plan, axis, C.byref(k)))
        kernel = {v: name for name, v in self.MODES.items()}[k.value]
        return {"partitions": P.value, "max_rows": M.value, "window": W.value,
                "trunc_coeff": t.value, "kernel": kernel}

    # -- device-level entry points ---------------------------------------------
    MODES = {"auto": 0, "lane": 1, "wave": 2, "explicit": 3}

    def set_launch(self, axis: int, mode: str = "auto", rows_cap: int = 0) -> None:
        """Launch shape of one axis: kernel mapping ("auto", "lane": a lane
        per line, "wave": a wavefront per line) and rows-per-partition cap
        (0 = default); the knobs the autotuner sets."""
        self._check_axis(axis)
        if mode not in self.MODES:
            raise ValueError(f"mode must be one of {sorted(self.MODES)}")
        nat.check(self._lib.tpf_ccd_plan_set_launch(self._plan, int(axis), self.MODES[mode],
                                                     int(rows_cap)))

    def set_store_align(self, axis: int, align: int = -1) -> None:
        """Store chunk (elements) of the explicit-inverse kernel on a strided
        axis: 8 or 16 write whole aligned chunks, 0 the lines' own stores,
        -1 the default.  Results do not depend on it."""
        self._check_axis(axis)
        nat.check(self._lib.tpf_ccd_plan_set_store_align(self._plan, int(axis), int(align)))

    def differentiate_into(self, f: torch.Tensor, d1, d2, axis: int,
                           apply_metric: bool = True) -> None:
        """Zero-copy form: f, d1, d2 are contiguous CUDA float64 tensors of
        grid.shape (d1 or d2 may be None).  Asynchronous on torch's stream.
        On a non-uniform grid the metric maps the xi derivatives to x unless
        ``apply_metric`` is False."""
        self._check_axis(axis)
        self._sync_metric()
        plan = self._plan if apply_metric else self._raw_plan()
        nat.check(self._lib.tpf_ccd_differentiate(
            plan, ptr(f), ptr(d1), ptr(d2), axis, stream_handle()))

    def apply_into(self, f: torch.Tensor, d1, d2, axis_mask: int) -> None:
        """All axes in ``axis_mask`` (bit a = axis a); d1/d2 are (k, *shape)."""
        self._sync_metric()
        nat.check(self._lib.tpf_ccd_apply(
            self._plan, ptr(f), ptr(d1), ptr(d2), int(axis_mask), stream_handle()))

    def apply_fields_into(self, fs, d1s, d2s, axis_mask: int) -> None:
        """Axis a (bit a of ``axis_mask``) differentiates fs[a] along a into
        d1s[a] / d2s[a], in one launch when the axes share a launch shape.
        fs, d1s, d2s are sequences of ndim contiguous CUDA float64 tensors of
        grid.shape (entries of unmasked axes, and d1s / d2s themselves, may
        be None)."""
        n = self.ndim
        arr = C.c_void_p * n

        def pack(ts):
            if ts is None:
                return None
            if len(ts) != n:
                raise ValueError("apply_fields_into: one entry per axis")
            return arr(*[ptr(t) if t is not None else None for t in ts])
        self._sync_metric()
        nat.check(self._lib.tpf_ccd_apply_fields(self._plan, pack(fs), pack(d1s), pack(d2s),
                                                 int(axis_mask), stream_handle()))

    # -- reference API -----------------------------------------------------------
    def differentiate(self, data, axis: int, bc_left=None, bc_right=None,
                      apply_metric: bool = True):
        return self._run(data, axis, bc_left, bc_right, want1=True, want2=True,
                         apply_metric=apply_metric)

    def first_derivative(self, data, axis: int, bc_left=None, bc_right=None,
                         apply_metric: bool = True):
        return self._run(data, axis, bc_left, bc_right, want1=True, want2=False,
                         apply_metric=apply_metric)[0]

    def second_derivative(self, data, axis: int, bc_left=None, bc_right=None,
                          apply_metric: bool = True):
        return self._run(data, axis, bc_left, bc_right, want1=False, want2=True,
                         apply_metric=apply_metric)[1]

    def differentiate_raw(self, data, axis: int):
        """1-D input of length N[axis]+1: the wall-closure solve in xi with
        the solver's spacing and no metric, on any axis (ccd_solver.py:
        174-203; the grid metric is built from it)."""
        arr = data if isinstance(data, torch.Tensor) else np.asarray(data)
        n = self.shape[axis]
        if arr.reshape(-1).shape[0] != n:
            raise ValueError(f"differentiate_raw expects {n} values")
        full = [self.shape[a] if a == axis else 1 for a in range(self.ndim)]
        # A plan over a single line: same spacing, wall closure along `axis`.
        lib = self._lib
        h = list(self._h)
        bcs = [nat.BC_WALL] * self.ndim
        plan = C.c_void_p()
        nat.check(lib.tpf_ccd_plan_create(C.byref(plan), self.ndim, nat.i64arr(full),
                                          nat.darr(h), nat.iarr(bcs), nat.PREC_FP64))
        try:
            io = Handoff(arr)
            f = io.to_device(arr, shape=full)
            d1 = torch.empty_like(f)
            d2 = torch.empty_like(f)
            nat.check(lib.tpf_ccd_differentiate(plan, ptr(f), ptr(d1), ptr(d2), axis,
                                                stream_handle()))
            return io.back(d1.reshape(-1)), io.back(d2.reshape(-1))
        finally:
            lib.tpf_ccd_plan_destroy(plan)

    def enforce_wall_neumann(self, grad, ax: int) -> None:
        """Zero the wall planes normal to ``ax`` in place
        (ccd_boundary_helpers.py:52-61)."""
        if self._kinds[ax] != "wall":
            return
        lo = [slice(None)] * grad.ndim
        hi = [slice(None)] * grad.ndim
        lo[ax], hi[ax] = 0, -1
        grad[tuple(lo)] = 0.0
        grad[tuple(hi)] = 0.0

    # -- internals -----------------------------------------------------------
    def _check_axis(self, axis):
        if not 0 <= axis < self.ndim:
            raise ValueError(f"axis {axis} out of range for ndim={self.ndim}")

    def _run(self, data, axis, bc_left, bc_right, want1, want2, apply_metric=True):
        if bc_left is not None or bc_right is not None:
            raise NotImplementedError(
                "bc_left/bc_right overrides (testing only in the reference, "
                "ccd_solver.py:152-153) are not provided by the HIP path")
        self._check_axis(axis)
        require_gpu()
        io = Handoff(data)
        f = io.to_device(data)
        if tuple(f.shape) != self.shape:
            raise ValueError(f"data shape {tuple(f.shape)} != grid.shape {self.shape}")
        d1 = torch.empty_like(f) if want1 else None
        d2 = torch.empty_like(f) if want2 else None
        self.differentiate_into(f, d1, d2, axis, apply_metric=apply_metric)
        return (io.back(d1) if want1 else None, io.back(d2) if want2 else None)


class CCDBatch:
    """Many independent CCD problems (uneven grids allowed) differentiated in
    one grid per launch shape (libtpf ``tpf_ccd_batch_*``).  No counterpart
    in the reference: the batching extension.

    ``solvers``: CCDSolver instances; ``axes``: per solver the axes to
    differentiate (default all).  ``apply_into(fs, d1s, d2s)`` takes CUDA
    FP64 tensors, with d1s[j] / d2s[j] of shape (len(axes[j]), *shape_j)."""

    def __init__(self, solvers, axes=None):
        self._lib = nat.lib()
        self.solvers = list(solvers)
        n = len(self.solvers)
        if n == 0:
            raise ValueError("CCDBatch needs at least one solver")
        axes = axes if axes is not None else [tuple(range(s.ndim)) for s in self.solvers]
        self.axes = [tuple(a) for a in axes]
        masks = (C.c_uint32 * n)(*[sum(1 << a for a in ax) for ax in self.axes])
        plans = (C.c_void_p * n)(*[s._plan.value for s in self.solvers])
        h = C.c_void_p()
        nat.check(self._lib.tpf_ccd_batch_create(C.byref(h), plans, masks, n))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.tpf_ccd_batch_destroy(h)
            self._h = None

    def apply_into(self, fs, d1s, d2s) -> None:
        n = len(self.solvers)
        if len(fs) != n or (d1s is not None and len(d1s) != n) or (d2s is not None and len(d2s) != n):
            raise ValueError("CCDBatch.apply_into: one field and output pair per problem")
        for j, (s, f) in enumerate(zip(self.solvers, fs)):
            shape = tuple(int(x) + 1 for x in s.grid.N)
            want = (len(self.axes[j]),) + shape
            if (not f.is_cuda or f.dtype != torch.float64 or not f.is_contiguous()
                    or tuple(f.shape) != shape):
                raise ValueError(f"CCDBatch: field {j} must be a contiguous CUDA FP64 {shape}")
            for o in ((d1s[j] if d1s is not None else None), (d2s[j] if d2s is not None else None)):
                if o is not None and (not o.is_cuda or o.dtype != torch.float64
                                      or not o.is_contiguous() or tuple(o.shape) != want):
                    raise ValueError(f"CCDBatch: outputs of problem {j} must be CUDA FP64 {want}")
        arr = C.c_void_p * n
        f_p = arr(*[f.data_ptr() for f in fs])
        d1_p = arr(*[(o.data_ptr() if o is not None else None) for o in d1s]) if d1s is not None else None
        d2_p = arr(*[(o.data_ptr() if o is not None else None) for o in d2s]) if d2s is not None else None
        nat.check(self._lib.tpf_ccd_batch_apply(self._h, f_p, d1_p, d2_p, stream_handle()))
