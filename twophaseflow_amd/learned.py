"""Learned preconditioner for the variable-density PPE and flexible CG.

No counterpart in the reference; this is the learned extension.  Structure:

* ``MGHierarchy`` / ``MGPreconditioner``: the FP32 geometric multigrid of
  libtpf (``tpf_mg_*``) on the reference CG system ``(-W L) p = -W rhs``:
  homogenised face conductances per level, red-black Gauss-Seidel, full
  weighting and trilinear interpolation, V- or W-cycles.
* ``UNet``: 4 levels (16, 32, 64, 64 channels), 3x3x3 convolutions in
  hand-written BF16 MFMA kernels with FP32 accumulation (libtpf
  ``tpf_conv3d_fwd`` / ``tpf_conv3d_wgrad``; the input gradient is the forward
  kernel with flipped, transposed weights).  Node-centred restriction
  (1/4, 1/2, 1/4 per axis) between levels and linear interpolation back;
  skip connections by channel concatenation.  PyTorch provides tensors,
  autograd bookkeeping and Adam; no PyTorch convolution is used.
* ``LearnedDeflationPreconditioner``: once per solve the U-Net maps the
  conductance field of one coarse level to k coarse vectors W; their
  prolongations span a deflation space applied after each multigrid cycle
  in the A-DEF2 form z = V(r) + Z E^-1 Z^T (r - A V(r)), E = Z^T A Z
  (Galerkin, FP64 Cholesky).  The preconditioner is linear in r during a
  solve.
* ``fcg_solve``: flexible CG with m-term orthogonalisation (Notay's FCG(m))
  in FP64; plain Jacobi-PCG (``tpf_ppe_solve``, warm-started) takes over on
  stagnation, and the FP64 true residual judges the result.
* ``train_deflation``: the network is trained on the residuals that the
  multigrid-FCG solver produces on random bubble fields (8-128 spheres,
  density ratio log-uniform in [10, 1000]) by maximising the share of the
  post-cycle error energy that the deflation space captures.
"""

from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import torch

from . import _native as nat
from .device import stream_handle

BF16 = torch.bfloat16


def _lib():
    return nat.lib()


def _p(t):
    return None if t is None else t.data_ptr()


# ---------------------------------------------------------------- convs ---
def conv_fwd(x, wt_bf16, bias, relu, out_f32):
    D, H, W, Cin = x.shape
    Cout = wt_bf16.shape[1]
    assert x.is_contiguous() and wt_bf16.is_contiguous() and x.dtype == BF16
    y = torch.empty((D, H, W, Cout), dtype=torch.float32 if out_f32 else BF16, device=x.device)
    nat.check(_lib().tpf_conv3d_fwd(_p(x), _p(wt_bf16), _p(bias), _p(y), D, H, W, Cin, Cout,
                                    int(bool(relu)), int(bool(out_f32)), stream_handle()))
    return y


def conv_wgrad(x, gy_bf16, gw):
    D, H, W, Cin = x.shape
    Cout = gy_bf16.shape[3]
    nat.check(_lib().tpf_conv3d_wgrad(_p(x), _p(gy_bf16), _p(gw), D, H, W, Cin, Cout,
                                      stream_handle()))


def flip_weights(wb):
    """W2[tap][ci][co] = Wt[26 - tap][co][ci]: the dgrad weights."""
    return wb.flip(0).transpose(1, 2).contiguous()


class _Conv(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w, b, relu, out_f32):
        wb = w.to(BF16).contiguous()
        y = conv_fwd(x, wb, b, relu, out_f32)
        ctx.save_for_backward(x, wb, y)
        ctx.relu = relu
        return y

    @staticmethod
    def backward(ctx, gy):
        x, wb, y = ctx.saved_tensors
        gy = gy.contiguous()
        if ctx.relu:
            gy = gy * (y > 0)
        gyb = gy.to(BF16).contiguous()
        gx = conv_fwd(gyb, flip_weights(wb), None, False, False)
        gw = torch.zeros(wb.shape, dtype=torch.float32, device=x.device)
        conv_wgrad(x, gyb, gw)
        gb = gy.float().sum(dim=(0, 1, 2))
        return gx, gw, gb, None, None


def conv(x, w, b, relu=True, out_f32=False):
    return _Conv.apply(x.contiguous(), w, b, relu, out_f32)


# ---------------------------------------------------- grid transfers ---
def restrict_axis(x, axis):
    """Node-centred full weighting along `axis`: 2m+1 -> m+1 nodes,
    y_i = x_2i / 2 + (x_2i-1 + x_2i+1) / 4 with even reflection at walls."""
    assert x.shape[axis] % 2 == 1, "node-centred transfer needs 2m+1 nodes"
    ev = x[_sl(x, axis, 0, None, 2)]
    od = x[_sl(x, axis, 1, None, 2)]
    left = torch.cat([od.narrow(axis, 0, 1), od], dim=axis)
    right = torch.cat([od, od.narrow(axis, od.shape[axis] - 1, 1)], dim=axis)
    return 0.5 * ev + 0.25 * (left + right)


def prolong_axis(x, axis):
    """Linear interpolation m+1 -> 2m+1 nodes along `axis`."""
    n = x.shape[axis]
    shape = list(x.shape)
    shape[axis] = 2 * n - 1
    y = torch.empty(shape, dtype=x.dtype, device=x.device)
    y[_sl(y, axis, 0, None, 2)] = x
    y[_sl(y, axis, 1, None, 2)] = 0.5 * (x.narrow(axis, 0, n - 1) + x.narrow(axis, 1, n - 1))
    return y


def _sl(x, axis, start, stop, step):
    s = [slice(None)] * x.dim()
    s[axis] = slice(start, stop, step)
    return tuple(s)


def restrict3(x, axes=(0, 1, 2)):
    for a in axes:
        x = restrict_axis(x, a)
    return x


def prolong3(x, axes=(0, 1, 2)):
    for a in axes:
        x = prolong_axis(x, a)
    return x


# ---------------------------------------------------------------- U-Net ---
class UNet(torch.nn.Module):
    """4-level U-Net on channels-last BF16 activations (input 16 channels:
    normalised residual, log10 density, zero padding)."""

    CIN = 16

    def __init__(self, ch=(16, 32, 64, 64), seed=0, k_out=16):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        c0, c1, c2, c3 = ch
        spec = {"e0a": (self.CIN, c0), "e0b": (c0, c0), "e1a": (c0, c1), "e1b": (c1, c1),
                "e2a": (c1, c2), "e2b": (c2, c2), "ba": (c2, c3), "bb": (c3, c3),
                "d2": (c3 + c2, c2), "d1": (c2 + c1, c1), "d0": (c1 + c0, c0), "out": (c0, k_out)}
        self.spec = spec
        for name, (ci, co) in spec.items():
            std = math.sqrt(2.0 / (27 * ci)) * (0.1 if name == "out" else 1.0)
            self.register_parameter(f"{name}_w", torch.nn.Parameter(
                torch.randn(27, co, ci, generator=g) * std))
            self.register_parameter(f"{name}_b", torch.nn.Parameter(torch.zeros(co)))

    def _c(self, name, x, relu=True, out_f32=False):
        return conv(x, getattr(self, f"{name}_w"), getattr(self, f"{name}_b"), relu, out_f32)

    def forward(self, x, all_channels=False):
        e0 = self._c("e0b", self._c("e0a", x))
        e1 = self._c("e1b", self._c("e1a", _down(e0)))
        e2 = self._c("e2b", self._c("e2a", _down(e1)))
        b = self._c("bb", self._c("ba", _down(e2)))
        d2 = self._c("d2", torch.cat([_up(b), e2], dim=3))
        d1 = self._c("d1", torch.cat([_up(d2), e1], dim=3))
        d0 = self._c("d0", torch.cat([_up(d1), e0], dim=3))
        y = self._c("out", d0, relu=False, out_f32=True)
        return y if all_channels else y[..., 0]


def _down(x):
    return restrict3(x.float()).to(BF16).contiguous()


def _up(x):
    return prolong3(x.float()).to(BF16).contiguous()


# ------------------------------------------------- multigrid backbone ---
class MGHierarchy:
    """FP32 geometric multigrid hierarchy of a PPE handle (libtpf ``tpf_mg_*``):
    homogenised face conductances per level, red-black Gauss-Seidel,
    full-weighting restriction, trilinear prolongation."""

    def __init__(self, ppe_solver, max_levels=12):
        self.s = ppe_solver
        self.h = ppe_solver._handle
        lib = _lib()
        lv = C.c_int()
        nat.check(lib.tpf_mg_setup(self.h, int(max_levels), C.byref(lv)))
        self.levels = lv.value
        self.shapes, self.pins = [], []
        for level in range(self.levels):
            n3 = (C.c_int64 * 3)()
            pin = C.c_int64()
            nat.check(lib.tpf_mg_level_info(self.h, level, n3, C.byref(pin)))
            self.shapes.append(tuple(int(v) for v in n3))
            self.pins.append(int(pin.value))

    def _new(self, level):
        return torch.empty(self.shapes[level], dtype=torch.float32, device="cuda")

    def vcycle(self, r, nu_pre=2, nu_post=2, coarse_sweeps=20, stop_level=-1, gamma=1):
        r = r.float().contiguous().view(self.shapes[0])
        z = self._new(0)
        nat.check(_lib().tpf_mg_vcycle(self.h, _p(r), _p(z), nu_pre, nu_post, coarse_sweeps,
                                       stop_level, gamma))
        return z

    def down(self, r, level, nu_pre=2):
        r = r.float().contiguous().view(self.shapes[0])
        bc = self._new(level)
        nat.check(_lib().tpf_mg_down(self.h, _p(r), level, nu_pre, _p(bc)))
        return bc

    def up(self, level, xc, nu_post=2):
        z = self._new(0)
        nat.check(_lib().tpf_mg_up(self.h, level, _p(xc.contiguous()), nu_post, _p(z)))
        return z

    def prolong(self, level, xc):
        z = self._new(0)
        nat.check(_lib().tpf_mg_prolong(self.h, level, _p(xc.contiguous()), _p(z)))
        return z

    def restrict(self, r, level):
        bc = self._new(level)
        nat.check(_lib().tpf_mg_restrict(self.h, _p(r.contiguous()), level, _p(bc)))
        return bc

    def apply(self, level, x):
        y = self._new(level)
        nat.check(_lib().tpf_mg_apply(self.h, level, _p(x.contiguous()), _p(y)))
        return y

    def smooth(self, level, b, x, sweeps):
        nat.check(_lib().tpf_mg_smooth(self.h, level, _p(b.contiguous()), _p(x), sweeps))
        return x

    def coeffs(self, level):
        out = torch.empty((4,) + self.shapes[level], dtype=torch.float32, device="cuda")
        nat.check(_lib().tpf_mg_coeffs(self.h, level, _p(out)))
        return out


class MGPreconditioner:
    """z = M(r): one classical multigrid cycle (no network); gamma 1 = V, 2 = W."""

    def __init__(self, hier: MGHierarchy, nu=2, coarse_sweeps=20, stop_level=-1, gamma=1):
        self.hier, self.nu, self.cs, self.stop, self.gamma = (hier, nu, coarse_sweeps,
                                                              stop_level, gamma)

    def __call__(self, r):
        return self.hier.vcycle(r, self.nu, self.nu, self.cs, self.stop,
                                self.gamma).view(r.shape)


def level_features(hier: MGHierarchy, level: int) -> torch.Tensor:
    """Per-node network inputs of a level (channels 1..4): log10 of the three
    face conductances relative to the unit-density value, and of 1/diag,
    scaled to O(1)."""
    co = hier.coeffs(level)
    n = hier.shapes[level]
    h = 1.0 / (max(n) - 1)
    feats = torch.empty((4,) + n, dtype=torch.float32, device="cuda")
    for a in range(3):
        feats[a] = torch.log10(co[a] / h + 1e-4) / 3.0
    feats[3] = torch.log10(co[3] * 6.0 * h + 1e-4) / 3.0
    return feats


class _Galerkin(torch.autograd.Function):
    """A_c w = R A P w for a batch of coarse vectors w [k, *shape] (the
    Galerkin coarse operator of the fine FP32 operator; symmetric, so the
    backward applies it again)."""

    @staticmethod
    def forward(ctx, w, hier, level):
        ctx.hier, ctx.level = hier, level
        return _galerkin(hier, level, w)

    @staticmethod
    def backward(ctx, g):
        return _galerkin(ctx.hier, ctx.level, g.contiguous()), None, None


def _galerkin(hier, level, w):
    out = torch.empty_like(w)
    pin0 = hier.pins[0]
    for j in range(w.shape[0]):
        f = hier.apply(0, hier.prolong(level, w[j]))
        f.view(-1)[pin0] = 0.0
        out[j] = hier.restrict(f, level)
    return out


class LearnedDeflationPreconditioner:
    """Multigrid V-cycle with a learned coarse space (A-DEF2 two-level form).

    Once per solve, the U-Net maps the conductance field of one coarse level
    (no residual) to k coarse vectors W (k = the network's output channels);
    Z = P W spans the deflation space, E = Z^T A Z = W^T (R A P) W is
    factored in FP64.  Each application is linear in r:
        z1 = V(r);  t = r - A z1;  z = z1 + Z E^-1 Z^T t
    (Z^T t = W^T R t and Z y = P (W y) are evaluated on the coarse level).
    With net=None and no vectors it is the V-cycle alone.

    ``vectors`` [k, *coarse shape] replaces the network's output (any fixed
    coarse space, e.g. per-bubble indicators).  ``galerkin="level"`` forms
    E from the hierarchy's own operator on that level (W^T A_l W, no fine
    work) instead of the exact R A P, which costs k fine-level applications.
    ``setup_times`` holds the network and E times (device-synchronised)."""

    def __init__(self, hier: MGHierarchy, net, level: int, nu=2, coarse_sweeps=20, ridge=1e-6,
                 gamma=1, vectors=None, galerkin="exact"):
        if galerkin not in ("exact", "level"):
            raise ValueError(f"galerkin must be 'exact' or 'level', got {galerkin!r}")
        self.hier, self.net, self.level = hier, net, level
        self.nu, self.cs, self.gamma = nu, coarse_sweeps, gamma
        self.setup_times = {}
        if vectors is not None:
            self.net = vectors
        if self.net is not None:
            t0 = time.perf_counter()
            with torch.no_grad():
                if vectors is not None:
                    W = vectors.float().clone().contiguous()
                    W.view(W.shape[0], -1)[:, hier.pins[level]] = 0.0
                    W = W / (W.reshape(W.shape[0], -1).norm(dim=1).view(-1, 1, 1, 1) + 1e-30)
                else:
                    W = deflation_vectors(hier, net, level)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                if galerkin == "exact":
                    AW = _galerkin(hier, level, W)
                else:
                    AW = torch.stack([hier.apply(level, W[j]) for j in range(W.shape[0])])
                    AW.view(W.shape[0], -1)[:, hier.pins[level]] = 0.0
                E = (W.reshape(W.shape[0], -1) @ AW.reshape(W.shape[0], -1).T).double()
            E = 0.5 * (E + E.T)
            E = E + ridge * torch.trace(E) / E.shape[0] * torch.eye(E.shape[0], device=E.device,
                                                                 dtype=E.dtype)
            self.W = W.reshape(W.shape[0], -1)
            self.L = torch.linalg.cholesky(E)
            torch.cuda.synchronize()
            self.setup_times = {"network": t1 - t0, "coarse_matrix": time.perf_counter() - t1}

    def __call__(self, r):
        shape = r.shape
        r32 = r.float().contiguous().view(self.hier.shapes[0])
        z1 = self.hier.vcycle(r32, self.nu, self.nu, self.cs, -1, self.gamma)
        if self.net is None:
            return z1.view(shape)
        t = r32 - self.hier.apply(0, z1)
        c = (self.W @ self.hier.restrict(t, self.level).reshape(-1)).double()
        y = torch.cholesky_solve(c[:, None], self.L)[:, 0].float()
        zc = (y @ self.W).view(self.hier.shapes[self.level])
        return (z1 + self.hier.prolong(self.level, zc)).view(shape)


def deflation_vectors(hier: MGHierarchy, net, level: int) -> torch.Tensor:
    """The network's k coarse vectors [k, *shape] (coarse pin zeroed,
    unit-norm columns); differentiable in the network parameters."""
    feats = level_features(hier, level)
    inp = torch.zeros(hier.shapes[level] + (UNet.CIN,), dtype=torch.float32, device="cuda")
    inp[..., 0] = 1.0
    inp[..., 1:5] = feats.permute(1, 2, 3, 0)
    y = net(inp.to(BF16).contiguous(), all_channels=True)        # [*shape, k]
    W = y.permute(3, 0, 1, 2).contiguous()
    mask = torch.ones(hier.shapes[level], dtype=torch.float32, device="cuda")
    mask.view(-1)[hier.pins[level]] = 0.0
    W = W * mask
    return W / (W.reshape(W.shape[0], -1).norm(dim=1).view(-1, 1, 1, 1) + 1e-30)


def deflation_loss(hier, net, level, samples, ridge=1e-6):
    """Minus the mean fraction of each sample's error energy captured by the
    deflation space: -t^T Q t / t^T V(t) with Q = Z E^-1 Z^T, over
    post-V-cycle residual samples t (pairs (t, V(t)))."""
    W = deflation_vectors(hier, net, level)
    k = W.shape[0]
    AW = _Galerkin.apply(W, hier, level)
    E = (W.reshape(k, -1) @ AW.reshape(k, -1).T).double()
    E = 0.5 * (E + E.T)
    E = E + ridge * torch.trace(E).detach() / k * torch.eye(k, device=E.device, dtype=E.dtype)
    loss = 0.0
    for t, vt in samples:
        c = (W.reshape(k, -1) @ hier.restrict(t, level).reshape(-1)).double()
        cap = (c * torch.linalg.solve(E, c)).sum()
        loss = loss - cap / float((t * vt).sum())
    return loss / len(samples)


# ----------------------------------------------------------- FCG solve ---
def pcg_fallback(ppe_solver, rhs, x, tol, maxiter):
    """Plain Jacobi-PCG (``tpf_ppe_solve`` with precond=1, FP64) warm-started
    from x, in place, with its own iteration budget.  Whatever route
    ``ppe_solver`` is configured for (GMRES, multigrid-FCG), the fallback is
    this one.  Non-convergence within `maxiter` is reported, not raised: the
    caller's FP64 true residual judges the result.  Returns (iterations,
    converged)."""
    lib = _lib()
    opts = nat.SolveOpts(tol=tol, maxiter=int(maxiter), precond=1,
                         check_every=int(getattr(ppe_solver, "check_every", 50)), use_x0=1,
                         mixed=0, inner_tol=0.0)
    rep = nat.SolveReport()
    status = lib.tpf_ppe_solve(ppe_solver._handle, _p(rhs), _p(x), C.byref(opts),
                               C.byref(rep))
    if status not in (nat.TPF_OK, nat.TPF_ERR_RUNTIME):
        nat.check(status)
    return int(rep.iterations), status == nat.TPF_OK


def fcg_solve(ppe_solver, rhs, precond, tol=1e-8, maxiter=2000, m=8, stall_window=30,
              stall_factor=0.7, restarts=3, fallback=True, record=(),
              fallback_maxiter=20000):
    """Flexible CG on the CG system in FP64 with a nonlinear preconditioner.

    FCG(m) (Notay 2000): each new direction is A-orthogonalised against the
    last m directions.  Progress is tracked on the best residual seen; when
    it fails to improve by `stall_factor` over `stall_window` iterations the
    recursion restarts from the FP64 true residual (at most `restarts` times)
    and then plain Jacobi-PCG (warm-started from x, at most
    `fallback_maxiter` iterations, ``pcg_fallback``) takes over.  The FP64
    true residual judges the result.  Residuals of the iterations listed in
    `record` are kept in info["records"].  Returns (p, info)."""
    lib = _lib()
    h = ppe_solver._handle

    def A64(v):
        out = torch.empty_like(v)
        nat.check(lib.tpf_ppe_system_apply(h, _p(v), _p(out)))
        return out

    b = torch.empty_like(rhs)
    nat.check(lib.tpf_ppe_system_rhs(h, _p(rhs), _p(b)))
    bn = float(b.norm())
    x = torch.zeros_like(b)
    r = b.clone()
    hist, best = [], []
    it, nrestart = 0, 0
    P, Q, PQ = [], [], []
    records = []
    reason = "maxiter"
    t0 = time.perf_counter()
    with torch.no_grad():
        while it < maxiter:
            if it in record:
                records.append(r.clone())
            rel = float(r.norm()) / bn
            hist.append(rel)
            best.append(min(rel, best[-1]) if best else rel)
            if rel <= tol:
                reason = "converged"
                break
            if len(best) > stall_window and best[-1] > stall_factor * best[-1 - stall_window]:
                if nrestart >= restarts:
                    reason = "stagnation"
                    break
                nrestart += 1
                r = b - A64(x)  # residual replacement
                P, Q, PQ = [], [], []
                best = [float(r.norm()) / bn]
                continue
            z = precond(r).double()
            p = z
            for pj, qj, pqj in zip(P, Q, PQ):
                p = p - (float(torch.dot(z.view(-1), qj.view(-1))) / pqj) * pj
            q = A64(p)
            pq = float(torch.dot(p.view(-1), q.view(-1)))
            alpha = float(torch.dot(p.view(-1), r.view(-1))) / pq
            x += alpha * p
            r -= alpha * q
            P.append(p); Q.append(q); PQ.append(pq)
            if len(P) > m:
                P.pop(0); Q.pop(0); PQ.pop(0)
            it += 1
    torch.cuda.synchronize()
    info = {"fcg_iterations": it, "fcg_exit": reason, "fcg_restarts": nrestart,
            "fcg_seconds": time.perf_counter() - t0, "fallback_iterations": 0}
    if reason != "converged" and fallback:
        info["fallback_iterations"], info["fallback_converged"] = pcg_fallback(
            ppe_solver, rhs, x, tol, fallback_maxiter)
    info["true_relres"] = ppe_solver.true_relres(rhs, x)
    info["history"] = hist
    info["records"] = records
    return x, info


def fcg_solve_native(ppe_solver, rhs, precond, tol=1e-8, maxiter=500, nu=2, gamma=2,
                     coarse_sweeps=20, w_levels=2, fallback=True, fallback_maxiter=20000):
    """The device FCG(1) of ``tpf_ppe_solve_mg`` (graph-replayed multigrid
    cycle, FP64 vectors and reductions, FP64 true-residual judge) with the
    learned coarse correction of ``precond`` (a LearnedDeflationPreconditioner)
    applied after every cycle on the device (``tpf_ppe_solve_mg_deflated``):
    z = V(r); z += P W E^-1 W^T R (r - A z).  Plain Jacobi-PCG from the FCG
    iterate (``pcg_fallback``) is the fallback when FCG stops above tol; the
    FP64 true residual judges the result."""
    if precond.net is None:
        raise ValueError("fcg_solve_native needs a learned coarse space (net is None)")
    lib = _lib()
    W = precond.W.float().contiguous()
    k = int(W.shape[0])
    einv = torch.cholesky_inverse(precond.L).double().cpu().numpy()
    einv = np.ascontiguousarray(0.5 * (einv + einv.T))
    p = torch.zeros_like(rhs)
    opts = nat.MgOpts(tol=tol, maxiter=maxiter, use_x0=0, nu=nu, gamma=gamma,
                      coarse_sweeps=coarse_sweeps, max_levels=0, w_levels=w_levels)
    rep = nat.SolveReport()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    status = lib.tpf_ppe_solve_mg_deflated(ppe_solver._handle, _p(rhs), _p(p), C.byref(opts),
                                            _p(W), k, int(precond.level),
                                            einv.ctypes.data_as(C.POINTER(C.c_double)),
                                            C.byref(rep))
    info = {"fcg_iterations": int(rep.iterations), "fcg_restarts": int(rep.refinements),
            "fcg_exit": "converged" if status == nat.TPF_OK else "maxiter",
            "fcg_seconds": time.perf_counter() - t0, "fallback_iterations": 0}
    if status not in (nat.TPF_OK, nat.TPF_ERR_RUNTIME):
        nat.check(status)
    if status != nat.TPF_OK and fallback:
        info["fallback_iterations"], info["fallback_converged"] = pcg_fallback(
            ppe_solver, rhs, p, tol, fallback_maxiter)
    info["true_relres"] = ppe_solver.true_relres(rhs, p)
    return p, info


# ------------------------------------------------------------- training ---
def make_problem(N, seed, nb_range=(8, 128), ratio_range=(10.0, 1000.0), return_bubbles=False):
    """Random bubble density and a divergence right-hand side on an N^3 wall
    grid (device tensors); with return_bubbles also the spheres' centres and
    radii and the density ratio."""
    from . import fields
    from .ccd import CCDSolver
    from .grid import Grid
    from .config import GridConfig
    rng = np.random.default_rng(seed)
    nb = int(rng.integers(nb_range[0], nb_range[1] + 1))
    ratio = float(np.exp(rng.uniform(np.log(ratio_range[0]), np.log(ratio_range[1]))))
    rmin, rmax = 0.03 * 128 / N if N < 128 else 0.03, 0.08 * 128 / N if N < 128 else 0.08
    centres, radii = fields.random_bubbles(nb, seed=seed + 7, ndim=3, rmin=min(rmin, 0.1),
                                           rmax=min(rmax, 0.2))
    _, rho = fields.bubbles_psi_rho((N,) * 3, (1.0,) * 3, centres, radii, 1.5 / N, 1.0,
                                    1.0 / ratio)
    g = Grid(GridConfig(ndim=3, N=(N,) * 3, L=(1.0,) * 3))
    ccd = CCDSolver(g, bc_type="wall")
    vel = [torch.from_numpy(fields.band_limited_periodic(N, 3, kmax=8, seed=seed * 3 + a)).cuda()
           for a in range(3)]
    rhs = fields.divergence(ccd, vel)
    if return_bubbles:
        return g, rho, rhs, (centres, radii, ratio)
    return g, rho, rhs


def level_of(hier: MGHierarchy, nodes: int) -> int:
    """Index of the level whose grid has `nodes` nodes per axis."""
    for level, shp in enumerate(hier.shapes):
        if max(shp) == nodes:
            return level
    raise ValueError(f"no level with {nodes} nodes per axis in {hier.shapes}")


def train_deflation(net, N=128, problems=200, record=(0, 2, 5, 9), level_nodes=33, lr=1e-3,
                    seed=1000, log=None, nu=2, gamma=1):
    """Trains `net` in place to produce the coarse space of
    LearnedDeflationPreconditioner from residuals that the multigrid-FCG
    solver itself produces: for each random bubble problem (8-128 spheres,
    density ratio log-uniform in [10, 1000]) the FCG residuals r_k of the
    listed iterations give samples t = r - A V(r), and deflation_loss is
    minimised with Adam (one step per problem)."""
    from .config import SimulationConfig, SolverConfig
    from .ppe import PPESolverFVMMatrixFree
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    hist = []
    t0 = time.perf_counter()
    cfg = SimulationConfig(solver=SolverConfig(pseudo_tol=1e-10, pseudo_maxiter=10 ** 6))
    cfg.solver.ppe_iteration_method = "cg"
    s = None
    for i in range(problems):
        g, rho, rhs = make_problem(N, seed + i)
        if s is None:
            s = PPESolverFVMMatrixFree(None, cfg, g)
        s.prepare_operator(rho)
        H = MGHierarchy(s)
        level = level_of(H, level_nodes)
        _, info = fcg_solve(s, rhs, MGPreconditioner(H, nu=nu, gamma=gamma), tol=1e-10,
                            maxiter=max(record) + 1, m=1, fallback=False, record=record)
        samples = []
        with torch.no_grad():
            for r in info["records"]:
                r32 = r.float().view(H.shapes[0])
                t = r32 - H.apply(0, H.vcycle(r32, nu, nu, 20, -1, gamma))
                t = t / (t.norm() + 1e-30)
                samples.append((t, H.vcycle(t, nu, nu, 20, -1, gamma)))
        loss = deflation_loss(H, net, level, samples)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        torch.nn.utils.clip_grad_norm_(net.parameters(), 1.0)
        opt.step()
        hist.append(float(loss.detach()))
        if log and (i + 1) % max(1, problems // 10) == 0:
            log(f"problem {i + 1}/{problems}: mean loss (last 20) {np.mean(hist[-20:]):.4f}")
    return {"loss_history": hist, "seconds": time.perf_counter() - t0}
