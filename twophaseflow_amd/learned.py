"""Learned preconditioner for the variable-density PPE and flexible CG.

No counterpart in the reference; this is the learned extension.  Structure:

* ``UNet``: 4 levels (16, 32, 64, 64 channels), 3x3x3 convolutions in
  hand-written BF16 MFMA kernels with FP32 accumulation (libtpf
  ``tpf_conv3d_fwd`` / ``tpf_conv3d_wgrad``; the input gradient is the forward
  kernel with flipped, transposed weights).  Node-centred restriction
  (1/4, 1/2, 1/4 per axis) between levels and linear interpolation back;
  skip connections by channel concatenation.  PyTorch provides tensors,
  autograd bookkeeping and Adam; no PyTorch convolution is used.
* ``TwoLevelPreconditioner``: damped-Jacobi pre-smoothing on the fine grid,
  restriction of the residual, the U-Net as the coarse-grid solver (input:
  normalised coarse residual and log10 density), interpolation with an
  energy-optimal step length, Jacobi post-smoothing.  Fine-grid matvecs are
  the FP32 CG-system operator of the PPE handle (``tpf_ppe_system_apply_f32``).
* ``fcg_solve``: flexible CG with one-term orthogonalisation (Notay's
  FCG(1)) on the reference CG system ``(-W L) p = -W rhs`` in FP64; plain
  Jacobi-PCG (``tpf_ppe_solve``, warm-started) takes over on stagnation, and
  the FP64 true residual judges the result.
* ``train``: the network is trained on residuals produced by the plain PCG on
  random bubble fields (8-128 spheres, density ratio log-uniform in
  [10, 1000]) by minimising the A-norm energy z.A z - 2 z.r of the
  preconditioner output, normalised by r.D^-1 r.
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

    def __init__(self, ch=(16, 32, 64, 64), seed=0):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        c0, c1, c2, c3 = ch
        spec = {"e0a": (self.CIN, c0), "e0b": (c0, c0), "e1a": (c0, c1), "e1b": (c1, c1),
                "e2a": (c1, c2), "e2b": (c2, c2), "ba": (c2, c3), "bb": (c3, c3),
                "d2": (c3 + c2, c2), "d1": (c2 + c1, c1), "d0": (c1 + c0, c0), "out": (c0, 16)}
        self.spec = spec
        for name, (ci, co) in spec.items():
            std = math.sqrt(2.0 / (27 * ci)) * (0.1 if name == "out" else 1.0)
            self.register_parameter(f"{name}_w", torch.nn.Parameter(
                torch.randn(27, co, ci, generator=g) * std))
            self.register_parameter(f"{name}_b", torch.nn.Parameter(torch.zeros(co)))

    def _c(self, name, x, relu=True, out_f32=False):
        return conv(x, getattr(self, f"{name}_w"), getattr(self, f"{name}_b"), relu, out_f32)

    def forward(self, x):
        e0 = self._c("e0b", self._c("e0a", x))
        e1 = self._c("e1b", self._c("e1a", _down(e0)))
        e2 = self._c("e2b", self._c("e2a", _down(e1)))
        b = self._c("bb", self._c("ba", _down(e2)))
        d2 = self._c("d2", torch.cat([_up(b), e2], dim=3))
        d1 = self._c("d1", torch.cat([_up(d2), e1], dim=3))
        d0 = self._c("d0", torch.cat([_up(d1), e0], dim=3))
        return self._c("out", d0, relu=False, out_f32=True)[..., 0]


def _down(x):
    return restrict3(x.float()).to(BF16).contiguous()


def _up(x):
    return prolong3(x.float()).to(BF16).contiguous()


# ------------------------------------------------- fine-grid operator ---
class _A32(torch.autograd.Function):
    """FP32 CG-system matvec of a PPE handle; A is symmetric on vectors whose
    pin entry is zero, so the backward is A itself."""

    @staticmethod
    def forward(ctx, p, handle):
        out = torch.empty_like(p)
        nat.check(_lib().tpf_ppe_system_apply_f32(handle, _p(p.contiguous()), _p(out)))
        ctx.handle = handle
        return out

    @staticmethod
    def backward(ctx, g):
        out = torch.empty_like(g)
        nat.check(_lib().tpf_ppe_system_apply_f32(ctx.handle, _p(g.contiguous()), _p(out)))
        return out, None


class TwoLevelPreconditioner:
    """z = M(r): Jacobi, U-Net coarse correction, Jacobi (all FP32 inside)."""

    def __init__(self, ppe_solver, net: UNet, omega: float = 0.6):
        self.s = ppe_solver
        self.net = net
        self.omega = omega
        lib = _lib()
        h = ppe_solver._handle
        nat.check(lib.tpf_ppe_prepare_f32(h))
        dinv = torch.empty(ppe_solver.shape, dtype=torch.float64, device="cuda")
        nat.check(lib.tpf_ppe_jacobi_inv(h, _p(dinv)))
        self.dinv = dinv.float()
        self.pin = ppe_solver._pin_dof
        self.dinv.view(-1)[self.pin] = 0.0
        rho = ppe_solver._rho.float()
        self.logrho_c = (restrict3(torch.log10(rho)) / 3.0).contiguous()

    def A(self, p):
        return _A32.apply(p, self.s._handle)

    def __call__(self, r, train=False):
        r = r.float() if r.dtype != torch.float32 else r
        om = self.omega
        z1 = om * self.dinv * r
        t = r - self.A(z1)
        rc = restrict3(t)
        scale = rc.square().mean().sqrt() + 1e-30
        inp = torch.zeros(rc.shape + (UNet.CIN,), dtype=torch.float32, device=r.device)
        inp[..., 0] = rc / scale
        inp[..., 1] = self.logrho_c
        ec = self.net(inp.to(BF16).contiguous()) * scale
        pe = prolong3(ec)
        pe = pe * self._pin_mask()
        ape = self.A(pe)
        gamma = (pe * t).sum() / ((pe * ape).sum() + 1e-300)
        z2 = z1 + gamma * pe
        t2 = t - gamma * ape
        z = z2 + om * self.dinv * t2
        return z * self._pin_mask()

    def _pin_mask(self):
        if not hasattr(self, "_mask"):
            m = torch.ones(self.s.shape, dtype=torch.float32, device="cuda")
            m.view(-1)[self.pin] = 0.0
            self._mask = m
        return self._mask


# ----------------------------------------------------------- FCG solve ---
def fcg_solve(ppe_solver, rhs, precond, tol=1e-8, maxiter=2000, m=8, stall_window=30,
              stall_factor=0.7, restarts=3, fallback=True):
    """Flexible CG on the CG system in FP64 with a nonlinear preconditioner.

    FCG(m) (Notay 2000): each new direction is A-orthogonalised against the
    last m directions.  Progress is tracked on the best residual seen; when
    it fails to improve by `stall_factor` over `stall_window` iterations the
    recursion restarts from the FP64 true residual (at most `restarts` times)
    and then plain Jacobi-PCG (warm-started from x) takes over.  The FP64
    true residual judges the result.  Returns (p, info)."""
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
    reason = "maxiter"
    t0 = time.perf_counter()
    with torch.no_grad():
        while it < maxiter:
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
        ppe_solver.solve_into(rhs, None, x, use_x0=True)
        info["fallback_iterations"] = ppe_solver.last_diagnostics["iterations"]
    info["true_relres"] = ppe_solver.true_relres(rhs, x)
    info["history"] = hist
    return x, info


# ------------------------------------------------------------- training ---
def make_problem(N, seed, nb_range=(8, 128), ratio_range=(10.0, 1000.0)):
    """Random bubble density and a divergence right-hand side on an N^3 wall
    grid (device tensors)."""
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
    return g, rho, rhs


def pcg_residual_samples(ppe_solver, rhs, rho, iters=(0, 10, 40, 120)):
    """System residuals r_k = b - A x_k of the plain PCG after k iterations."""
    lib = _lib()
    h = ppe_solver._handle
    ppe_solver.prepare_operator(rho)
    b = torch.empty_like(rhs)
    nat.check(lib.tpf_ppe_system_rhs(h, _p(rhs), _p(b)))
    out = []
    for k in iters:
        x = torch.zeros_like(rhs)
        if k > 0:
            opts = nat.SolveOpts(tol=1e-30, maxiter=k, precond=1, check_every=min(k, 32),
                                 use_x0=0, mixed=0, inner_tol=0.0)
            rep = nat.SolveReport()
            lib.tpf_ppe_solve(h, _p(rhs), _p(x), C.byref(opts), C.byref(rep))  # stops at maxiter
        ax = torch.empty_like(x)
        nat.check(lib.tpf_ppe_system_apply(h, _p(x), _p(ax)))
        out.append((b - ax).float())
    return out


def energy_loss(M, r):
    z = M(r)
    az = M.A(z)
    num = (z * az).sum() - 2.0 * (z * r).sum()
    den = (r * M.dinv * r).sum() + 1e-300
    return num / den


def train(net, N=64, problems=64, epochs=2, lr=2e-3, seed=100, log=None):
    """Trains `net` in place on PCG residuals of random bubble problems."""
    from .config import SimulationConfig, SolverConfig
    from .ppe import PPESolverFVMMatrixFree
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    hist = []
    data = []
    t0 = time.perf_counter()
    for i in range(problems):
        g, rho, rhs = make_problem(N, seed + i)
        cfg = SimulationConfig(solver=SolverConfig(pseudo_tol=1e-8, pseudo_maxiter=10 ** 6))
        cfg.solver.ppe_iteration_method = "cg"
        s = PPESolverFVMMatrixFree(None, cfg, g)
        samples = pcg_residual_samples(s, rhs, rho)
        data.append((s, rho, samples))
    t_data = time.perf_counter() - t0
    for ep in range(epochs):
        for (s, rho, samples) in data:
            M = TwoLevelPreconditioner(s, net)
            for r in samples:
                r = r / (r.norm() + 1e-30)
                loss = energy_loss(M, r)
                opt.zero_grad(set_to_none=True)
                loss.backward()
                torch.nn.utils.clip_grad_norm_(net.parameters(), 1.0)
                opt.step()
                hist.append(float(loss.detach()))
        if log:
            log(f"epoch {ep}: mean energy {np.mean(hist[-len(data) * 4:]):.4f}")
    return {"loss_history": hist, "data_seconds": t_data,
            "train_seconds": time.perf_counter() - t0 - t_data}
