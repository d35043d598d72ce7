"""Oracle: 3x3x3 "same" convolution and its two gradients, FP64 NumPy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  No counterpart in
the reference (the learned preconditioner is this build's extension).  Pins:
``conv3d`` against an independent direct quadruple loop on tiny shapes, and
``conv3d_dgrad`` / ``conv3d_wgrad`` against central finite differences of
<y, g> (tests/test_oracle_unet.py), so a wrong flip or transpose fails.

Layouts follow the device kernels: x [D,H,W,Cin], weights Wt [27, Cout, Cin]
with tap = (dz+1)*9 + (dy+1)*3 + (dx+1), zero padding of one node.
"""

from __future__ import annotations

import numpy as np


def taps():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def _shift(x, dz, dy, dx):
    """s[z,y,x] = x[z+dz, y+dy, x+dx] with zeros outside."""
    D, H, W = x.shape[:3]
    s = np.zeros_like(x)
    zs, ze = max(0, -dz), min(D, D - dz)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    s[zs:ze, ys:ye, xs:xe] = x[zs + dz:ze + dz, ys + dy:ye + dy, xs + dx:xe + dx]
    return s


def conv3d(x, wt, bias=None, relu=False):
    y = np.zeros(x.shape[:3] + (wt.shape[1],))
    for t, (dz, dy, dx) in enumerate(taps()):
        y += _shift(x, dz, dy, dx) @ wt[t].T
    if bias is not None:
        y += bias
    if relu:
        y = np.maximum(y, 0.0)
    return y


def conv3d_dgrad(gy, wt):
    """d<conv3d(x), gy>/dx (no activation)."""
    gx = np.zeros(gy.shape[:3] + (wt.shape[2],))
    for t, (dz, dy, dx) in enumerate(taps()):
        gx += _shift(gy @ wt[t], -dz, -dy, -dx)
    return gx


def conv3d_wgrad(x, gy):
    """d<conv3d(x), gy>/dWt, shape [27, Cout, Cin]."""
    Cin, Cout = x.shape[3], gy.shape[3]
    gw = np.zeros((27, Cout, Cin))
    for t, (dz, dy, dx) in enumerate(taps()):
        gw[t] = gy.reshape(-1, Cout).T @ _shift(x, dz, dy, dx).reshape(-1, Cin)
    return gw


def conv3d_direct(x, wt):
    """Quadruple-loop definition (tiny shapes only)."""
    D, H, W, Cin = x.shape
    Cout = wt.shape[1]
    y = np.zeros((D, H, W, Cout))
    for z in range(D):
        for yy in range(H):
            for xx in range(W):
                for t, (dz, dy, dx) in enumerate(taps()):
                    zz, y2, x2 = z + dz, yy + dy, xx + dx
                    if 0 <= zz < D and 0 <= y2 < H and 0 <= x2 < W:
                        y[z, yy, xx] += wt[t] @ x[zz, y2, x2]
    return y


def bf16_round(a):
    """Round float64 values to the nearest bfloat16 (ties to even), as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)
