"""GPU checks of the learned preconditioner's BF16 MFMA convolutions against
the FP64 CPU oracle (the preconditioner itself: tests/test_gpu_mg.py).

Tolerances: inputs are BF16-representable and the kernels accumulate in FP32
over K = 27*Cin <= 3456 terms, so FP32 outputs match the oracle to 1e-4
relative (max-norm); BF16 outputs (forward activations, input gradients) to
one BF16 rounding, 1e-2 relative.
"""

import sys

import numpy as np
import pytest
import torch

from oracle import unet as ou
from twophaseflow_amd import learned as L

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def _bf(a):
    return torch.from_numpy(ou.bf16_round(a)).to(torch.bfloat16).cuda()


@pytest.mark.parametrize("shape,cout,relu", [((5, 6, 17, 16), 16, False), ((3, 4, 9, 32), 48, True),
                                             ((4, 4, 20, 48), 64, False), ((7, 3, 33, 16), 32, True)])
def test_conv_forward_matches_oracle(shape, cout, relu):
    _announce(f"conv fwd {shape}->{cout} relu={relu}")
    rng = np.random.default_rng(1)
    x = ou.bf16_round(rng.standard_normal(shape))
    w = ou.bf16_round(rng.standard_normal((27, cout, shape[3])) * 0.1)
    b = rng.standard_normal(cout).astype(np.float32)
    y = L.conv_fwd(_bf(x), _bf(w), torch.from_numpy(b).cuda(), relu, True).cpu().numpy()
    ref = ou.conv3d(x, w, b.astype(np.float64), relu)
    assert _rel(y, ref) <= 1e-4, _rel(y, ref)
    yb = L.conv_fwd(_bf(x), _bf(w), torch.from_numpy(b).cuda(), relu, False).float().cpu().numpy()
    assert _rel(yb, ref) <= 1e-2


@pytest.mark.parametrize("shape,cout", [((4, 5, 18, 16), 16), ((3, 3, 11, 32), 16), ((3, 4, 9, 16), 48),
                                        ((3, 3, 17, 128), 64),
                                        # K = 32 wgrad (channels 16 / 32 / 64; voxel counts
                                        # that leave a partial last chunk)
                                        ((5, 6, 13, 64), 64), ((4, 7, 9, 32), 64),
                                        ((3, 5, 21, 64), 16), ((6, 5, 7, 16), 32),
                                        ((9, 9, 9, 32), 32), ((4, 5, 6, 16), 64)])
def test_conv_gradients_match_oracle(shape, cout):
    _announce(f"conv grads {shape}->{cout}")
    rng = np.random.default_rng(2)
    x = ou.bf16_round(rng.standard_normal(shape))
    w = ou.bf16_round(rng.standard_normal((27, cout, shape[3])) * 0.1)
    g = ou.bf16_round(rng.standard_normal(shape[:3] + (cout,)))
    xt = _bf(x).requires_grad_(True)
    wt = torch.from_numpy(w).float().cuda().requires_grad_(True)
    bt = torch.zeros(cout, device="cuda", requires_grad=True)
    y = L.conv(xt, wt, bt, relu=False, out_f32=True)
    y.backward(torch.from_numpy(g).float().cuda())
    gx_ref = ou.conv3d_dgrad(g, w)
    gw_ref = ou.conv3d_wgrad(x, g)
    assert _rel(xt.grad.float().cpu().numpy(), gx_ref) <= 1e-2
    assert _rel(wt.grad.cpu().numpy(), gw_ref) <= 1e-4
    assert _rel(bt.grad.cpu().numpy(), g.reshape(-1, cout).sum(0)) <= 1e-4


def _solver(N, rho, tol=1e-8):
    from twophaseflow_amd import PPESolverFVMMatrixFree, SimulationConfig, SolverConfig
    from twophaseflow_amd import Grid, GridConfig
    cfg = SimulationConfig(solver=SolverConfig(pseudo_tol=tol, pseudo_maxiter=10 ** 6))
    cfg.solver.ppe_iteration_method = "cg"
    s = PPESolverFVMMatrixFree(None, cfg, Grid(GridConfig(ndim=3, N=(N,) * 3, L=(1.0,) * 3)))
    s.prepare_operator(rho)
    return s
