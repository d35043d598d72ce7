"""GPU parity of the Rhie-Chow face-velocity divergence
(spatial/rhie_chow.py:55-284) against the oracle (oracle/rhie_chow.py):
1-D to 3-D, wall / periodic / mixed, with and without the balanced-force
bracket, NumPy and CUDA inputs; 1e-12 of the output's maximum."""

import sys

import numpy as np
import pytest
import torch

from oracle.rhie_chow import OracleRhieChow
from twophaseflow_amd import CCDSolver, Grid, GridConfig
from twophaseflow_amd.rhie_chow import RhieChowInterpolator

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,bc,L,bf", [
    ((40,), "wall", (1.0,), False),
    ((32,), "periodic", (2.0,), True),
    ((24, 16), "wall", (1.0, 0.5), True),
    ((24, 16), ("periodic", "wall"), (1.0, 1.0), False),
    ((12, 10, 8), "periodic", (1.0, 0.8, 0.6), True),
    ((12, 10, 8), "wall", (1.0, 1.0, 1.0), False),
    # several 16-plane slabs along axis 0, mixed boundaries
    ((40, 6, 5), ("wall", "periodic", "wall"), (2.0, 1.0, 1.0), True),
    ((33, 7, 6), ("periodic", "wall", "periodic"), (1.0, 0.5, 1.0), False),
    # planes of n1*n2 > 256 nodes over several workgroups, rows longer than a
    # wavefront (the 257^3 launch shape's paths)
    ((20, 4, 80), ("wall", "periodic", "wall"), (1.0, 0.4, 1.0), True),
    ((20, 5, 80), ("periodic", "wall", "periodic"), (1.0, 0.4, 1.0), False),
    ((17, 9, 130), "wall", (1.0, 0.5, 2.0), True),
])
def test_rhie_chow_matches_oracle(N, bc, L, bf):
    print(f"[gpu-case] rhie-chow N={N} bc={bc} bf={bf}", file=sys.stderr, flush=True)
    rng = np.random.default_rng(sum(N) + (7 if bf else 0))
    g = Grid(GridConfig(ndim=len(N), N=N, L=L))
    shape = g.shape
    vel = [rng.standard_normal(shape) for _ in N]
    p = rng.standard_normal(shape)
    rho = 1.0 + 99.0 * rng.random(shape)
    kw = {}
    if bf:
        kw = {"kappa": rng.standard_normal(shape), "psi": rng.random(shape), "we": 3.0}
    o = OracleRhieChow(N, L, bc)
    want = o.face_velocity_divergence(vel, p, rho, 0.05, **kw)
    rc = RhieChowInterpolator(None, g, CCDSolver(g, None, bc), bc)
    got = rc.face_velocity_divergence(vel, p, rho, 0.05, **kw)
    assert isinstance(got, np.ndarray)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    got_t = rc.face_velocity_divergence([torch.from_numpy(v).cuda() for v in vel],
                                        torch.from_numpy(p).cuda(),
                                        torch.from_numpy(rho).cuda(), 0.05,
                                        **{k: (torch.from_numpy(v).cuda()
                                               if isinstance(v, np.ndarray) else v)
                                           for k, v in kw.items()})
    assert isinstance(got_t, torch.Tensor)
    assert np.abs(got_t.cpu().numpy() - want).max() <= 1e-12 * np.abs(want).max()
