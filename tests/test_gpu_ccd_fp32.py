"""GPU checks of the FP32-refined CCD mode (FP32 solves, FP64 residual,
one refinement; FP64 inputs and outputs).

Error against the FP64 dense-LU oracle (and the FP64 kernel at 257^3),
relative max-norm per output.  One refinement of an FP32 solve whose
relative error is ~kappa * 2^-24 leaves ~(kappa * 2^-24)^2 of |x|; the bound
written here is 1e-12 on every case (the FP64 bound), and the measured errors
are printed (random fields ~1e-14, the band-limited 257^3 field 5e-13 on f'').
"""

import sys

import numpy as np
import pytest
import torch

from oracle import ccd as occd
from twophaseflow_amd import CCDSolver, Grid, GridConfig
from twophaseflow_amd.fields import band_limited_periodic

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _announce(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


@pytest.mark.parametrize("N,bc", [((64, 48), "wall"), ((64, 48), "periodic"),
                                  ((40, 24, 32), "wall"), ((256, 40), "periodic"),
                                  ((255, 30), "wall"), ((120, 36), "wall_periodic")])
def test_fp32_refined_matches_fp64_oracle(N, bc):
    _announce(f"ccd fp32-refined N={N} {bc}")
    L = (1.0,) * len(N)
    g = Grid(GridConfig(ndim=len(N), N=N, L=L))
    s = CCDSolver(g, bc_type=bc, precision="fp32_refined")
    f = np.random.default_rng(9).standard_normal(tuple(n + 1 for n in N))
    o = occd.OracleCCD(N, L, bc)
    for axis in range(len(N)):
        d1, d2 = s.differentiate(f, axis)
        r1, r2 = o.differentiate(f, axis)
        e1 = np.max(np.abs(d1 - r1)) / np.max(np.abs(r1))
        e2 = np.max(np.abs(d2 - r2)) / np.max(np.abs(r2))
        print(f"[gpu-case]   axis {axis}: f' {e1:.2e}  f'' {e2:.2e}", file=sys.stderr, flush=True)
        assert e1 <= TOL and e2 <= TOL, (axis, e1, e2)


def test_fp32_refined_full_size_257_cubed():
    """All axes of the 257^3 band-limited periodic field against the FP64 kernel."""
    _announce("ccd fp32-refined 257^3")
    N = 256
    f = torch.from_numpy(band_limited_periodic(N, 3, kmax=32, seed=2)).cuda()
    g = Grid(GridConfig(ndim=3, N=(N,) * 3, L=(1.0,) * 3))
    a = CCDSolver(g, bc_type="periodic")
    b = CCDSolver(g, bc_type="periodic", precision="fp32_refined")
    r1 = torch.empty((3,) + g.shape, dtype=torch.float64, device="cuda")
    r2 = torch.empty_like(r1)
    o1 = torch.empty_like(r1)
    o2 = torch.empty_like(r1)
    a.apply_into(f, r1, r2, 0b111)
    b.apply_into(f, o1, o2, 0b111)
    e1 = float((o1 - r1).abs().max() / r1.abs().max())
    e2 = float((o2 - r2).abs().max() / r2.abs().max())
    print(f"[gpu-case]   257^3: f' {e1:.2e}  f'' {e2:.2e}", file=sys.stderr, flush=True)
    assert e1 <= TOL and e2 <= TOL


def test_fp32_refined_rejects_long_lines():
    """Lines beyond 16 partitions x 16 rows are refused when the plan is built
    (the refined solve keeps both iterates in registers)."""
    g = Grid(GridConfig(ndim=2, N=(1024, 16), L=(1.0, 1.0)))
    with pytest.raises(ValueError, match="FP32-refined"):
        CCDSolver(g, bc_type="periodic", precision="fp32_refined")
    s = CCDSolver(Grid(GridConfig(ndim=2, N=(64, 16), L=(1.0, 1.0))), bc_type="periodic",
                  precision="fp32_refined")
    with pytest.raises(ValueError, match="FP32-refined"):
        s.set_launch(0, "wave")
        s.differentiate(np.zeros((65, 17)), 0)
