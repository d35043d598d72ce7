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
                                  ((255, 30), "wall"), ((120, 36), "wall_periodic"),
                                  ((6, 700), "wall_periodic"), ((600, 7), "wall")])
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


@pytest.mark.parametrize("bc", ["periodic", "wall"])
def test_fp32_refined_1024_squared(bc):
    """Config 2's 2-D 1024^2 field (band-limited, |k| <= 32) in FP32 solves +
    FP64 refinement, both axes: the strided axis 0 and the contiguous axis 1
    (wavefront kernel, 64 partitions x 16 rows).  f' to 1e-12 against the
    FP64 dense LU; f'' to max(1e-12, 2 x the dense LU's own error against the
    roundoff-free result): the closed-form spectral result on periodic lines
    (the FP64 rule of test_gpu_ccd.py), the long-double solve of the same
    system on wall lines (smooth non-periodic field; the dense LU's closure
    row is 1.3e-10 of max|f''| off there).  The FP64 kernel is held to the
    same bound on the wall field."""
    _announce(f"ccd fp32-refined 1025^2 {bc}")
    N = 1024
    g = Grid(GridConfig(ndim=2, N=(N, N), L=(1.0, 1.0)))
    if bc == "periodic":
        f = band_limited_periodic(N, 2, kmax=32, seed=2)
    else:
        x = np.linspace(0.0, 1.0, N + 1)
        X, Y = np.meshgrid(x, x, indexing="ij")
        f = np.exp(np.sin(np.pi * X) * np.cos(2 * np.pi * Y)) + X * Y ** 2
    s = CCDSolver(g, bc_type=bc, precision="fp32_refined")
    o = occd.OracleCCD((N, N), (1.0, 1.0), bc)
    ft = torch.from_numpy(f).cuda()
    for axis in (0, 1):
        d1, d2 = (t.cpu().numpy() for t in s.differentiate(ft, axis))
        r1, r2 = o.differentiate(f, axis)
        e1 = np.max(np.abs(d1 - r1)) / np.max(np.abs(r1))
        e2 = np.max(np.abs(d2 - r2)) / np.max(np.abs(r2))
        if bc == "periodic":
            s1, s2 = occd.periodic_spectral(f, axis, 1.0 / N)
        else:
            s1, s2 = occd.differentiate_wall_extended(f, axis, o._axis_info(axis))
        bound = max(TOL, 2.0 * np.max(np.abs(r2 - s2)) / np.max(np.abs(s2)))
        e2s = np.max(np.abs(d2 - s2)) / np.max(np.abs(s2))
        assert e2s <= bound, (axis, e2s, bound)
        if bc == "wall":
            c64 = CCDSolver(g, bc_type=bc)
            _, q2 = c64.differentiate(ft, axis)
            e64 = float(np.max(np.abs(q2.cpu().numpy() - s2)) / np.max(np.abs(s2)))
            print(f"[gpu-case]   axis {axis}: FP64 kernel f'' {e64:.2e} vs long double",
                  file=sys.stderr, flush=True)
            assert e64 <= bound, (axis, e64, bound)
            e2 = e2s  # judged against the roundoff-free result
        print(f"[gpu-case]   axis {axis} ({s.axis_info(axis)['kernel']}): f' {e1:.2e}  "
              f"f'' {e2:.2e} (bound {bound:.1e})", file=sys.stderr, flush=True)
        assert e1 <= TOL and e2 <= bound, (axis, e1, e2, bound)


def test_fp32_refined_rejects_lines_beyond_1024_unknowns():
    """Lines beyond 64 partitions x 16 rows are refused when the plan is built;
    a forced wavefront mapping of a short line is accepted."""
    g = Grid(GridConfig(ndim=2, N=(2048, 8), L=(1.0, 1.0)))
    with pytest.raises(ValueError, match="too long|FP32-refined"):
        CCDSolver(g, bc_type="periodic", precision="fp32_refined")
    s = CCDSolver(Grid(GridConfig(ndim=2, N=(64, 16), L=(1.0, 1.0))), bc_type="periodic",
                  precision="fp32_refined")
    s.set_launch(0, "wave", 8)
    f = np.random.default_rng(4).standard_normal((65, 17))
    d1, d2 = s.differentiate(f, 0)
    r1, r2 = occd.OracleCCD((64, 16), (1.0, 1.0), "periodic").differentiate(f, 0)
    assert np.max(np.abs(d1 - r1)) <= TOL * np.max(np.abs(r1))
    assert np.max(np.abs(d2 - r2)) <= TOL * np.max(np.abs(r2))
    with pytest.raises(ValueError, match="explicit|FP32-refined"):
        s.set_launch(0, "explicit")
