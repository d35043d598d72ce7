"""GPU parity of the dissipative-CCD level-set advection
(levelset/advection_dccd.py:84-152) against the oracle (oracle/dccd.py): one
TVD-RK3 step on 1-D to 3-D periodic / wall / mixed grids, with clipping,
with and without mass correction, NumPy and CUDA inputs; 1e-12 of max|psi|.
Also: the reference's in-place image synchronisation of the input."""

import sys

import numpy as np
import pytest
import torch

from oracle.dccd import OracleDCCD
from twophaseflow_amd import CCDSolver, Grid, GridConfig
from twophaseflow_amd.dccd import DissipativeCCDAdvection

pytestmark = pytest.mark.gpu


def _psi(shape, rng):
    grids = np.meshgrid(*[np.linspace(0, 1, n) for n in shape], indexing="ij")
    r = np.sqrt(sum((g - 0.5) ** 2 for g in grids))
    psi = 1.0 / (1.0 + np.exp(np.clip((r - 0.3) / 0.06, -500, 500)))
    return psi + 0.01 * rng.standard_normal(shape)  # some nodes outside [0, 1]


@pytest.mark.parametrize("N,ccd_bc,ls_bc,mass", [
    ((48,), "periodic", "periodic", False),
    ((40,), "wall", "neumann", True),
    ((32, 24), "periodic", "periodic", True),
    ((32, 24), "wall", "neumann", False),
    ((24, 20), ("periodic", "wall"), ("periodic", "neumann"), True),
    ((16, 12, 10), "periodic", "periodic", False),
    ((16, 12, 10), "wall", "neumann", True),
])
def test_dccd_step_matches_oracle(N, ccd_bc, ls_bc, mass):
    print(f"[gpu-case] dccd N={N} bc={ccd_bc} mass={mass}", file=sys.stderr, flush=True)
    rng = np.random.default_rng(sum(N))
    g = Grid(GridConfig(ndim=len(N), N=N, L=(1.0,) * len(N)))
    psi = _psi(g.shape, rng)
    vel = [0.5 * rng.standard_normal(g.shape) for _ in N]
    o = OracleDCCD(N, (1.0,) * len(N), ccd_bc, ls_bc, mass_correction=mass)
    want = o.advance(psi.copy(), vel, 0.002)
    adv = DissipativeCCDAdvection(None, g, CCDSolver(g, None, ccd_bc), bc=ls_bc,
                                  mass_correction=mass)
    host = psi.copy()
    got = adv.advance(host, vel, 0.002)
    assert isinstance(got, np.ndarray)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    assert np.array_equal(host, o.sync(psi.copy()))  # input images synchronised in place
    t = torch.from_numpy(psi.copy()).cuda()
    got_t = adv.advance(t, [torch.from_numpy(v).cuda() for v in vel], 0.002)
    assert np.abs(got_t.cpu().numpy() - want).max() <= 1e-12 * np.abs(want).max()
    assert np.array_equal(t.cpu().numpy(), o.sync(psi.copy()))


@pytest.mark.parametrize("shape,per", [
    ((9,), (1,)), ((9,), (0,)), ((7, 6), (1, 0)), ((7, 6), (0, 1)), ((7, 6), (1, 1)),
    ((5, 6, 7), (1, 1, 1)), ((5, 6, 7), (0, 1, 1)), ((5, 6, 7), (1, 0, 1)), ((5, 6, 7), (0, 0, 0)),
])
def test_sync_periodic_images_in_place_and_copy(shape, per):
    """tpf_sync_periodic_images against the reference's axis-by-axis copies
    (core/boundary.py:116-133): out-of-place (every node written) and in place
    (only the image planes change; no launch without a periodic axis)."""
    import ctypes as C
    from twophaseflow_amd import _native as nat
    print(f"[gpu-case] image sync {shape} periodic={per}", file=sys.stderr, flush=True)
    rng = np.random.default_rng(len(shape) * 10 + sum(per))
    a = rng.standard_normal(shape)
    want = a.copy()
    for ax, p in enumerate(per):
        if p:
            src = [slice(None)] * len(shape)
            dst = [slice(None)] * len(shape)
            src[ax], dst[ax] = 0, -1
            want[tuple(dst)] = want[tuple(src)]
    lib = nat.lib()
    dims, pa = nat.i64arr(shape), nat.iarr(per)
    t_in = torch.from_numpy(a).cuda()
    t_out = torch.full_like(t_in, float("nan"))
    nat.check(lib.tpf_sync_periodic_images(len(shape), dims, pa, C.c_void_p(t_in.data_ptr()),
                                           C.c_void_p(t_out.data_ptr()), None))
    t_io = torch.from_numpy(a).cuda()
    nat.check(lib.tpf_sync_periodic_images(len(shape), dims, pa, C.c_void_p(t_io.data_ptr()),
                                           C.c_void_p(t_io.data_ptr()), None))
    torch.cuda.synchronize()
    assert np.array_equal(t_out.cpu().numpy(), want)
    assert np.array_equal(t_io.cpu().numpy(), want)
