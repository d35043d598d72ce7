"""Seeded sweep of CCD shapes and boundary conditions on the GPU against the
dense-LU oracle: every axis by a single call (the default kernel per axis:
explicit partition inverse on strided axes, recurrences on the contiguous
one, wavefront per line for few long lines) and all axes in one apply.
Bound: 1e-12 relative max-norm, as for the fixed cases."""

import sys

import numpy as np
import pytest
import torch

from oracle import ccd as occd
from twophaseflow_amd import CCDSolver, Grid, GridConfig

pytestmark = pytest.mark.gpu


def _cases(count=24, seed=2024):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nd = int(rng.integers(2, 4))
        N = []
        kinds = []
        for _a in range(nd):
            kind = "periodic" if rng.random() < 0.5 else "wall"
            lo = 3 if kind == "periodic" else 5
            hi = 300 if nd == 2 else 90
            N.append(int(rng.integers(lo, hi)))
            kinds.append(kind)
        out.append((tuple(N), tuple(kinds)))
    return out


@pytest.mark.parametrize("N,kinds", _cases())
def test_ccd_random_shapes_match_oracle(N, kinds):
    print(f"[gpu-case] ccd sweep N={N} bc={kinds}", file=sys.stderr, flush=True)
    L = [1.0 + 0.1 * a for a in range(len(N))]
    g = Grid(GridConfig(ndim=len(N), N=N, L=tuple(L)))
    ccd = CCDSolver(g, bc_type=kinds)
    o = occd.OracleCCD(N, L, kinds)
    f_np = np.random.default_rng(sum(N)).standard_normal(g.shape)
    f = torch.from_numpy(f_np).cuda()
    refs = [o.differentiate(f_np, a) for a in range(len(N))]
    d1 = torch.empty((len(N),) + g.shape, dtype=torch.float64, device="cuda")
    d2 = torch.empty_like(d1)
    ccd.apply_into(f, d1, d2, (1 << len(N)) - 1)
    for a in range(len(N)):
        e1, e2 = ccd.differentiate(f, a)
        for got in ((e1, e2), (d1[a], d2[a])):
            for x, r in zip(got, refs[a]):
                rel = float(np.max(np.abs(x.cpu().numpy() - r)) / np.max(np.abs(r)))
                assert rel <= 1e-12, (a, ccd.axis_info(a)["kernel"], rel)
