"""GPU checks of the batched CCD launch and the launch-shape autotuner.

A batch runs the same per-line arithmetic (and kernel choice) as
tpf_ccd_apply on each problem's axes, so every problem's f' and f'' must
equal the single-problem results bit for bit;
other rows caps change the partitioning and agree with the dense-LU oracle
to 1e-12 (the CCD parity bound).
"""

import sys

import numpy as np
import pytest
import torch

from oracle import ccd as occd
from twophaseflow_amd import CCDBatch, CCDSolver, Grid, GridConfig
from twophaseflow_amd import autotune

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def _solver(N, bc):
    return CCDSolver(Grid(GridConfig(ndim=len(N), N=tuple(N), L=(1.0,) * len(N))), bc_type=bc)


def test_batch_equals_single_problem_launches():
    _announce("ccd batch uneven 2-D/3-D")
    specs = [((128, 128), "wall"), ((256, 256), "periodic"), ((512, 512), "wall"),
             ((64, 96), "periodic"), ((32, 40, 48), "wall"), ((256, 256), "wall")]
    solvers = [_solver(N, bc) for N, bc in specs]
    axes = [tuple(range(len(N))) for N, _ in specs]
    axes[3] = (1,)                      # a subset of axes
    batch = CCDBatch(solvers, axes)
    rng = np.random.default_rng(0)
    fs = [torch.from_numpy(rng.standard_normal(tuple(n + 1 for n in N))).cuda()
          for N, _ in specs]
    d1s = [torch.empty((len(a),) + tuple(f.shape), dtype=torch.float64, device="cuda")
           for a, f in zip(axes, fs)]
    d2s = [torch.empty_like(d) for d in d1s]
    batch.apply_into(fs, d1s, d2s)
    for j, (s, f, a) in enumerate(zip(solvers, fs, axes)):
        r1 = torch.empty_like(d1s[j])
        r2 = torch.empty_like(d2s[j])
        s.apply_into(f, r1, r2, sum(1 << ax for ax in a))
        for k, ax in enumerate(a):
            assert torch.equal(d1s[j][k], r1[k]) and torch.equal(d2s[j][k], r2[k]), (j, ax)
    # repeated calls reuse the device tables (pointers restaged per call)
    fs2 = [f * 2.0 for f in fs]
    batch.apply_into(fs2, d1s, d2s)
    torch.cuda.synchronize()
    s0 = solvers[0]
    r1 = torch.empty_like(d1s[0])
    r2 = torch.empty_like(d2s[0])
    s0.apply_into(fs2[0], r1, r2, 0b11)
    assert torch.equal(d1s[0][1], r1[1])
    # a strided axis alone runs the explicit inverse in the batch as in a single call
    sb = _solver((128, 128), "wall")
    bb = CCDBatch([sb], [(0,)])
    o1 = torch.empty((1, 129, 129), dtype=torch.float64, device="cuda")
    o2 = torch.empty_like(o1)
    bb.apply_into([fs[0]], [o1], [o2])
    e1, e2 = sb.differentiate(fs[0], 0)
    assert sb.axis_info(0)["kernel"] == "explicit"
    assert torch.equal(o1[0], e1) and torch.equal(o2[0], e2)


def test_batch_groups_above_the_kernel_argument_limit():
    """A group of more than 128 slots (here 140) takes its array pointers from
    the restaged device slot table instead of the kernel arguments."""
    _announce("ccd batch 70 x 2-D 24x20 (140 slots in one group)")
    solvers = [_solver((24, 20), "wall") for _ in range(70)]
    batch = CCDBatch(solvers)
    rng = np.random.default_rng(7)
    fs = [torch.from_numpy(rng.standard_normal((25, 21))).cuda() for _ in solvers]
    d1s = [torch.empty((2, 25, 21), dtype=torch.float64, device="cuda") for _ in solvers]
    d2s = [torch.empty_like(d) for d in d1s]
    for rep in range(2):  # the second call restages the table behind the first
        scale = 1.0 + rep
        batch.apply_into([f * scale for f in fs], d1s, d2s)
        torch.cuda.synchronize()
        for s, f, a, b in zip(solvers, fs, d1s, d2s):
            r1 = torch.empty_like(a)
            r2 = torch.empty_like(b)
            s.apply_into(f * scale, r1, r2, 0b11)
            assert torch.equal(a, r1) and torch.equal(b, r2)


def test_batch_rejects_bad_shapes():
    s = _solver((32, 32), "wall")
    batch = CCDBatch([s])
    f = torch.zeros((33, 33), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        batch.apply_into([f], [torch.zeros((1, 33, 33), dtype=torch.float64, device="cuda")], None)


@pytest.mark.parametrize("N,bc", [((256, 64), "periodic"), ((255, 40), "wall"),
                                  ((1024, 24), "periodic"), ((1024, 24), "wall")])
def test_launch_shapes_keep_parity(N, bc):
    """Every admissible (mapping, rows cap) of axis 0 against the dense-LU oracle."""
    _announce(f"ccd launch shapes N={N} {bc}")
    s = _solver(N, bc)
    f = np.random.default_rng(3).standard_normal(tuple(n + 1 for n in N))
    o = occd.OracleCCD(N, (1.0, 1.0), bc)
    r1, r2 = o.differentiate(f, 0)
    for mode, cap in autotune.admissible_shapes(N[0] + 1, bc == "periodic"):
        s.set_launch(0, mode, cap)
        d1, d2 = s.differentiate(f, 0)
        info = s.axis_info(0)
        assert info["max_rows"] == cap
        for d, r in ((d1, r1), (d2, r2)):
            assert np.max(np.abs(d - r)) <= 1e-12 * np.max(np.abs(r)), (mode, cap)
    with pytest.raises(ValueError):
        s.set_launch(0, "lane", 12)
    with pytest.raises(ValueError):
        s.set_launch(0, "wave", 32)


def test_autotune_picks_an_admissible_cap(tmp_path, monkeypatch):
    _announce("ccd autotune 2-D 513^2")
    monkeypatch.setenv("TPF_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    N = (512, 512)
    s = _solver(N, "wall")
    res = autotune.tune_ccd(s, reps=3)
    for axis, r in res.items():
        assert (r["mode"], r["rows_cap"]) in autotune.admissible_shapes(N[axis] + 1, False)
        assert s.axis_info(axis)["max_rows"] == r["rows_cap"]
    f = np.random.default_rng(4).standard_normal((513, 513))
    o = occd.OracleCCD(N, (1.0, 1.0), "wall")
    for axis in (0, 1):
        d1, _ = s.differentiate(f, axis)
        r1, _ = o.differentiate(f, axis)
        assert np.max(np.abs(d1 - r1)) <= 1e-12 * np.max(np.abs(r1))
    s2 = _solver(N, "wall")
    res2 = autotune.tune_ccd(s2, reps=3)     # served from the cache
    assert all(r["cached"] for r in res2.values())
