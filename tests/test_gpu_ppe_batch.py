"""Batched Jacobi-PCG (tpf_ppe_batch_*) and concurrent multigrid-FCG
(tpf_ppe_solve_mg_many) against single solves.

Every problem of a batch runs the arithmetic and reduction order of
tpf_ppe_solve, so its solution and report must equal the single solve's bit
for bit: uneven 2-D and 3-D grids, two-phase densities up to ratio 1000, a
zero right-hand side, initial guesses, skipped problems, and a failing
problem reported without disturbing the others.  The single solve itself is
pinned against the oracle's direct solve in test_gpu_ppe.py.
"""

import sys

import numpy as np
import pytest
import torch

from twophaseflow_amd import (Grid, GridConfig, PPEBatch, PPESolverFVMMatrixFree,
                              SimulationConfig, SolverConfig)

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def _solver(N, tol=1e-10, maxiter=20000, check_every=32, pre="jacobi"):
    cfg = SimulationConfig(grid=GridConfig(ndim=len(N), N=tuple(N), L=(1.0,) * len(N)),
                           solver=SolverConfig(pseudo_tol=tol, pseudo_maxiter=maxiter))
    cfg.solver.ppe_iteration_method = "cg"
    cfg.solver.ppe_check_every = check_every
    cfg.solver.ppe_preconditioner = pre
    return PPESolverFVMMatrixFree(None, cfg, Grid(cfg.grid))


def _problem(shape, ratio, seed, zero_rhs=False):
    rng = np.random.default_rng(seed)
    axes = [np.linspace(0.0, 1.0, n) for n in shape]
    X = np.meshgrid(*axes, indexing="ij")
    r = np.sqrt(sum((x - 0.5) ** 2 for x in X))
    psi = 1.0 / (1.0 + np.exp(-(r - 0.3) / 0.03))
    rho = 1.0 / ratio + (1.0 - 1.0 / ratio) * psi
    rhs = np.zeros(shape) if zero_rhs else rng.standard_normal(shape)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t(rho), t(rhs)


CASES = [((32, 32), 1000.0), ((40, 24), 10.0), ((64, 64), 300.0), ((12, 10, 14), 1000.0),
         ((128, 128), 1000.0)]


def _single(s, rho, rhs, x0=None):
    p = torch.empty_like(rhs) if x0 is None else x0.clone()
    s.solve_into(rhs, rho, p, use_x0=x0 is not None)
    return p, dict(s.last_diagnostics)


def test_batch_equals_single_solves():
    _announce("ppe batch: 5 uneven problems vs single solves")
    solvers = [_solver(N) for N, _ in CASES]
    probs = [_problem(s.shape, ratio, 10 + i) for i, (s, (_, ratio)) in
             enumerate(zip(solvers, CASES))]
    singles = [_single(s, rho, rhs) for s, (rho, rhs) in zip(solvers, probs)]
    batch = PPEBatch(solvers)
    for s, (rho, _) in zip(solvers, probs):
        s.prepare_operator(rho)
    xs = [torch.empty_like(rhs) for _, rhs in probs]
    for x in xs:
        x.fill_(float("nan"))  # x is written before it is read (zero initial guess)
    reps = batch.solve_into([rhs for _, rhs in probs], xs)
    for i, ((p, d), x, rep) in enumerate(zip(singles, xs, reps)):
        assert torch.equal(p, x), f"problem {i}: max diff {(p - x).abs().max().item():.3e}"
        assert rep["iterations"] == d["iterations"], i
        assert rep["relres"] == d["relres"] and rep["true_relres"] == d["true_relres"], i
        assert rep["true_relres"] <= 1e-10
    # a second call reuses the captured graph with new right-hand sides
    probs2 = [_problem(s.shape, ratio, 50 + i) for i, (s, (_, ratio)) in
              enumerate(zip(solvers, CASES))]
    for s, (rho, _) in zip(solvers, probs2):
        s.prepare_operator(rho)
    xs2 = [torch.empty_like(rhs) for _, rhs in probs2]
    batch.solve_into([rhs for _, rhs in probs2], xs2)
    for i, (s, (rho, rhs), x) in enumerate(zip(solvers, probs2, xs2)):
        p, _ = _single(s, rho, rhs)
        assert torch.equal(p, x), f"second call, problem {i}"


def test_batch_initial_guess_zero_rhs_and_skips():
    _announce("ppe batch: initial guesses, zero rhs, skipped problems")
    shapes = [(33, 33), (17, 29), (21, 21)]
    solvers = [_solver(tuple(n - 1 for n in sh)) for sh in shapes]
    probs = [_problem(sh, 100.0, 3 + i, zero_rhs=(i == 1)) for i, sh in enumerate(shapes)]
    gen = torch.Generator(device="cuda").manual_seed(7)
    x0 = [torch.randn(sh, dtype=torch.float64, device="cuda", generator=gen) for sh in shapes]
    singles = [_single(s, rho, rhs, x) for s, (rho, rhs), x in zip(solvers, probs, x0)]
    batch = PPEBatch(solvers)
    for s, (rho, _) in zip(solvers, probs):
        s.prepare_operator(rho)
    xs = [x.clone() for x in x0]
    batch.solve_into([rhs for _, rhs in probs], xs, use_x0=True)
    for i, ((p, d), x) in enumerate(zip(singles, xs)):
        assert torch.equal(p, x), f"problem {i}"
        assert batch.last_reports[i]["iterations"] == d["iterations"]
    # b = 0: SciPy's cg returns b itself whatever the initial guess
    assert torch.count_nonzero(singles[1][0]).item() == 0
    assert singles[1][1]["iterations"] == 0 and singles[1][1]["relres"] == 0.0
    # skip the middle problem: the others still equal their single solves
    xs = [x.clone() for x in x0]
    keep = xs[1].clone()
    reps = batch.solve_into([probs[0][1], None, probs[2][1]], [xs[0], None, xs[2]],
                            use_x0=True)
    assert reps[1] is None
    assert torch.equal(xs[1], keep)
    assert torch.equal(xs[0], singles[0][0]) and torch.equal(xs[2], singles[2][0])


def test_batch_reports_failing_problem():
    _announce("ppe batch: maxiter failure isolated to one problem")
    easy, hard = _solver((8, 8), maxiter=60), _solver((64, 64), maxiter=60)
    pe, ph = _problem(easy.shape, 2.0, 1), _problem(hard.shape, 1000.0, 2)
    batch = PPEBatch([easy, hard])
    easy.prepare_operator(pe[0])
    hard.prepare_operator(ph[0])
    xs = [torch.empty_like(pe[1]), torch.empty_like(ph[1])]
    reps = batch.solve_into([pe[1], ph[1]], xs, raise_on_failure=False)
    assert reps[0]["converged"] and not reps[1]["converged"]
    assert reps[1]["iterations"] == 60
    p, _ = _single(easy, *pe)
    assert torch.equal(p, xs[0])
    with pytest.raises(RuntimeError, match="problem\\(s\\) 1"):
        batch.solve_into([pe[1], ph[1]], xs)


def test_batch_rejects_mismatched_solvers():
    a, b = _solver((8, 8)), _solver((8, 8), tol=1e-8)
    with pytest.raises(ValueError, match="share"):
        PPEBatch([a, b])


def _single_any(s, rho, rhs, x0=None):
    """A single solve that may end above tol (its result is still compared)."""
    p = torch.empty_like(rhs) if x0 is None else x0.clone()
    try:
        s.solve_into(rhs, rho, p, use_x0=x0 is not None)
    except RuntimeError:
        pass
    return p, dict(s.last_diagnostics)


def test_multigrid_batch_equals_single_solves():
    """Every problem of a multigrid batch runs the kernels of its own
    tpf_ppe_solve_mg on its own stream: solutions and reports equal the
    single solves' bit for bit (uneven 2-D and 3-D grids, ratio up to 1000,
    a zero right-hand side, initial guesses, a skipped problem, a problem
    that reaches maxiter above tol)."""
    _announce("ppe batch, multigrid FCG: uneven problems vs single solves")
    cases = CASES + [((24, 24), 5.0)]
    solvers = [_solver(N, pre="mg") for N, _ in cases]
    probs = [_problem(s.shape, ratio, 20 + i, zero_rhs=(i == len(cases) - 1))
             for i, (s, (_, ratio)) in enumerate(zip(solvers, cases))]
    singles = [_single(s, rho, rhs) for s, (rho, rhs) in zip(solvers, probs)]
    batch = PPEBatch(solvers)
    assert batch.mode == "mg"
    for s, (rho, _) in zip(solvers, probs):
        s.prepare_operator(rho)
    xs = [torch.full_like(rhs, float("nan")) for _, rhs in probs]
    reps = batch.solve_into([rhs for _, rhs in probs], xs)
    for i, ((p, d), x, rep) in enumerate(zip(singles, xs, reps)):
        assert torch.equal(p, x), f"problem {i}: max diff {(p - x).abs().max().item():.3e}"
        for k in ("iterations", "relres", "true_relres", "refinements"):
            assert rep[k] == d[k], (i, k)
        assert rep["converged"] and rep["true_relres"] <= 1e-10
    assert torch.count_nonzero(xs[-1]).item() == 0 and reps[-1]["iterations"] == 0
    # initial guesses, and the second problem sits out
    gen = torch.Generator(device="cuda").manual_seed(9)
    x0 = [torch.randn(s.shape, dtype=torch.float64, device="cuda", generator=gen)
          for s in solvers]
    ref = [_single(s, rho, rhs, x) for s, (rho, rhs), x in zip(solvers, probs, x0)]
    for s, (rho, _) in zip(solvers, probs):
        s.prepare_operator(rho)
    xs = [x.clone() for x in x0]
    keep = xs[1].clone()
    rhs = [r for _, r in probs]
    rhs[1] = None
    reps = batch.solve_into(rhs, [x if r is not None else None for x, r in zip(xs, rhs)],
                            use_x0=True)
    assert reps[1] is None and torch.equal(xs[1], keep)
    for i in (0, 2, 3, 4, 5):
        assert torch.equal(xs[i], ref[i][0]), f"initial guess, problem {i}"
        assert reps[i]["iterations"] == ref[i][1]["iterations"], i


def test_multigrid_batch_reports_failing_problem():
    """maxiter 3: the easy problem starts from its converged solution (another
    solver's), the hard one from zero and ends above tol; only the hard one
    is reported, and both equal their single solves."""
    _announce("ppe batch, multigrid FCG: maxiter failure isolated to one problem")
    easy = _solver((16, 16), maxiter=3, pre="mg")
    hard = _solver((64, 64), maxiter=3, pre="mg")
    pe, ph = _problem(easy.shape, 2.0, 1), _problem(hard.shape, 1000.0, 2)
    x0e, _ = _single(_solver((16, 16), maxiter=200, pre="mg"), *pe)
    x0h = torch.zeros_like(ph[1])
    se, sh = _single_any(easy, *pe, x0=x0e), _single_any(hard, *ph, x0=x0h)
    assert se[1]["true_relres"] <= 1e-10 and sh[1]["true_relres"] > 1e-10
    batch = PPEBatch([easy, hard])
    easy.prepare_operator(pe[0])
    hard.prepare_operator(ph[0])
    xs = [x0e.clone(), x0h.clone()]
    reps = batch.solve_into([pe[1], ph[1]], xs, use_x0=True, raise_on_failure=False)
    assert reps[0]["converged"] and not reps[1]["converged"]
    assert reps[1]["iterations"] == 3 == sh[1]["iterations"]
    assert torch.equal(se[0], xs[0]) and torch.equal(sh[0], xs[1])
    xs = [x0e.clone(), x0h.clone()]
    with pytest.raises(RuntimeError, match="problem\\(s\\) 1"):
        batch.solve_into([pe[1], ph[1]], xs, use_x0=True)


def test_multigrid_batch_rejects_mixed_solvers():
    with pytest.raises(ValueError, match="'mg'"):
        PPEBatch([_solver((8, 8), pre="mg"), _solver((8, 8))])
    with pytest.raises(ValueError, match="'mg'"):
        PPEBatch([_solver((8, 8)), _solver((8, 8), pre="mg")])
    with pytest.raises(ValueError, match="share"):
        PPEBatch([_solver((8, 8), pre="mg"), _solver((8, 8), tol=1e-8, pre="mg")])
    s = _solver((8, 8), pre="mg")
    with pytest.raises(ValueError, match="twice"):
        PPEBatch([s, s])
