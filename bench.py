"""Benchmark of the MI355X CCD operator and variable-density PPE.

Workload (one step): the CCD operator, f' and f'' along every axis of a 3-D
257^3-node periodic band-limited field (|k| <= 32, amplitude |k|^-2, seed 2)
in FP64: one launch for the two strided axes and one for the contiguous
axis, each with all lines of its axes.  `value` is
Mpts/s = grid points x axes / second over the timed steps, inputs resident in
HBM.  Also reported on the same line:

* roofline of the dominant CCD kernel (the explicit partition-inverse grid
  of the two strided axes, the larger of the step's two launches):
  algorithmic bytes per launch (read f once, write f' and f'' of both axes:
  40 B per point) over its mean launch time from HIP events on the launch
  stream;
* 2-D 1025^2 periodic all-axes CCD throughput;
* the 257^3 density-ratio-1000 PPE (64 gas spheres, r in [0.03, 0.08],
  eps = 1.5h, rhs = CCD divergence of a seeded smooth field): wall time to
  ||r||/||b|| <= 1e-8 for plain FP64 Jacobi-PCG and for FP32-inner/FP64-
  refined PCG, each with its FP64 true residual;
* cpu_baseline: the reference CPU algorithm (dense LU per axis via SciPy
  lu_solve, the oracle restatement) on a bounded slab of the same field,
  timed on this host.

`python bench.py --gpus N --steps K --warmup W`; N > 1 exits: this build is
single-GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "256^3 PPE time-to-1e-8: learned-FCG vs plain PCG; CCD Mpts/s & HBM %"
HBM_PEAK_GBS = 8000.0
# measured HBM bytes of one ccd_explicit_kernel<16, 8, true> launch over the
# strided axes 0 and 1 of 257^3 (KiB per dispatch, mean of 3; counter scale
# from a 1 GiB copy: FETCH_SIZE x 2, WRITE_SIZE x 1): FETCH_SIZE 161995 x 2 +
# WRITE_SIZE 692443 x 1
TRAFFIC_BYTES_CCD = int((161_995 * 2 + 692_443) * 1024)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_ccd(ccd, f, d1, d2, mask, steps, warmup, ndim):
    import torch
    for _ in range(warmup):
        ccd.apply_into(f, d1, d2, mask)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ccd.apply_into(f, d1, d2, mask)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def event_ms(fn, reps):
    """Mean duration of fn's launches from HIP events on torch's current
    stream (after two untimed calls)."""
    import torch
    fn()
    fn()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for a, b in ev:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    return sum(a.elapsed_time(b) for a, b in ev) / reps


def per_axis_event_times(ccd, f, d1, d2, reps):
    """Mean duration of each axis launch from HIP events on torch's stream
    (the stream the kernels are launched on)."""
    import torch
    n = f.numel()
    out = []
    for axis in range(f.dim()):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(reps)]
        for a, b in ev:
            a.record()
            ccd.differentiate_into(f, d1[axis], d2[axis], axis)
            b.record()
        torch.cuda.synchronize()
        ms = sum(a.elapsed_time(b) for a, b in ev) / reps
        out.append({"axis": axis, "ms": ms,
                    "GBps": 24.0 * n / (ms * 1e-3) / 1e9,
                    "info": ccd.axis_info(axis)})
    return out


def host_cores() -> int:
    """Cores this process may run on, capped at 16: the GPU box gives one
    GPU's job a 16-core share whatever the machine's core count (and sets
    OMP_NUM_THREADS to 16); more BLAS threads than that share oversubscribe."""
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    return min(cores, 16)


def cpu_baseline_ccd(N=256, seed=2):
    """Reference CPU path (dense LU per axis with SciPy lu_solve, the oracle
    restatement of ccd_solver_helpers.py:406-465) on the full 257^3 field, all
    three axes, with every host core of this job's share (host_cores); plus a
    1-thread figure on a 17-plane slab."""
    from threadpoolctl import threadpool_limits
    from oracle import ccd as occd
    from twophaseflow_amd.fields import band_limited_periodic
    f = band_limited_periodic(N, 3, kmax=32, seed=seed)
    info = occd.build_periodic_axis(N + 1, 1.0 / N)
    cores = host_cores()
    with threadpool_limits(limits=cores):
        t0 = time.perf_counter()
        for axis in range(3):
            occd.differentiate_periodic(f, axis, info)
        t_all = time.perf_counter() - t0
    slab = f[:, :, :17].copy()
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        occd.differentiate_periodic(slab, 0, info)
        occd.differentiate_periodic(slab, 1, info)
        t_one = time.perf_counter() - t0
    return {"value": round(3 * f.size / t_all / 1e6, 2), "unit": "Mpts/s", "cores": cores,
            "kind": "port",
            "sample": f"full workload: dense-LU CCD (SciPy lu_solve, oracle restatement) on the "
                      f"257^3 periodic field, 3 axes, {cores} BLAS threads, {t_all:.2f} s",
            "one_thread_Mpts_per_s": round(2 * slab.size / t_one / 1e6, 2),
            "one_thread_sample": f"257x257x17 slab, axes 0 and 1, {t_one:.2f} s"}


def ppe_case(N=256, nb=64, seed=3, ratio=1e-3):
    import numpy as np
    import torch
    from twophaseflow_amd import CCDSolver, Grid, GridConfig
    from twophaseflow_amd import fields
    g = Grid(GridConfig(ndim=3, N=(N, N, N), L=(1.0, 1.0, 1.0)))
    centres, radii = fields.random_bubbles(nb, seed=seed, ndim=3)
    psi, rho = fields.bubbles_psi_rho(g.N, g.L, centres, radii, 1.5 / N, 1.0, ratio)
    ccd = CCDSolver(g, bc_type="wall")
    vel = []
    for a in range(3):
        u = fields.band_limited_periodic(N, 3, kmax=8, seed=4 + a)
        vel.append(torch.from_numpy(u).cuda())
    rhs = fields.divergence(ccd, vel)
    del vel
    torch.cuda.synchronize()
    return g, rho, rhs


def ppe_solve_timed(g, rho, rhs, precision, tol=1e-8, maxiter=200000, inner_tol=0.0,
                    method="cg", line_solver="pcr"):
    import torch
    from twophaseflow_amd import PPESolverFVMMatrixFree, SimulationConfig, SolverConfig
    cfg = SimulationConfig(solver=SolverConfig(pseudo_tol=tol, pseudo_maxiter=maxiter))
    cfg.grid = None
    cfg.solver.ppe_iteration_method = method
    if method == "gmres":
        cfg.solver.ppe_restart = 80
        cfg.solver.ppe_line_solver = line_solver
    cfg.solver.ppe_precision = precision
    cfg.solver.ppe_check_every = 64
    cfg.solver.ppe_inner_tol = inner_tol
    s = PPESolverFVMMatrixFree(None, cfg, g)
    p = torch.empty_like(rhs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    diag = s.solve_into(rhs, rho, p)  # includes operator setup (Jacobi diagonal)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    if method == "gmres":
        tr = diag["true_relres"]  # FP64 ||rhs - L p|| / ||rhs|| recomputed at exit
    else:
        tr = s.true_relres(rhs, p)
    return {"seconds": t, "iterations": diag["iterations"], "refinements": diag["refinements"],
            "true_relres": tr, "converged": bool(tr <= tol)}


def mg_native_timed(g, rho, rhs, cycle="W", tol=1e-8, reps=2):
    """Native multigrid-FCG (tpf_ppe_solve_mg through the PPE plugin with
    ppe_preconditioner="mg"); time includes operator and hierarchy setup."""
    import torch
    from twophaseflow_amd import PPESolverFVMMatrixFree, SimulationConfig, SolverConfig
    cfg = SimulationConfig(solver=SolverConfig(pseudo_tol=tol, pseudo_maxiter=500))
    cfg.solver.ppe_iteration_method = "cg"
    cfg.solver.ppe_preconditioner = "mg"
    cfg.solver.ppe_mg_cycle = cycle
    s = PPESolverFVMMatrixFree(None, cfg, g)
    p = torch.empty_like(rhs)
    out = None
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d = s.solve_into(rhs, rho, p)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        out = {"seconds": t, "iterations": d["iterations"], "true_relres": d["true_relres"],
               "converged": bool(d["true_relres"] <= tol),
               "cycle": f"{cycle}(2,2)" + (", W on the top 2 levels" if cycle == "W" else "")}
    return out


def mg_solve_timed(g, rho, rhs, gamma=2, net=None, tol=1e-8, reps=2):
    """Multigrid-FCG (net=None) or learned-deflation FCG on the CG system;
    time includes the hierarchy setup (and the network + Galerkin setup)."""
    import torch
    from twophaseflow_amd import PPESolverFVMMatrixFree, SimulationConfig, SolverConfig
    from twophaseflow_amd import learned as L
    cfg = SimulationConfig(solver=SolverConfig(pseudo_tol=tol, pseudo_maxiter=10 ** 6))
    cfg.solver.ppe_iteration_method = "cg"
    s = PPESolverFVMMatrixFree(None, cfg, g)
    out = None
    for _ in range(reps):
        s.invalidate_cache()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.prepare_operator(rho)
        H = L.MGHierarchy(s)
        if net is None:
            M = L.MGPreconditioner(H, nu=2, gamma=gamma)
        else:
            M = L.LearnedDeflationPreconditioner(H, net, L.level_of(H, 33), nu=2, gamma=gamma)
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t0
        p, info = L.fcg_solve(s, rhs, M, tol=tol, maxiter=500, m=1)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        out = {"seconds": t, "setup_seconds": t_setup, "iterations": info["fcg_iterations"],
               "fallback_iterations": info["fallback_iterations"], "exit": info["fcg_exit"],
               "true_relres": info["true_relres"], "converged": bool(info["true_relres"] <= tol)}
    return out


def learned_native_timed(g, rho, rhs, net, tol=1e-8, reps=2):
    """Learned coarse space in the device FCG (tpf_ppe_solve_mg_deflated,
    the native W(2,2) depth-2 cycle + the correction after every cycle);
    time includes the hierarchy, network and Galerkin setup."""
    import torch
    from twophaseflow_amd import PPESolverFVMMatrixFree, SimulationConfig, SolverConfig
    from twophaseflow_amd import learned as L
    cfg = SimulationConfig(solver=SolverConfig(pseudo_tol=tol, pseudo_maxiter=10 ** 6))
    cfg.solver.ppe_iteration_method = "cg"
    s = PPESolverFVMMatrixFree(None, cfg, g)
    out = None
    for _ in range(reps):
        s.invalidate_cache()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.prepare_operator(rho)
        H = L.MGHierarchy(s)
        M = L.LearnedDeflationPreconditioner(H, net, L.level_of(H, 33), nu=2, gamma=2)
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t0
        p, info = L.fcg_solve_native(s, rhs, M, tol=tol, maxiter=500, gamma=2, w_levels=2)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        out = {"seconds": t, "setup_seconds": t_setup, "iterations": info["fcg_iterations"],
               "fallback_iterations": info["fallback_iterations"], "exit": info["fcg_exit"],
               "true_relres": info["true_relres"], "converged": bool(info["true_relres"] <= tol)}
    return out


def bubble_indicator_vectors(hier, level, centres, radii, k):
    """Per-bubble coarse vectors on one multigrid level (one smoothed sphere
    indicator per bubble, the k largest bubbles; coarse pin zeroed, unit
    norm): the hand-built space that deflation for bubbly flows uses, a
    control for the network's space."""
    import numpy as np
    import torch
    shp = hier.shapes[level]
    h = 1.0 / (max(shp) - 1)
    axes = [torch.linspace(0.0, 1.0, n, device="cuda", dtype=torch.float64) for n in shp]
    X = torch.meshgrid(*axes, indexing="ij")
    order = np.argsort(-np.asarray(radii))[:k]
    W = torch.empty((len(order),) + shp, dtype=torch.float32, device="cuda")
    for j, b in enumerate(order):
        d = torch.sqrt(sum((X[a] - float(centres[b][a])) ** 2 for a in range(3)))
        W[j] = (0.5 * (1.0 - torch.tanh((d - float(radii[b])) / h))).float()
    W.view(W.shape[0], -1)[:, hier.pins[level]] = 0.0
    return W / (W.reshape(W.shape[0], -1).norm(dim=1).view(-1, 1, 1, 1) + 1e-30)


def deflated_native_timed(g, rho, rhs, make, tol=1e-8, reps=2):
    """Device FCG with a deflation space built by make(hier, level) -> the
    preconditioner; time includes the hierarchy and the space's setup."""
    import torch
    from twophaseflow_amd import PPESolverFVMMatrixFree, SimulationConfig, SolverConfig
    from twophaseflow_amd import learned as L
    cfg = SimulationConfig(solver=SolverConfig(pseudo_tol=tol, pseudo_maxiter=10 ** 6))
    cfg.solver.ppe_iteration_method = "cg"
    s = PPESolverFVMMatrixFree(None, cfg, g)
    out = None
    for _ in range(reps):
        s.invalidate_cache()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.prepare_operator(rho)
        H = L.MGHierarchy(s)
        M = make(H, L.level_of(H, 33))
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t0
        p, info = L.fcg_solve_native(s, rhs, M, tol=tol, maxiter=500, gamma=2, w_levels=2,
                                     fallback=False)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        out = {"seconds": round(t, 5), "setup_seconds": round(t_setup, 5),
               "space_seconds": {k: round(v, 5) for k, v in M.setup_times.items()},
               "iterations": info["fcg_iterations"],
               "per_iteration_ms": round((t - t_setup) / max(1, info["fcg_iterations"]) * 1e3, 3),
               "true_relres": info["true_relres"], "converged": bool(info["true_relres"] <= tol)}
    return out


def learned_study(train_problems=400, ks=(16, 64), held_out=(9001, 9002, 9003)):
    """Controlled study of the learned coarse space against the same device
    W(2,2) multigrid-FCG without it: networks with k = 16 and 64 output
    vectors trained on `train_problems` random 129^3 bubble problems, the
    trained spaces evaluated at 129^3 and 257^3 (the 64-bubble ratio-1000
    case and held-out random problems) with the exact Galerkin E and with E
    from the level operator, beside per-bubble indicator vectors (the
    hand-built deflation space for bubbly flows) and untrained networks."""
    import numpy as np
    import torch
    from twophaseflow_amd import learned as L
    from twophaseflow_amd import fields
    res = {"train_problems": train_problems, "train": {}, "cases": []}
    nets = {}
    for k in ks:
        log(f"learned study: training k={k} on {train_problems} problems (129^3)")
        torch.manual_seed(0)
        net = L.UNet(seed=0, k_out=k).cuda()
        tr = L.train_deflation(net, N=128, problems=train_problems, lr=1e-3, seed=1000, gamma=2,
                               log=log)
        lh = np.asarray(tr["loss_history"])
        marks = [n for n in (20, 50, 100, 200, 400, 800, 1600, 2000) if n <= len(lh)]
        res["train"][f"k{k}"] = {"seconds": round(tr["seconds"], 1),
                                 "mean_loss_of_20_before": {str(n): round(float(lh[n - 20:n].mean()), 4)
                                                            for n in marks}}
        log(f"  {res['train'][f'k{k}']}")
        nets[k] = net
    cases = []
    centres, radii = fields.random_bubbles(64, seed=3, ndim=3)
    cases.append(("257^3, 64 bubbles, ratio 1000", lambda: ppe_case(N=256) + ((centres, radii, 1000.0),)))
    for N in (128, 256):
        for sd in held_out:
            cases.append((f"{N + 1}^3 held-out seed {sd}",
                          (lambda N=N, sd=sd: L.make_problem(N, sd, return_bubbles=True))))
    for name, mk in cases:
        g, rho, rhs, (cen, rad, ratio) = mk()
        row = {"case": name, "bubbles": int(len(rad)), "density_ratio": round(float(ratio), 1)}
        row["no_space"] = mg_native_timed(g, rho, rhs, "W")
        for k in ks:
            net = nets[k]
            row[f"trained_k{k}"] = deflated_native_timed(
                g, rho, rhs, lambda H, lv, net=net: L.LearnedDeflationPreconditioner(
                    H, net, lv, nu=2, gamma=2))
            row[f"trained_k{k}_level_E"] = deflated_native_timed(
                g, rho, rhs, lambda H, lv, net=net: L.LearnedDeflationPreconditioner(
                    H, net, lv, nu=2, gamma=2, galerkin="level"))
        unt = L.UNet(seed=5, k_out=max(ks)).cuda()
        row[f"untrained_k{max(ks)}"] = deflated_native_timed(
            g, rho, rhs, lambda H, lv: L.LearnedDeflationPreconditioner(H, unt, lv, nu=2, gamma=2))
        kb = min(64, len(rad))
        row[f"bubble_indicators_k{kb}"] = deflated_native_timed(
            g, rho, rhs, lambda H, lv: L.LearnedDeflationPreconditioner(
                H, None, lv, nu=2, gamma=2,
                vectors=bubble_indicator_vectors(H, lv, cen, rad, kb)))
        log(f"  {json.dumps(row)}")
        res["cases"].append(row)
        del g, rho, rhs
    return res


def ccd_batch_bench(reps=30):
    """48 independent 2-D wall problems, 16 each at 129^2, 257^2 and 513^2
    nodes (the uneven-batch sizes), all axes: one batched call (one grid per
    launch shape) against one tpf_ccd_apply per problem."""
    import numpy as np
    import torch
    from twophaseflow_amd import CCDBatch, CCDSolver, Grid, GridConfig
    rng = np.random.default_rng(5)
    solvers, fs, d1s, d2s = [], [], [], []
    for N in (128, 256, 512):
        for _ in range(16):
            g = Grid(GridConfig(ndim=2, N=(N, N), L=(1.0, 1.0)))
            solvers.append(CCDSolver(g, bc_type="wall"))
            f = torch.from_numpy(rng.standard_normal((N + 1, N + 1))).cuda()
            fs.append(f)
            d1s.append(torch.empty((2,) + tuple(f.shape), dtype=torch.float64, device="cuda"))
            d2s.append(torch.empty_like(d1s[-1]))
    batch = CCDBatch(solvers)
    pts = sum(2 * f.numel() for f in fs)

    def timed(fn, warm=5, n=reps):
        """Median and mean of per-call HIP-event times after `warm` calls
        (single calls here have shown rare multi-ms stalls of the stream)."""
        for _ in range(warm):
            fn()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(n)]
        for a, b in ev:
            a.record()
            fn()
            b.record()
        torch.cuda.synchronize()
        ts = sorted(a.elapsed_time(b) for a, b in ev)
        return ts[len(ts) // 2], sum(ts) / len(ts)

    def loop():
        for s_, f, a, b in zip(solvers, fs, d1s, d2s):
            s_.apply_into(f, a, b, 0b11)

    t_b, m_b = timed(lambda: batch.apply_into(fs, d1s, d2s))
    t_l, m_l = timed(loop)
    return {"problems": 48, "sizes": "16 x 129^2, 16 x 257^2, 16 x 513^2, wall, both axes",
            "batched_ms": round(t_b, 4), "per_problem_launches_ms": round(t_l, 4),
            "timing": "median of per-call HIP-event times after 5 warm-up calls",
            "mean_ms": {"batched": round(m_b, 4), "per_problem_launches": round(m_l, 4)},
            "batched_Mpts_per_s": round(pts / (t_b * 1e-3) / 1e6, 1),
            "algorithmic_GBps": round(40 * (pts / 2) / (t_b * 1e-3) / 1e9, 1),
            "launches": {"batched": 3, "per_problem": 48}}


def copy_bandwidth_GBps(reps=10):
    """Device-to-device copy of 1 GiB (read + write counted), event-timed."""
    import torch
    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return 2 * src.numel() * 8 / (ms * 1e-3) / 1e9


def _cpu_cg(rho, rhs, tol, maxiter, progress_every=0, label=""):
    """The oracle's SciPy Jacobi-CG on host arrays: (seconds, iterations,
    FP64 true relres).  progress_every > 0 prints a line every that many
    iterations (long solves)."""
    import numpy as np
    import scipy.sparse.linalg as spla
    from oracle import ppe as oppe
    N = rho.shape[0] - 1
    o = oppe.OraclePPE((N,) * 3)
    o.prepare_operator(rho)
    n = rho.size
    A = spla.LinearOperator((n, n), matvec=o.system_matvec, dtype=np.float64)
    M = spla.LinearOperator((n, n), matvec=o.system_precond, dtype=np.float64)
    it = [0]
    t0 = time.perf_counter()

    def cb(_xk):
        it[0] += 1
        if progress_every and it[0] % progress_every == 0:
            log(f"  CPU CG {label}: {it[0]} iterations, {time.perf_counter() - t0:.0f} s")
    x0 = np.zeros(n)
    sol, info = spla.cg(A, o.system_rhs(rhs), x0=x0, M=M, maxiter=maxiter, atol=0.0, rtol=tol,
                        callback=cb)
    sec = time.perf_counter() - t0
    p = sol.reshape(rho.shape)
    p.ravel()[o.pin] = 0.0
    return sec, it[0], float(o.relres(rhs, p)), info == 0


def cpu_baseline_ppe(rho_dev, rhs_dev, gpu_iterations, k_iters=10, full=()):
    """The oracle's SciPy Jacobi-CG (the reference's CG form) on the host:
    K iterations on the 257^3 case with every core of the job's share and
    with one thread (s/iteration, extrapolated to the GPU's iteration count),
    and full solves to 1e-8 of the cases in `full`: (label, rho_dev, rhs_dev,
    gpu_result) tuples of the same construction at smaller N.  The CG's
    vector work and the matrix-free operator are NumPy elementwise code, so
    the BLAS thread count barely changes the time."""
    import numpy as np
    from threadpoolctl import threadpool_limits
    from oracle import ppe as oppe
    cores = host_cores()
    rho = rho_dev.cpu().numpy()
    rhs = rhs_dev.cpu().numpy()
    N = rho.shape[0] - 1
    per_it = {}
    for th in (cores, 1):
        o = oppe.OraclePPE((N,) * 3)
        with threadpool_limits(th):
            t0 = time.perf_counter()
            try:
                o.solve_cg(rhs, rho, tol=1e-30, maxiter=k_iters)
            except RuntimeError:
                pass  # stops at maxiter by design
            per_it[th] = (time.perf_counter() - t0) / k_iters
    out = {"kind": "port", "cores": cores,
           "sample": f"oracle SciPy Jacobi-CG, {k_iters} iterations on the {N + 1}^3 case "
                     f"(incl. operator setup)",
           "seconds_per_iteration": round(per_it[cores], 4),
           "one_thread_seconds_per_iteration": round(per_it[1], 4),
           "extrapolated_seconds_to_1e-8": round(per_it[cores] * gpu_iterations, 1),
           "extrapolation_iterations": gpu_iterations}
    for label, r_dev, b_dev, gpu in full:
        r = r_dev.cpu().numpy()
        b = b_dev.cpu().numpy()
        with threadpool_limits(cores):
            sec, its, rel, ok = _cpu_cg(r, b, 1e-8, 10 ** 6, progress_every=500, label=label)
        out[f"full_solve_{label}"] = {
            "seconds": round(sec, 2), "iterations": its, "true_relres": rel, "converged": ok,
            "gpu_seconds": round(gpu["seconds"], 4), "gpu_iterations": gpu["iterations"],
            "gpu_vs_cpu": round(sec / gpu["seconds"], 1)}
    return out


def unet_mfma(reps=10):
    """BF16 MFMA 3x3x3 convolution throughput against the dense BF16 peak, at
    the U-Net's coarse-level shape and at larger ones: the forward kernel, the
    input gradient (the forward kernel on dY with flipped, transposed weights)
    and the weight gradient, and the MFMA fraction of one forward + backward
    pass (3 x 2 x 27 Cin Cout n^3 flops over the three kernels' time)."""
    import torch
    from twophaseflow_amd import learned as L

    def ev(fn):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    out = []
    for (n, cin, cout) in ((33, 16, 16), (65, 64, 64), (129, 32, 32)):
        x = torch.randn((n, n, n, cin), device="cuda").to(L.BF16).contiguous()
        gy = torch.randn((n, n, n, cout), device="cuda").to(L.BF16).contiguous()
        w = (torch.randn((27, cout, cin), device="cuda") * 0.05).to(L.BF16).contiguous()
        wf = L.flip_weights(w)
        b = torch.zeros(cout, device="cuda")
        gw = torch.zeros((27, cout, cin), device="cuda")
        t_f = ev(lambda: L.conv_fwd(x, w, b, True, False))
        t_d = ev(lambda: L.conv_fwd(gy, wf, None, False, False))
        t_w = ev(lambda: L.conv_wgrad(x, gy, gw))
        flops = 2.0 * 27 * cin * cout * n ** 3
        tf = flops / (t_f * 1e-3) / 1e12
        out.append({"grid": n, "cin": cin, "cout": cout, "ms": round(t_f, 4),
                    "TFLOPs": round(tf, 1), "frac_of_bf16_dense_peak": round(tf / 2500.0, 4),
                    "dgrad_ms": round(t_d, 4), "wgrad_ms": round(t_w, 4),
                    "wgrad_TFLOPs": round(flops / (t_w * 1e-3) / 1e12, 1),
                    "fwd_bwd_frac": round(3 * flops / ((t_f + t_d + t_w) * 1e-3) / 2.5e15, 4)})
    return out


def fccd_ppe_bench(reps=20, cpu=True):
    """FCCD matrix-free PPE (PPESolverFCCDMatrixFree, default configuration).

    Matvec at 257^3 on the C3 density (wall, 64 bubbles, ratio 1000): the CCD
    second derivative of every axis (d1 not written) followed by the fused
    face-gradient / divergence kernel; algorithmic bytes 32 + 48 = 80 B per
    node (CCD: read p, write 3 q; fused: read p, rho, q0..q2, write L p).
    Solves to 1e-8: GMRES(20) preconditioned by one FVM multigrid cycle per
    step at 257^3 and 65^3; Jacobi-GMRES(80) at 65^3 (restart 20, the
    reference's default, stalls there).  CPU: the oracle's apply on the 65^3
    case (one thread of NumPy/SciPy)."""
    import numpy as np
    import torch
    from twophaseflow_amd import (CCDSolver, PPESolverFCCDMatrixFree, SimulationConfig,
                                  SolverConfig)

    def solver(g, tol=1e-8, maxiter=1000, restart=80, pre="jacobi", coupling=None):
        cfg = SimulationConfig(solver=SolverConfig(pseudo_tol=tol, pseudo_maxiter=maxiter))
        cfg.solver.ppe_preconditioner = pre
        cfg.solver.ppe_restart = restart
        if coupling is not None:
            cfg.solver.ppe_coefficient_scheme = "phase_separated"
            cfg.solver.ppe_interface_coupling_scheme = coupling
        return PPESolverFCCDMatrixFree(None, cfg, g, None, bc_type="wall")

    def solve(s, rhs, rho, reps=2):
        x = torch.empty_like(rhs)
        for _ in range(reps):  # the last of reps (the first builds graphs / hierarchy)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            d = s.solve_into(rhs, rho, x)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
        return x, {"seconds": round(t, 4), "iterations": d["iterations"],
                   "restart_cycles": d["restart_cycles"],
                   "relative_l2": d["ppe_linear_relative_l2"]}

    def events(fn, n):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(n)]
        fn()
        for a, b in ev:
            a.record()
            fn()
            b.record()
        torch.cuda.synchronize()
        return sorted(a.elapsed_time(b) for a, b in ev)[n // 2]

    out = {}
    g, rho, rhs = ppe_case(N=256)
    s = solver(g)
    s.prepare_operator(rho)
    p = torch.randn_like(rho)
    Lp = torch.empty_like(p)
    npts = p.numel()
    ms_apply = events(lambda: s.apply_into(p, Lp), reps)
    ccd = CCDSolver(g, bc_type="wall")
    q = torch.empty((3,) + tuple(p.shape), dtype=torch.float64, device="cuda")
    ms_ccd = events(lambda: ccd.apply_into(p, None, q, 0b111), reps)
    ms_fused = ms_apply - ms_ccd
    _, out["solve_257_ratio1000_mg"] = solve(solver(g, maxiter=100, restart=20, pre="mg"),
                                             rhs, rho)
    # the reference's defect correction (ppe/defect_correction.py) with this
    # FCCD operator over the FVM multigrid-FCG solve as L_L
    from twophaseflow_amd import PPESolverFVMMatrixFree
    from twophaseflow_amd.defect_correction import PPESolverDefectCorrection
    bcfg = SimulationConfig(solver=SolverConfig(pseudo_tol=1e-8, pseudo_maxiter=200))
    bcfg.solver.ppe_iteration_method = "cg"
    bcfg.solver.ppe_preconditioner = "mg"
    dc = PPESolverDefectCorrection(None, g, PPESolverFVMMatrixFree(None, bcfg, g),
                                   solver(g, pre="mg"), max_corrections=40, tolerance=1e-8)
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dc.solve(rhs, rho)
        torch.cuda.synchronize()
        t_dc = time.perf_counter() - t0
    dd = dc.last_diagnostics
    out["solve_257_ratio1000_defect_correction"] = {
        "seconds": round(t_dc, 4), "corrections": int(dd["ppe_dc_iterations"]),
        "relative_l2": dd["ppe_dc_final_relative_l2"], "converged": bool(dd["ppe_dc_converged"]),
        "base": "FVM multigrid-FCG, tol 1e-8"}
    del dc
    # the pipeline's production configuration (ns_pipeline.py:195-196):
    # phase_separated coefficients with the affine cut-face jump; psi from the
    # bubbles' density, kappa_lg = -1/0.05 (gas spheres of radius ~0.05),
    # sigma 0.01
    psi = (rho - 1e-3) / (1.0 - 1e-3)
    kappa = torch.full_like(rho, -20.0)
    sa = solver(g, maxiter=100, restart=20, pre="mg", coupling="affine_jump")
    sa.set_interface_jump_context(psi=psi, kappa=kappa, sigma=0.01)
    sa.prepare_operator(rho)
    ms_affine = events(lambda: sa.apply_into(p, Lp), reps)
    _, out["solve_257_ratio1000_affine_jump_mg"] = solve(sa, rhs, rho)
    out["solve_257_ratio1000_affine_jump_mg"]["apply_ms"] = round(ms_affine, 4)
    del sa, psi, kappa
    out["apply_257"] = {
        "ms": round(ms_apply, 4), "timing": "median of per-call HIP events",
        "algorithmic_GBps": round(80 * npts / (ms_apply * 1e-3) / 1e9, 1),
        "ccd_q_only_ms": round(ms_ccd, 4),
        "ccd_q_only_GBps": round(32 * npts / (ms_ccd * 1e-3) / 1e9, 1),
        "fused_face_divergence_ms": round(ms_fused, 4),
        "fused_face_divergence_GBps": round(48 * npts / (ms_fused * 1e-3) / 1e9, 1)}
    del s, q, p, Lp, rho, rhs
    torch.cuda.empty_cache()
    g, rho, rhs = ppe_case(N=64)
    _, out["solve_65_ratio1000_mg"] = solve(solver(g, maxiter=100, restart=20, pre="mg"),
                                            rhs, rho)
    s = solver(g)
    x, out["solve_65_ratio1000_jacobi_gmres80"] = solve(s, rhs, rho, reps=1)
    y = torch.empty_like(x)
    ms65 = events(lambda: s.apply_into(x, y), reps)
    out["apply_65_ms"] = round(ms65, 4)
    if cpu:
        from oracle.fccd import OracleFCCDPPE
        o = OracleFCCDPPE((64,) * 3, (1.0,) * 3, "wall")
        o.prepare_operator(rho.cpu().numpy())
        xh = x.cpu().numpy()
        t0 = time.perf_counter()
        for _ in range(3):
            o.apply(xh)
        cpu_s = (time.perf_counter() - t0) / 3
        out["cpu_oracle_apply_65"] = {"seconds": round(cpu_s, 4), "kind": "port",
                                      "sample": "oracle FCCD apply (NumPy/SciPy dense-LU CCD) "
                                                "on the 65^3 case, 3 calls",
                                      "gpu_speedup": round(cpu_s / (ms65 * 1e-3), 1)}
    return out


def rhie_chow_bench(reps=20):
    """Rhie-Chow face-velocity divergence (the PPE right-hand side of the
    projection step) on the 257^3 wall case: one all-axes CCD launch for the
    pressure gradient (d1 only) plus the fused face-flux divergence kernel.
    Algorithmic bytes of the kernel: u (3), p, rho, dp (3) in, div out =
    72 B per node."""
    import numpy as np
    import torch
    from twophaseflow_amd import CCDSolver
    from twophaseflow_amd.rhie_chow import RhieChowInterpolator
    g, rho, rhs = ppe_case(N=256)
    ccd = CCDSolver(g, bc_type="wall")
    rc = RhieChowInterpolator(None, g, ccd, "wall")
    rng = np.random.default_rng(9)
    vel = [torch.from_numpy(rng.standard_normal(tuple(rho.shape))).cuda() for _ in range(3)]
    p = torch.randn_like(rho)
    rc.face_velocity_divergence(vel, p, rho, 0.01)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for a, b in ev:
        a.record()
        rc.face_velocity_divergence(vel, p, rho, 0.01)
        b.record()
    torch.cuda.synchronize()
    ms = sorted(a.elapsed_time(b) for a, b in ev)[reps // 2]
    npts = p.numel()
    return {"ms": round(ms, 4), "timing": "median of per-call HIP events (CCD d1 + divergence "
            "kernel + the call's allocations)", "Mpts_per_s": round(npts / (ms * 1e-3) / 1e6, 1),
            "algorithmic_GBps_incl_ccd": round((72 + 32) * npts / (ms * 1e-3) / 1e9, 1)}


def dccd_bench(reps=10):
    """One dissipative-CCD TVD-RK3 level-set step on 257^3 periodic (64
    bubbles, a seeded velocity): 3 stages x (flux kernel, 3 CCD launches,
    fused filter / stage kernel)."""
    import numpy as np
    import torch
    from twophaseflow_amd import CCDSolver, Grid, GridConfig, fields
    from twophaseflow_amd.dccd import DissipativeCCDAdvection
    N = 256
    g = Grid(GridConfig(ndim=3, N=(N,) * 3, L=(1.0,) * 3))
    c, r = fields.random_bubbles(64, seed=3, ndim=3)
    psi, _ = fields.bubbles_psi_rho(g.N, g.L, c, r, 1.5 / N, 1.0, 1e-3)
    rng = np.random.default_rng(11)
    vel = [torch.from_numpy(0.5 * rng.standard_normal(g.shape)).cuda() for _ in range(3)]
    adv = DissipativeCCDAdvection(None, g, CCDSolver(g, bc_type="periodic"), bc="periodic")
    adv.advance(psi, vel, 1e-3)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for a, b in ev:
        a.record()
        adv.advance(psi, vel, 1e-3)
        b.record()
    torch.cuda.synchronize()
    ms = sorted(a.elapsed_time(b) for a, b in ev)[reps // 2]
    return {"ms_per_step": round(ms, 4), "timing": "median of per-call HIP events",
            "Mpts_steps_per_s": round(psi.numel() / (ms * 1e-3) / 1e6, 1)}


def curvature_bench(reps=10):
    """CurvatureCalculatorPsi on the 257^3 periodic bubble field (64 bubbles),
    DCCD filter on (eps 0.05): CCD d1/d2, 6 filtered fields, 3 mixed CCD
    derivatives, the curvature kernel."""
    import torch
    from twophaseflow_amd import CCDSolver, CurvatureCalculatorPsi, Grid, GridConfig, fields
    N = 256
    g = Grid(GridConfig(ndim=3, N=(N,) * 3, L=(1.0,) * 3))
    c, r = fields.random_bubbles(64, seed=3, ndim=3)
    psi, _ = fields.bubbles_psi_rho(g.N, g.L, c, r, 1.5 / N, 1.0, 1e-3)
    calc = CurvatureCalculatorPsi(None, CCDSolver(g, bc_type="periodic"))
    calc.compute(psi)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for a, b in ev:
        a.record()
        calc.compute(psi)
        b.record()
    torch.cuda.synchronize()
    ms = sorted(a.elapsed_time(b) for a, b in ev)[reps // 2]
    return {"ms": round(ms, 4), "timing": "median of per-call HIP events (incl. the call's "
            "allocations)"}


def reinit_bench(reps=5):
    """SplitReinitializer (4 pseudo-steps, mass correction) on the 257^3 wall
    bubble field, eps = 1.5 h."""
    import torch
    from twophaseflow_amd import CCDSolver, Grid, GridConfig, SplitReinitializer, fields
    N = 256
    g = Grid(GridConfig(ndim=3, N=(N,) * 3, L=(1.0,) * 3))
    c, r = fields.random_bubbles(64, seed=3, ndim=3)
    psi, _ = fields.bubbles_psi_rho(g.N, g.L, c, r, 1.5 / N, 1.0, 1e-3)
    ri = SplitReinitializer(None, g, CCDSolver(g, bc_type="wall"), 1.5 / N)
    ri.reinitialize(psi)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    for a, b in ev:
        a.record()
        ri.reinitialize(psi)
        b.record()
    torch.cuda.synchronize()
    ms = sorted(a.elapsed_time(b) for a, b in ev)[reps // 2]
    return {"ms": round(ms, 3), "pseudo_steps": 4, "timing": "median of per-call HIP events"}


def legacy_config1_bench(steps=20, cpu=True):
    """The 2-D 64 x 64 legacy rising-bubble loop (config 1, extension
    "none"): CFL time step, DCCD advection + split reinit, properties,
    curvature, UCCD6 / Picard-CN predictor, Rhie-Chow right-hand side,
    Jacobi-PCG PPE to 1e-10, corrector, non-finite check; `steps` steps timed
    by the host clock after a 2-step warm-up on another simulation object.
    The oracle's run of the same steps on the host is the CPU side."""
    import numpy as np
    import torch
    from twophaseflow_amd import (FluidConfig, GridConfig, NumericsConfig, SimulationBuilder,
                                  SimulationConfig, SolverConfig)
    n = 64

    def cfg():
        c = SimulationConfig(grid=GridConfig(ndim=2, N=(n, n), L=(1.0, 1.0)),
                             fluid=FluidConfig(Re=100.0, Fr=1.0, We=10.0, rho_ratio=0.1,
                                               mu_ratio=0.1),
                             numerics=NumericsConfig(advection_scheme="dissipative_ccd",
                                                     cfl_number=0.3, extension_method="none"),
                             solver=SolverConfig(ppe_solver_type="fvm_iterative",
                                                 pseudo_tol=1e-10))
        c.solver.ppe_iteration_method = "cg"
        return c

    x = np.linspace(0.0, 1.0, n + 1)
    X, Y = np.meshgrid(x, x, indexing="ij")
    psi0 = 1.0 / (1.0 + np.exp(-(np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2) - 0.25) / (1.5 / n)))

    def run(k):
        sim = SimulationBuilder(cfg()).build()
        sim.psi.data = psi0
        sim._update_properties()
        sim._update_curvature()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its = 0
        for _ in range(k):
            dt = sim.cfl_calc.compute([sim.velocity[a] for a in range(2)], sim.mu.data,
                                      sim.rho.data)
            sim._step_with_retry(dt)
            its += sim.ppe_solver.last_diagnostics["iterations"]
        torch.cuda.synchronize()
        return time.perf_counter() - t0, sim, its

    run(2)
    sec, sim, its = run(steps)
    out = {"steps": steps, "seconds": round(sec, 4), "steps_per_s": round(steps / sec, 2),
           "ms_per_step": round(sec / steps * 1e3, 3), "pcg_iterations": its,
           "div_max_final": sim._diagnostics.divergence_max(sim),
           "mass_final": sim._diagnostics.volume(sim)}
    if cpu:
        from oracle import legacy_step as ls
        o = ls.OracleLegacySimulation((n, n), (1.0, 1.0))
        o.psi = psi0.copy()
        o.update_properties()
        o.update_curvature()
        t0 = time.perf_counter()
        for _ in range(steps):
            o.step_forward(o.cfl_dt())
        cs = time.perf_counter() - t0
        try:
            cores = len(os.sched_getaffinity(0))
        except Exception:
            cores = os.cpu_count() or 1
        out["cpu_baseline"] = {"kind": "port", "seconds": round(cs, 4),
                               "steps_per_s": round(steps / cs, 2), "cores": cores,
                               "sample": f"oracle legacy step (NumPy / SciPy, dense-LU CCD, "
                                         f"SciPy CG), the same {steps} steps"}
        out["gpu_vs_cpu"] = round(cs / sec, 2)
    return out


def uneven_bubble_batch_bench(steps=2, sizes=(128, 256, 512), per_size=16,
                              preconditioner="jacobi"):
    """The uneven batch of 48 seeded 2-D rising bubbles (16 each at 128^2,
    256^2, 512^2; radius 0.15-0.3, density ratio 10-1000, We 1-50) stepped
    together by BatchedSimulation: per simulation the CFL dt and steps 1-5,
    one batched solve of all 48 PPEs to 1e-10, then the corrector.  The PPE
    is the reference's Jacobi-PCG (one launch per PCG step over all problems,
    captured iteration chunks) or, with preconditioner="mg", multigrid FCG
    (the 48 solves advance together, one stream each).  One warm-up batched
    step, then `steps` batched steps timed by the host clock; aggregate
    simulation steps per second.  The same simulations then take one step
    each one after another (single PPE solves) for comparison."""
    import torch
    from twophaseflow_amd import BatchedSimulation, bubble_simulation
    from twophaseflow_amd.fields import uneven_bubble_batch
    cases = uneven_bubble_batch(seed=5, sizes=sizes, per_size=per_size)
    t0 = time.perf_counter()
    sims = [bubble_simulation(c, ppe_preconditioner=preconditioner) for c in cases]
    batch = BatchedSimulation(sims)
    batch.initialize()
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    batch.step()  # warm-up: graph capture, hierarchies, first-touch allocations
    torch.cuda.synchronize()
    its = []
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.step()
        its.append(max(s.ppe_solver.last_diagnostics["iterations"] for s in sims))
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    n = len(sims)
    t0 = time.perf_counter()
    for s in sims:
        dt = s.cfl_calc.compute([s.velocity[a] for a in range(2)], s.mu.data, s.rho.data)
        s._step_with_retry(dt)
    torch.cuda.synchronize()
    seq = time.perf_counter() - t0
    finite = all(not s._has_nonfinite_state() for s in sims)
    ppe = {"jacobi": "Jacobi-PCG to 1e-10 (CG form)",
           "mg": "multigrid FCG to 1e-10 (CG form, one W(2,2) FP32 cycle per iteration)"}
    return {"simulations": n, "grids": [f"{N + 1}^2" for N in sizes], "per_grid": per_size,
            "batched_steps": steps, "seconds": round(sec, 3),
            "sim_steps_per_s": round(n * steps / sec, 2),
            "ms_per_batched_step": round(sec / steps * 1e3, 1),
            "max_ppe_iterations_per_step": its, "retries": batch.retries,
            "sequential_one_step_each_s": round(seq, 3),
            "sequential_sim_steps_per_s": round(n / seq, 2),
            "batched_vs_sequential": round((n * steps / sec) / (n / seq), 2),
            "build_seconds": round(build_s, 1), "all_finite": finite,
            "ppe": ppe[preconditioner], "timing": "host clock, synchronized"}


def profile_mode(args):
    import torch
    from twophaseflow_amd import CCDSolver, Grid, GridConfig
    from twophaseflow_amd.fields import band_limited_periodic
    if args.profile == "ccd":
        N = 256
        f = torch.from_numpy(band_limited_periodic(N, 3, kmax=32, seed=2)).cuda()
        g = Grid(GridConfig(ndim=3, N=(N, N, N), L=(1.0, 1.0, 1.0)))
        ccd = CCDSolver(g, bc_type="periodic")
        d1 = torch.empty((3,) + g.shape, dtype=torch.float64, device="cuda")
        d2 = torch.empty_like(d1)
        for _ in range(args.steps):  # explicit grid (axes 0, 1) + wavefront kernel (axis 2)
            ccd.apply_into(f, d1, d2, 0b111)
        for _ in range(args.steps):  # single-axis calls
            for axis in range(3):
                ccd.differentiate_into(f, d1[axis], d2[axis], axis)
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"profile ccd: {args.steps} all-axes steps (explicit grid of axes 0-1, wavefront "
            f"kernel on axis 2) and {args.steps} x 3 single-axis launches on 257^3, then a "
            f"1 GiB copy")
    elif args.profile == "fccd":
        from twophaseflow_amd import PPESolverFCCDMatrixFree, SimulationConfig, SolverConfig
        gp, rho, rhs = ppe_case(N=256)
        cfg = SimulationConfig(solver=SolverConfig())
        s = PPESolverFCCDMatrixFree(None, cfg, gp, None, bc_type="wall")
        s.prepare_operator(rho)
        out = torch.empty_like(rhs)
        for _ in range(args.steps):
            s.apply_into(rhs, out)
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"profile fccd: {args.steps} FCCD PPE matvecs on 257^3, then a 1 GiB copy")
    elif args.profile == "n3":
        log(f"profile n3 Rhie-Chow 257^3: {rhie_chow_bench(reps=args.steps)}")
        log(f"profile n3 DCCD step 257^3: {dccd_bench(reps=args.steps)}")
        log(f"profile n3 psi curvature 257^3: {curvature_bench(reps=args.steps)}")
        log(f"profile n3 split reinit 257^3: {reinit_bench(reps=args.steps)}")
    elif args.profile == "legacy":
        log(f"profile legacy config-1 64^2, 2 + {args.steps} steps: "
            f"{legacy_config1_bench(steps=args.steps, cpu=False)}")
    elif args.profile == "batch":
        log(f"profile 48 uneven bubbles, 1 + {args.steps} batched steps: "
            f"{uneven_bubble_batch_bench(steps=args.steps)}")
    elif args.profile == "batch_mg":
        log(f"profile 48 uneven bubbles, multigrid FCG, 1 + {args.steps} batched steps: "
            f"{uneven_bubble_batch_bench(steps=args.steps, preconditioner='mg')}")
    elif args.profile == "ppe":
        gp, rho, rhs = ppe_case(N=128)
        r = ppe_solve_timed(gp, rho, rhs, "fp64")
        log(f"profile ppe 129^3: {r}")
    else:
        gp, rho, rhs = ppe_case(N=256)
        r = mg_native_timed(gp, rho, rhs, "W", reps=1)
        log(f"profile mg 257^3 native W(2,2)-FCG: {r}")


def cpu_protocol():
    """Full CPU PPE solves of the 65^3 and 129^3 cases (64 bubbles, ratio
    1000, the 257^3 case's construction) to ||r||/||b|| <= 1e-8 with the
    oracle's SciPy Jacobi-CG, beside the GPU's FP64 PCG on the same inputs;
    65^3 also with one thread.  Prints one JSON line."""
    from threadpoolctl import threadpool_limits
    cores = host_cores()
    out = {"cores": cores}
    for N in (64, 128):
        g, rho, rhs = ppe_case(N=N)
        gpu = ppe_solve_timed(g, rho, rhs, "fp64")
        r, b = rho.cpu().numpy(), rhs.cpu().numpy()
        log(f"CPU CG {N + 1}^3 on {cores} cores (GPU: {gpu['iterations']} iterations, "
            f"{gpu['seconds']:.3f} s)")
        with threadpool_limits(cores):
            sec, its, rel, ok = _cpu_cg(r, b, 1e-8, 10 ** 6, progress_every=250,
                                        label=f"{N + 1}^3")
        entry = {"cpu_seconds": round(sec, 2), "cpu_iterations": its, "cpu_true_relres": rel,
                 "cpu_converged": ok, "gpu_seconds": round(gpu["seconds"], 4),
                 "gpu_iterations": gpu["iterations"], "gpu_true_relres": gpu["true_relres"],
                 "gpu_vs_cpu": round(sec / gpu["seconds"], 1)}
        if N == 64:
            log("CPU CG 65^3 on one thread")
            with threadpool_limits(1):
                sec1, its1, _, _ = _cpu_cg(r, b, 1e-8, 10 ** 6, progress_every=250,
                                           label="65^3 one thread")
            entry["cpu_one_thread_seconds"] = round(sec1, 2)
            entry["cpu_one_thread_iterations"] = its1
        out[f"{N + 1}^3"] = entry
        log(f"  {entry}")
        del rho, rhs
    print(json.dumps({"cpu_protocol": out}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 200 steps of 0.3 ms: the fixed cost of the two synchronisations around
    # the timed region (~1 ms) stays under 2 % of it
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-ppe", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ppe-n", type=int, default=256)
    ap.add_argument("--train-problems", type=int, default=100,
                    help="random 129^3 bubble problems used to train the learned coarse space")
    ap.add_argument("--profile", choices=["ccd", "ppe", "mg", "fccd", "n3", "legacy", "batch",
                                         "batch_mg"],
                    default=None,
                    help="run only one workload (for rocprofv3): 'ccd' = K all-axes CCD steps "
                         "on 257^3 plus one 1 GiB device copy (counter scale reference); "
                         "'ppe' = one FP64 PCG solve of the 129^3 ratio-1000 case; "
                         "'mg' = one W(2,2) multigrid-FCG solve of the 257^3 case; "
                         "'fccd' = K FCCD PPE matvecs on the 257^3 case; "
                         "'n3' = K calls each of the Rhie-Chow divergence, the DCCD step, "
                         "the psi curvature and the split reinitialisation on 257^3; "
                         "'legacy' = 2 + K steps of the 2-D 64^2 legacy rising-bubble loop; "
                         "'batch' = 1 + K batched steps of the 48 uneven 2-D bubbles "
                         "(Jacobi-PCG PPE); 'batch_mg' = the same with multigrid FCG")
    ap.add_argument("--cpu-protocol", action="store_true",
                    help="only the CPU PPE protocol: full oracle Jacobi-CG solves to 1e-8 of the "
                         "65^3 and 129^3 ratio-1000 cases (all cores of the job's share, then "
                         "one thread at 65^3) beside the GPU's PCG; minutes of CPU time")
    ap.add_argument("--learned-study", action="store_true",
                    help="only the learned coarse-space study: k = 16 and 64 networks trained on "
                         "--train-problems problems, evaluated at 129^3 and 257^3 against the "
                         "same device multigrid-FCG without them, per-bubble indicator vectors "
                         "and untrained networks (one JSON line; about ten minutes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from twophaseflow_amd import CCDSolver, Grid, GridConfig
    from twophaseflow_amd.fields import band_limited_periodic
    torch.cuda.set_device(0)
    if args.profile:
        return profile_mode(args)
    if args.cpu_protocol:
        return cpu_protocol()
    if args.learned_study:
        print(json.dumps({"learned_study": learned_study(args.train_problems)}), flush=True)
        return None

    # ---- CCD 3-D 257^3 periodic, all axes, FP64
    N = 256
    log("generating 257^3 band-limited field")
    f = torch.from_numpy(band_limited_periodic(N, 3, kmax=32, seed=2)).cuda()
    g = Grid(GridConfig(ndim=3, N=(N, N, N), L=(1.0, 1.0, 1.0)))
    ccd = CCDSolver(g, bc_type="periodic")
    d1 = torch.empty((3,) + g.shape, dtype=torch.float64, device="cuda")
    d2 = torch.empty_like(d1)
    log("timing 3-D CCD")
    sec = time_ccd(ccd, f, d1, d2, 0b111, args.steps, args.warmup, 3)
    npts = f.numel()
    mpts = 3 * npts / sec / 1e6
    axes = per_axis_event_times(ccd, f, d1, d2, max(args.steps, 5))
    # One all-axes step is two launches: the explicit partition-inverse grid of
    # the strided axes 0 and 1 (dominant), then the wavefront kernel on the
    # contiguous axis 2.  Algorithmic bytes: 8 (1 + 2 x 2) B = 40 B per point
    # for the strided grid (f read once, f' and f'' of two axes written),
    # 24 B per point for axis 2.  Mean launch times from HIP events on torch's
    # stream, the stream both kernels run on.
    reps = max(args.steps, 5)
    dom_ms = event_ms(lambda: ccd.apply_into(f, d1[:2], d2[:2], 0b011), reps)
    ax2_ms = event_ms(lambda: ccd.differentiate_into(f, d1[2], d2[2], 2), reps)
    step_ms = event_ms(lambda: ccd.apply_into(f, d1, d2, 0b111), reps)
    bytes_per_launch = 8 * (1 + 2 * 2) * npts
    achieved = bytes_per_launch / (dom_ms * 1e-3) / 1e9
    all_axes_gbs = 8.0 * 7 * npts / sec / 1e9
    strided = {"axes": [0, 1], "kernel": "ccd_explicit_kernel<16, 8, true> (one grid)",
               "ms": round(dom_ms, 4),
               "algorithmic_GBps": round(8 * (1 + 2 * 2) * npts / (dom_ms * 1e-3) / 1e9, 1)}
    step_kernels = [
        {"kernel": "ccd_explicit_kernel<16, 8, true>", "axes": [0, 1], "ms": round(dom_ms, 4),
         "bytes": bytes_per_launch},
        {"kernel": "ccd_wave_kernel<16, 4>", "axes": [2], "ms": round(ax2_ms, 4),
         "bytes": 24 * npts,
         "algorithmic_GBps": round(24 * npts / (ax2_ms * 1e-3) / 1e9, 1)},
        {"all_axes_step_ms_hip_events": round(step_ms, 4),
         "algorithmic_GBps": round(56 * npts / (step_ms * 1e-3) / 1e9, 1)}]
    log("timing 3-D CCD, FP32 solves + FP64 refinement")
    ccd_r = CCDSolver(g, bc_type="periodic", precision="fp32_refined")
    r1 = torch.empty_like(d1)
    r2 = torch.empty_like(d2)
    sec_r = time_ccd(ccd_r, f, r1, r2, 0b111, args.steps, args.warmup, 3)
    refined = {"Mpts_per_s": round(3 * npts / sec_r / 1e6, 2),
               "ms_per_step": round(sec_r * 1e3, 4),
               "max_rel_err_vs_fp64": {
                   "d1": float((r1 - d1).abs().max() / d1.abs().max()),
                   "d2": float((r2 - d2).abs().max() / d2.abs().max())}}
    del d1, d2, r1, r2

    # ---- CCD 2-D 1025^2 periodic, all axes
    log("timing 2-D CCD")
    N2 = 1024
    f2 = torch.from_numpy(band_limited_periodic(N2, 2, kmax=32, seed=2)).cuda()
    g2 = Grid(GridConfig(ndim=2, N=(N2, N2), L=(1.0, 1.0)))
    ccd2 = CCDSolver(g2, bc_type="periodic")
    e1 = torch.empty((2,) + g2.shape, dtype=torch.float64, device="cuda")
    e2 = torch.empty_like(e1)
    sec2 = time_ccd(ccd2, f2, e1, e2, 0b11, args.steps, args.warmup, 2)
    mpts2 = 2 * f2.numel() / sec2 / 1e6
    log("timing 2-D CCD, FP32 solves + FP64 refinement")
    ccd2r = CCDSolver(g2, bc_type="periodic", precision="fp32_refined")
    q1 = torch.empty_like(e1)
    q2 = torch.empty_like(e2)
    sec2r = time_ccd(ccd2r, f2, q1, q2, 0b11, args.steps, args.warmup, 2)
    refined2 = {"Mpts_per_s": round(2 * f2.numel() / sec2r / 1e6, 2),
                "ms_per_step": round(sec2r * 1e3, 4),
                "kernel": "ccd_wave_kernel_ref<16, 1> (64 partitions x 16 rows, two refinements)",
                "max_rel_err_vs_fp64": {
                    "d1": float((q1 - e1).abs().max() / e1.abs().max()),
                    "d2": float((q2 - e2).abs().max() / e2.abs().max())}}
    del e1, e2, q1, q2

    result = {
        "metric": METRIC,
        "value": round(mpts, 2),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded band-limited periodic field, |k|<=32, amp |k|^-2)",
        "config": {"workload": "CCD f',f'' all axes, 3-D 257^3 nodes periodic, FP64",
                   "grid": [N + 1] * 3, "axes": 3, "lines_per_axis": (N + 1) ** 2},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counter runs
                     # recorded in profiles/ccd_all_axes_257cubed_periodic_fp64.md
                     # (scale factors 2 and 1 from a 1 GiB copy in the same runs)
                     "traffic": TRAFFIC_BYTES_CCD,
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE (profiles/)",
                     "kernel": "ccd_explicit_kernel<16, 8, true> (one grid, strided axes 0 and 1)",
                     "bytes_per_launch": bytes_per_launch, "launch_ms": round(dom_ms, 4)},
        "ccd_step_kernels": step_kernels,
        "hbm_copy_GBps": round(copy_bandwidth_GBps(), 1),
        "ccd_all_axes_GBps": round(all_axes_gbs, 1),
        "ccd_axes": [{"axis": a["axis"], "ms": round(a["ms"], 4), "GBps": round(a["GBps"], 1),
                      "kernel": a["info"]["kernel"]} for a in axes],
        "ccd_strided_axes_fused": strided,
        "ccd_2d_1025": {"Mpts_per_s": round(mpts2, 2), "ms_per_step": round(sec2 * 1e3, 4)},
        "ccd_2d_1025_fp32_refined": refined2,
        "ccd_fp32_refined_257": refined,
        "ccd_batch_uneven_2d": ccd_batch_bench(),
    }

    if not args.no_ppe:
        log(f"building PPE case {args.ppe_n}^3")
        gp, rho, rhs = ppe_case(N=args.ppe_n)
        res = {}
        for key, prec, itol in (("fp64", "fp64", 0.0), ("mixed", "mixed", 1e-4),
                                ("mixed_1e-3", "mixed", 1e-3), ("mixed_1e-5", "mixed", 1e-5)):
            log(f"PPE solve {key}")
            res[key] = ppe_solve_timed(gp, rho, rhs, prec, inner_tol=itol)
            log(f"  {res[key]}")
        for key, cyc in (("mg_fcg_native_W22", "W"), ("mg_fcg_native_V22", "V")):
            log(f"PPE solve {key}")
            res[key] = mg_native_timed(gp, rho, rhs, cyc)
            log(f"  {res[key]}")
        for key, gam in (("mg_fcg_V22", 1), ("mg_fcg_W22", 2)):
            log(f"PPE solve {key} (host FCG loop)")
            res[key] = mg_solve_timed(gp, rho, rhs, gamma=gam)
            log(f"  {res[key]}")
        import numpy as np
        from twophaseflow_amd import learned as L
        log(f"training the learned coarse space on {args.train_problems} problems (129^3)")
        torch.manual_seed(0)
        net = L.UNet(seed=0).cuda()
        tr = L.train_deflation(net, N=128, problems=args.train_problems, lr=1e-3, seed=1000,
                               gamma=2)
        log(f"  trained in {tr['seconds']:.1f} s")
        res["learned"] = mg_solve_timed(gp, rho, rhs, gamma=2, net=net)
        log(f"  learned: {res['learned']}")
        res["untrained"] = mg_solve_timed(gp, rho, rhs, gamma=2, net=L.UNet(seed=5).cuda())
        log("learned coarse space in the device FCG")
        res["learned_native"] = learned_native_timed(gp, rho, rhs, net)
        log(f"  learned (device FCG): {res['learned_native']}")
        lh = tr["loss_history"]
        pcg_s = res["fp64"]["seconds"]
        result["ppe"] = {
            "grid": [args.ppe_n + 1] * 3, "density_ratio": 1000, "bubbles": 64, "tol": 1e-8,
            "plain_pcg_fp64": res["fp64"], "pcg_mixed": res["mixed"],
            "pcg_mixed_inner_1e-3": res["mixed_1e-3"], "pcg_mixed_inner_1e-5": res["mixed_1e-5"],
            "mg_fcg_native_W22": res["mg_fcg_native_W22"],
            "mg_fcg_native_V22": res["mg_fcg_native_V22"],
            "speedup_native_mg_vs_plain_pcg": round(pcg_s / res["mg_fcg_native_W22"]["seconds"], 2),
            "host_loop_mg_fcg_V22": res["mg_fcg_V22"], "host_loop_mg_fcg_W22": res["mg_fcg_W22"],
            "learned_fcg": dict(res["learned"], backbone="W(2,2) multigrid",
                                coarse_space="16 U-Net vectors on the 33^3 level",
                                train_problems=args.train_problems,
                                train_seconds=round(tr["seconds"], 1),
                                train_loss_first10=float(np.mean(lh[:10])),
                                train_loss_last10=float(np.mean(lh[-10:]))),
            "untrained_net_fcg": res["untrained"],
            "speedup_learned_vs_plain_pcg": round(pcg_s / res["learned"]["seconds"], 2),
            "speedup_mg_W22_vs_plain_pcg": round(pcg_s / res["mg_fcg_W22"]["seconds"], 2),
            "learned_vs_same_cycle_without_net": round(
                res["mg_fcg_W22"]["seconds"] / res["learned"]["seconds"], 3),
            "learned_fcg_device": dict(res["learned_native"],
                                       backbone="W(2,2), W on the top 2 levels (device FCG)"),
            "speedup_learned_device_vs_plain_pcg": round(
                pcg_s / res["learned_native"]["seconds"], 2),
            "learned_device_vs_same_cycle_without_net": round(
                res["mg_fcg_native_W22"]["seconds"] / res["learned_native"]["seconds"], 3),
        }
        if not args.no_cpu:
            log("CPU baseline (oracle SciPy Jacobi-CG: 10 iterations at 257^3 with all cores "
                "and one thread; full solve at 65^3)")
            gs, rhos, rhss = ppe_case(N=64)
            gpu65 = ppe_solve_timed(gs, rhos, rhss, "fp64")
            result["ppe"]["plain_pcg_fp64_65"] = gpu65
            result["ppe"]["cpu_baseline"] = cpu_baseline_ppe(
                rho, rhs, res["fp64"]["iterations"], full=[("65^3", rhos, rhss, gpu65)])
            del rhos, rhss
        del rho, rhs
        log("U-Net BF16 MFMA convolution throughput")
        result["unet_conv_mfma"] = unet_mfma()
        log("GMRES + line_pcr (reference default route) on 65^3")
        gq, rhoq, rhsq = ppe_case(N=64)
        result["ppe"]["gmres_line_pcr_65"] = ppe_solve_timed(gq, rhoq, rhsq, "fp64", maxiter=500,
                                                             method="gmres")
        del rhoq, rhsq
        log("split reinitialisation, 257^3 wall")
        result["reinit_split_257"] = reinit_bench()
        log(f"  {result['reinit_split_257']}")
        log("psi curvature, 257^3 periodic")
        result["curvature_psi_257"] = curvature_bench()
        log(f"  {result['curvature_psi_257']}")
        log("DCCD level-set TVD-RK3 step, 257^3 periodic")
        result["dccd_advection_257"] = dccd_bench()
        log(f"  {result['dccd_advection_257']}")
        log("Rhie-Chow face-velocity divergence, 257^3 wall")
        result["rhie_chow_257"] = rhie_chow_bench()
        log(f"  {result['rhie_chow_257']}")
        log("FCCD matrix-free PPE: 257^3 matvec, 65^3 Jacobi-GMRES(80) solve")
        result["fccd_ppe"] = fccd_ppe_bench(cpu=not args.no_cpu)
        log(f"  {result['fccd_ppe']}")
        log("legacy config-1 loop: 2-D 64^2 rising bubble, 20 steps")
        result["legacy_config1_64"] = legacy_config1_bench(cpu=not args.no_cpu)
        log(f"  {result['legacy_config1_64']}")
        log("48 uneven 2-D rising bubbles stepped together (batched PPE)")
        result["uneven_bubble_batch_48"] = uneven_bubble_batch_bench()
        log(f"  {result['uneven_bubble_batch_48']}")
        result["uneven_bubble_batch_48_mg"] = uneven_bubble_batch_bench(preconditioner="mg")
        log(f"  {result['uneven_bubble_batch_48_mg']}")

    if not args.no_cpu:
        log("CPU baseline (dense-LU CCD slab)")
        result["cpu_baseline"] = cpu_baseline_ccd()

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
