"""Launch-shape autotuner for the CCD line kernels.

No counterpart in the reference.  For each axis of a CCDSolver it times the
admissible launch shapes on the solver's own grid with HIP events and keeps
the fastest: the lane-per-line kernel with rows caps 8/16/32/64 (<= 16
partitions) and the wavefront-per-line kernel with rows caps 8/16 (<= 64
partitions), each partition holding >= 2 (periodic) or 4 (wall) rows.  Results are cached per (nodes, bc, lines) key in a
JSON file (``TPF_AUTOTUNE_CACHE`` or ``~/.cache/twophaseflow_amd/ccd_tune.json``)
so later solvers of the same shape reuse them without timing.
"""

from __future__ import annotations

import json
import os

import torch

from . import _native as nat

CANDIDATES = (8, 16, 32, 64)
MAX_PARTS = 16


def _cache_path() -> str:
    return os.environ.get("TPF_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "twophaseflow_amd",
                                       "ccd_tune.json"))


def _load() -> dict:
    try:
        with open(_cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _save(table: dict) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as fh:
            json.dump(table, fh, indent=1, sort_keys=True)
    except OSError:
        pass  # a read-only home only loses the cache


def admissible_shapes(n_nodes: int, periodic: bool, strided: bool = True):
    """(mode, rows_cap) pairs the plan accepts for one axis ("explicit" only
    on strided axes)."""
    n = n_nodes - 1 if periodic else n_nodes - 2
    min_rows = 2 if periodic else 4
    out = []
    shapes = [("lane", CANDIDATES, MAX_PARTS), ("wave", (8, 16), 64)]
    if strided:
        shapes.append(("explicit", (8, 16), MAX_PARTS))
    for mode, caps, max_parts in shapes:
        for m in caps:
            parts = -(-n // m)
            if parts <= max_parts and (parts == 1 or n // parts >= min_rows):
                out.append((mode, m))
    return out


def tune_ccd(solver, reps: int = 10, use_cache: bool = True) -> dict:
    """Times every admissible launch shape on each axis of `solver`, applies
    the fastest, and returns {axis: {"mode", "rows_cap", "ms": {shape: time}}}."""
    shape = tuple(int(n) + 1 for n in solver.grid.N)
    f = torch.randn(shape, dtype=torch.float64, device="cuda")
    d1 = torch.empty_like(f)
    d2 = torch.empty_like(f)
    table = _load() if use_cache else {}
    result = {}
    lines = 1
    for s in shape:
        lines *= s
    for axis, kind in enumerate(solver._kinds):
        periodic = kind == "periodic"
        key = f"{shape[axis]}:{kind}:{lines // shape[axis]}:{axis == len(shape) - 1}"
        if use_cache and key in table:
            solver.set_launch(axis, table[key]["mode"], int(table[key]["rows_cap"]))
            result[axis] = dict(table[key], cached=True)
            continue
        times = {}
        for mode, m in admissible_shapes(shape[axis], periodic, strided=axis < len(shape) - 1):
            solver.set_launch(axis, mode, m)
            solver.differentiate_into(f, d1, d2, axis)  # builds the tables
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                solver.differentiate_into(f, d1, d2, axis)
            e1.record()
            torch.cuda.synchronize()
            times[(mode, m)] = e0.elapsed_time(e1) / reps
        best = min(times, key=times.get)
        solver.set_launch(axis, best[0], best[1])
        entry = {"mode": best[0], "rows_cap": best[1],
                 "ms": {f"{k[0]}{k[1]}": round(v, 5) for k, v in times.items()}}
        table[key] = entry
        result[axis] = dict(entry, cached=False)
    if use_cache:
        _save(table)
    return result
