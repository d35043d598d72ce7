"""Boundary-type helpers with the reference semantics (``core/boundary.py``)."""

from __future__ import annotations

import numpy as np


def axis_kinds(bc_type, ndim: int):
    """Expand a compact boundary name into 'wall'/'periodic' per axis
    (core/boundary.py:45-66)."""
    if isinstance(bc_type, (tuple, list)):
        axes = tuple(str(v).strip().lower() for v in bc_type)
        if len(axes) != ndim:
            raise ValueError(f"len(boundary axes)={len(axes)} != ndim={ndim}")
        if any(a not in ("wall", "periodic") for a in axes):
            raise ValueError(f"boundary axes must be wall|periodic, got {axes!r}")
        return axes
    v = str(getattr(bc_type, "value", bc_type)).strip().lower()
    if v == "wall":
        return ("wall",) * ndim
    if v == "periodic":
        return ("periodic",) * ndim
    if v in ("periodic_wall", "x_periodic_y_wall") and ndim == 2:
        return ("periodic", "wall")
    if v in ("wall_periodic", "x_wall_y_periodic") and ndim == 2:
        return ("wall", "periodic")
    raise ValueError(f"unsupported boundary type {bc_type!r} for ndim={ndim}")


def is_periodic_axis(bc_type, axis: int, ndim: int) -> bool:
    return axis_kinds(bc_type, ndim)[axis] == "periodic"


def pin_dof(N, bc_type="wall") -> int:
    """Gauge pin: centre node for wall, node 0 for all-periodic
    (core/boundary.py:181-191)."""
    kinds = axis_kinds(bc_type, len(N))
    if all(k == "periodic" for k in kinds):
        return 0
    shape = tuple(n + 1 for n in N)
    return int(np.ravel_multi_index(tuple(n // 2 for n in N), shape))


class BoundarySpec:
    """Minimal ``BoundarySpec`` (core/boundary.py:143-204)."""

    def __init__(self, bc_type, shape, N):
        self.bc_type = bc_type
        self.shape = tuple(shape)
        self.N = tuple(N)

    @property
    def pin_dof(self) -> int:
        return pin_dof(self.N, self.bc_type)
