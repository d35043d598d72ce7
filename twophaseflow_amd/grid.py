"""Uniform node-centred grid with the reference ``Grid`` surface the CCD and
PPE paths use (``core/grid.py:36-81``, ``cell_volumes`` :1123-1170)."""

from __future__ import annotations

import numpy as np

from .boundary import axis_kinds


class Grid:
    """N cells / N+1 nodes per axis; coords = linspace(0, L, N+1)."""

    def __init__(self, grid_config, backend=None):
        self.ndim = int(grid_config.ndim)
        self.N = tuple(int(n) for n in grid_config.N)
        self.L = tuple(float(x) for x in grid_config.L)
        if getattr(grid_config, "alpha_grid", 1.0) != 1.0:
            raise ValueError("twophaseflow_amd supports uniform grids only (alpha_grid=1)")
        self.uniform = True
        self.backend = backend
        self.coords = [np.linspace(0.0, self.L[a], self.N[a] + 1) for a in range(self.ndim)]
        self.h = [np.full(self.N[a] + 1, self.L[a] / self.N[a]) for a in range(self.ndim)]
        self.shape = tuple(n + 1 for n in self.N)
        self.bc_type = getattr(grid_config, "bc_type", "wall")

    def set_boundary_type(self, bc_type) -> None:
        self.bc_type = bc_type

    def meshgrid(self):
        return np.meshgrid(*self.coords, indexing="ij")

    def cell_volumes(self, bc_type=None):
        """Nodal control volumes: full h on wall axes (grid.py:66-69), the
        periodic image node carries 0 (grid.py:1160-1168)."""
        kinds = axis_kinds(self.bc_type if bc_type is None else bc_type, self.ndim)
        vol = None
        for a in range(self.ndim):
            w = np.asarray(self.h[a], dtype=float).copy()
            if kinds[a] == "periodic":
                cw = np.diff(self.coords[a])
                w[:-1] = 0.5 * (np.roll(cw, 1) + cw)
                w[-1] = 0.0
            vol = w if vol is None else np.multiply.outer(vol, w)
        return vol

    def __repr__(self) -> str:
        return f"Grid(ndim={self.ndim}, N={self.N}, L={self.L})"
