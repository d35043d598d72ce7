"""Node-centred grid with the reference ``Grid`` surface the CCD and PPE
paths use (``core/grid.py:36-81``, ``cell_volumes`` :1123-1170), and the
grid metric of non-uniform axes (``core/metrics.py:22-88``)."""

from __future__ import annotations

import numpy as np

from .boundary import axis_kinds


def require_uniform(grid, owner: str) -> None:
    """Components other than ``CCDSolver`` are built for uniform grids: a
    grid with ``alpha_grid > 1`` raises NotImplementedError here."""
    if grid is not None and not getattr(grid, "uniform", True):
        raise NotImplementedError(f"{owner}: non-uniform grids (alpha_grid > 1) are supported "
                                  "by CCDSolver only")


def node_spacing(coords) -> np.ndarray:
    """Per-node spacing of a coordinate vector: the end cells' widths at the
    ends, the mean of the two adjacent cells inside (core/grid.py:1078-1084)."""
    c = np.asarray(coords, dtype=float)
    w = np.diff(c)
    out = np.empty(c.shape[0])
    out[0], out[-1] = w[0], w[-1]
    out[1:-1] = 0.5 * (w[:-1] + w[1:])
    return out


class Grid:
    """N cells / N+1 nodes per axis; coords start as linspace(0, L, N+1).

    ``alpha_grid > 1`` marks the grid non-uniform as in the reference
    (``uniform`` is ``alpha_grid <= 1``): CCD then differentiates in the
    computational coordinate xi in [0, 1] and maps the result with the metric
    ``J = dxi/dx``, ``dJ_dxi``.  The reference's interface-fitted coordinate
    generator (``update_from_levelset``) is not part of this package; set the
    coordinates of an axis with ``set_coordinates`` (or assign ``coords[a]``
    and ``h[a]`` and call ``_build_metrics(ccd)``, as the reference's tests
    do)."""

    def __init__(self, grid_config, backend=None):
        self.ndim = int(grid_config.ndim)
        self.N = tuple(int(n) for n in grid_config.N)
        self.L = tuple(float(x) for x in grid_config.L)
        self.alpha_grid = float(getattr(grid_config, "alpha_grid", 1.0))
        self.backend = backend
        self.coords = [np.linspace(0.0, self.L[a], self.N[a] + 1) for a in range(self.ndim)]
        self.h = [np.full(self.N[a] + 1, self.L[a] / self.N[a]) for a in range(self.ndim)]
        self.shape = tuple(n + 1 for n in self.N)
        self.bc_type = getattr(grid_config, "bc_type", "wall")
        self._build_metrics()

    @property
    def uniform(self) -> bool:
        """True when alpha_grid <= 1 (core/grid.py:1104-1107)."""
        return self.alpha_grid <= 1.0

    def _build_metrics(self, ccd=None) -> None:
        """J = dxi/dx and dJ/dxi per axis (core/metrics.py:22-88).

        Non-uniform grid with a CCD solver: x(xi) is differentiated by the
        CCD wall solve in xi (``ccd.differentiate_raw``), J = 1 / x_xi and
        dJ/dxi = -x_xixi / x_xi^2.  Otherwise the coordinates must be
        uniform: J = (1/N) / h and dJ/dxi its second-order differences
        (central inside, one-sided at the ends); non-uniform coordinates
        without a CCD solver raise ValueError."""
        J, dJ = [], []
        for a in range(self.ndim):
            if ccd is not None and not self.uniform:
                d1, d2 = ccd.differentiate_raw(self.coords[a], axis=a)
                d1 = np.asarray(d1.cpu() if hasattr(d1, "cpu") else d1, dtype=float)
                d2 = np.asarray(d2.cpu() if hasattr(d2, "cpu") else d2, dtype=float)
                Ja = 1.0 / d1
                dJa = -d2 / (d1 ** 2)
            else:
                w = np.diff(np.asarray(self.coords[a], dtype=float))
                even = w.size == 0 or np.allclose(w, w[0], rtol=1.0e-12, atol=1.0e-14)
                if not self.uniform and not even:
                    raise ValueError("non-uniform grid metrics require a CCDSolver; "
                                     "low-order metric substitution is not permitted")
                dxi = 1.0 / self.N[a]
                Ja = dxi / np.asarray(self.h[a], dtype=float)
                dJa = np.zeros_like(Ja)
                if Ja.size > 1:
                    dJa[1:-1] = (Ja[2:] - Ja[:-2]) / (2.0 * dxi)
                    dJa[0] = (Ja[1] - Ja[0]) / dxi
                    dJa[-1] = (Ja[-1] - Ja[-2]) / dxi
            J.append(Ja)
            dJ.append(dJa)
        self.J, self.dJ_dxi = J, dJ

    def set_coordinates(self, axis: int, coords, ccd=None) -> None:
        """Node coordinates of one axis (increasing, N[axis]+1 values from 0
        to L[axis]); node spacing and the metric of every axis are rebuilt
        (with ``ccd``, a CCDSolver on this grid, for non-uniform ones)."""
        c = np.asarray(coords, dtype=float).reshape(-1)
        n = self.N[axis] + 1
        if c.shape[0] != n:
            raise ValueError(f"axis {axis} needs {n} coordinates")
        if not np.all(np.diff(c) > 0.0):
            raise ValueError("coordinates must increase strictly")
        if abs(c[0]) > 1e-12 * self.L[axis] or abs(c[-1] - self.L[axis]) > 1e-12 * self.L[axis]:
            raise ValueError(f"coordinates of axis {axis} must run from 0 to L = {self.L[axis]}")
        self.coords[axis] = c
        self.h[axis] = node_spacing(c)
        self._build_metrics(ccd)

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
