"""Oracle: FP64 NumPy restatement of the multigrid hierarchy (libtpf tpf_mg_*).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  No counterpart in
the reference (the multigrid backbone of the learned preconditioner); the
fine operator is the reference's CG system (ppe/fvm_matrixfree.py:163-221).

Arrays use the device layout: 3 slots (n0, n1, n2), leading slots of size 1
for 2-D and 1-D grids.  Pins (tests/test_oracle_mg.py):

* ``fine_conductances`` + ``apply`` equal ``OraclePPE.system_matvec`` on
  vectors with a zero pin entry (the reference CG system);
* ``coarse_conductances`` equals the 2h rediscretisation for constant
  density and the exact Schur complement (elimination of the odd nodes) on
  1-D grids with variable density;
* ``restrict`` is the exact adjoint of ``prolong``;
* ``cycle`` is symmetric and contracts the A-norm error.
"""

from __future__ import annotations

import numpy as np


def level_shapes(shape, pin, coarse_max=1024, max_levels=12):
    """Node counts and pin index per level (coarsening rule of tpf_mg_setup)."""
    shapes, pins = [tuple(shape)], [int(pin)]
    while len(shapes) < max_levels:
        f = shapes[-1]
        if int(np.prod(f)) <= coarse_max:
            break
        pc = np.unravel_index(pins[-1], f)
        ok, c = True, []
        for a in range(3):
            if f[a] == 1:
                c.append(1)
                continue
            if f[a] < 5 or f[a] % 2 == 0 or pc[a] % 2:
                ok = False
            c.append((f[a] - 1) // 2 + 1)
        if not ok:
            break
        shapes.append(tuple(c))
        pins.append(int(np.ravel_multi_index(tuple(p // 2 for p in pc), tuple(c))))
    return shapes, pins


def _dv(n, h):
    if n == 1:
        return np.ones(1)
    d = np.full(n, h)
    d[0] = d[-1] = 0.5 * h
    return d


def fine_conductances(rho, h):
    """T[a][i] = prod_{b != a} dv_b * 2/(rho_i + rho_{i+e_a}) / h_a for the
    face to the +1 neighbour along slot a (0 on the last plane)."""
    rho = np.asarray(rho, dtype=np.float64)
    n = rho.shape
    dvs = [_dv(n[b], h[b]).reshape([-1 if c == b else 1 for c in range(3)]) for b in range(3)]
    T = []
    for a in range(3):
        t = np.zeros(n)
        if n[a] > 1:
            area = np.ones(n)
            for b in range(3):
                if b != a and n[b] > 1:
                    area = area * dvs[b]
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[a] = slice(0, n[a] - 1)
            hi[a] = slice(1, n[a])
            t[tuple(lo)] = (area[tuple(lo)] * 2.0 / (rho[tuple(lo)] + rho[tuple(hi)]) / h[a])
        T.append(t)
    return T


def coarse_conductances(T, cshape):
    """Series combination of the two fine faces spanned by a coarse face,
    summed over the transverse fine faces with weights 1/2, 1, 1/2 per axis."""
    fshape = T[0].shape
    out = []
    for a in range(3):
        tc = np.zeros(cshape)
        if cshape[a] > 1:
            t = T[a]
            s_lo = [slice(None)] * 3
            s_hi = [slice(None)] * 3
            s_lo[a] = slice(0, fshape[a] - 2, 2)
            s_hi[a] = slice(1, fshape[a] - 1, 2)
            t1, t2 = t[tuple(s_lo)], t[tuple(s_hi)]
            with np.errstate(divide="ignore", invalid="ignore"):
                ser = np.where((t1 > 0) & (t2 > 0), t1 * t2 / (t1 + t2), 0.0)
            # transverse full weighting onto even fine nodes
            for b in range(3):
                if b == a or cshape[b] == 1:
                    continue
                ser = _weight_axis(ser, b)
            sl = [slice(None)] * 3
            sl[a] = slice(0, cshape[a] - 1)
            tc[tuple(sl)] = ser
        out.append(tc)
    return out


def _weight_axis(x, b):
    """y_I = x_{2I} + (x_{2I-1} + x_{2I+1}) / 2 along axis b (missing
    neighbours dropped)."""
    n = x.shape[b]
    m = (n - 1) // 2 + 1
    xm = np.moveaxis(x, b, 0)
    y = xm[0::2].copy()
    y[1:] += 0.5 * xm[1::2][: m - 1]
    y[:-1] += 0.5 * xm[1::2][: m - 1]
    return np.moveaxis(y, 0, b)


def diag_inv(T, pin):
    d = np.zeros(T[0].shape)
    for a in range(3):
        t = T[a]
        d += t
        sh = np.zeros_like(t)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a] = slice(1, None)
        hi[a] = slice(0, -1)
        sh[tuple(lo)] = t[tuple(hi)]
        d += sh
    with np.errstate(divide="ignore"):
        di = np.where(d > 0, 1.0 / np.where(d > 0, d, 1.0), 0.0)
    di.ravel()[pin] = 0.0
    return di


def _nbr_sum(T, x):
    s = np.zeros_like(x)
    for a in range(3):
        t = T[a]
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a] = slice(0, -1)
        hi[a] = slice(1, None)
        s[tuple(lo)] += t[tuple(lo)] * x[tuple(hi)]
        s[tuple(hi)] += t[tuple(lo)] * x[tuple(lo)]
    return s


def _diag(T):
    d = np.zeros(T[0].shape)
    for a in range(3):
        t = T[a]
        d += t
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a] = slice(1, None)
        hi[a] = slice(0, -1)
        d[tuple(lo)] += t[tuple(hi)]
    return d


def apply(T, x, pin):
    """y = A x, y[pin] = 0."""
    y = _diag(T) * x - _nbr_sum(T, x)
    y.ravel()[pin] = 0.0
    return y


def rbgs_half(T, dinv, b, x, color):
    """One red-black half sweep in place on nodes with (i0+i1+i2) % 2 == color."""
    idx = np.indices(x.shape).sum(axis=0) % 2 == color
    upd = (b + _nbr_sum(T, x)) * dinv
    x[idx] = upd[idx]
    return x


def smooth(T, dinv, b, x, sweeps, reverse):
    for _ in range(sweeps):
        for h in range(2):
            rbgs_half(T, dinv, b, x, 1 - h if reverse else h)
    return x


def restrict(r, cshape, cpin):
    """Full weighting (the transpose of trilinear interpolation); b_c[pin] = 0."""
    y = np.asarray(r, dtype=np.float64)
    for b in range(3):
        if cshape[b] > 1:
            y = _weight_axis(y, b)
    y = y.copy()
    y.ravel()[cpin] = 0.0
    return y


def prolong(xc, fshape, fpin):
    """Trilinear interpolation; the fine pin entry is 0."""
    y = np.asarray(xc, dtype=np.float64)
    for b in range(3):
        if fshape[b] > 1:
            ym = np.moveaxis(y, b, 0)
            out = np.empty((fshape[b],) + ym.shape[1:])
            out[0::2] = ym
            out[1::2] = 0.5 * (ym[:-1] + ym[1:])
            y = np.moveaxis(out, 0, b)
    y = y.copy()
    y.ravel()[fpin] = 0.0
    return y


class OracleMG:
    """Hierarchy and cycles of tpf_mg_* in FP64."""

    def __init__(self, rho3, h3, pin, max_levels=12):
        rho3 = np.asarray(rho3, dtype=np.float64)
        self.shapes, self.pins = level_shapes(rho3.shape, pin, max_levels=max_levels)
        self.T = [fine_conductances(rho3, h3)]
        for c in self.shapes[1:]:
            self.T.append(coarse_conductances(self.T[-1], c))
        self.dinv = [diag_inv(T, p) for T, p in zip(self.T, self.pins)]

    def apply(self, level, x):
        return apply(self.T[level], x, self.pins[level])

    def coarse_solve(self, level, b, sweeps):
        x = np.zeros(self.shapes[level])
        smooth(self.T[level], self.dinv[level], b, x, sweeps, False)
        smooth(self.T[level], self.dinv[level], b, x, sweeps, True)
        return x

    def cycle(self, b, nu=2, coarse_sweeps=20, gamma=1, stop_level=-1):
        last = len(self.shapes) - 1 if stop_level < 0 else stop_level
        return self._cycle(0, last, np.asarray(b, dtype=np.float64), None, nu, coarse_sweeps,
                           gamma)

    def _cycle(self, l, last, b, x, nu, cs, gamma):
        if l == last:
            return self.coarse_solve(l, b, cs)
        T, di = self.T[l], self.dinv[l]
        x = np.zeros(self.shapes[l]) if x is None else x
        smooth(T, di, b, x, nu, False)
        r = b - apply(T, x, self.pins[l])
        r.ravel()[self.pins[l]] = 0.0
        bc = restrict(r, self.shapes[l + 1], self.pins[l + 1])
        xc = None
        for k in range(gamma):
            if k > 0 and l + 1 == last:
                break
            xc = self._cycle(l + 1, last, bc, xc, nu, cs, gamma)
        x += prolong(xc, self.shapes[l], self.pins[l])
        smooth(T, di, b, x, nu, True)
        return x

    def prolong_chain(self, level, xc):
        y = xc
        for l in range(level - 1, -1, -1):
            y = prolong(y, self.shapes[l], self.pins[l])
        return y

    def restrict_chain(self, r, level):
        y = r
        for l in range(level):
            y = restrict(y, self.shapes[l + 1], self.pins[l + 1])
        return y
