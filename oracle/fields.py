"""Oracle: PPE inputs and post-projection outputs, FP64 NumPy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* sphere SDF union + CLS psi  ``simulation/initial_conditions/builder.py:100-172``,
                              ``simulation/initial_conditions/shape_primitives.py:21-53``
* density from psi            ``levelset/heaviside.py:167-187``
* cell volumes                ``core/grid.py:1123-1170`` (wall axes: full h at
                              boundary nodes; periodic: image node weight 0)
* mass correction             ``levelset/heaviside.py:118-149``
* velocity correction         ``coupling/velocity_corrector.py:39-89``
* CCD divergence              ``simulation/diagnostics.py:44-49``
"""

from __future__ import annotations

import numpy as np


def node_coords(N, L):
    return [np.linspace(0.0, L[a], N[a] + 1) for a in range(len(N))]


def gas_bubbles_psi(N, L, centres, radii, eps):
    """psi for gas spheres in a liquid background.

    phi_gas = min_k (|x-c_k| - r_k); phi = max(-inf, -phi_gas) = -phi_gas;
    psi = 1/(1+exp(clip(phi/eps, -500, 500))) -> psi ~ 1 in liquid.
    """
    X = np.meshgrid(*node_coords(N, L), indexing="ij")
    phi_gas = np.full(X[0].shape, np.inf)
    for c, r in zip(centres, radii):
        r2 = sum((X[a] - c[a]) ** 2 for a in range(len(N)))
        phi_gas = np.minimum(phi_gas, np.sqrt(r2) - r)
    phi = -phi_gas
    return 1.0 / (1.0 + np.exp(np.clip(phi / eps, -500.0, 500.0)))


def density(psi, rho_l, rho_g):
    """rho = rho_g + (rho_l - rho_g) psi (heaviside.py:185)."""
    return rho_g + (rho_l - rho_g) * psi


def cell_volumes(N, L, kinds):
    vol = None
    for a in range(len(N)):
        x = np.linspace(0.0, L[a], N[a] + 1)
        if kinds[a] == "periodic":
            cw = np.diff(x)
            w = np.empty_like(x)
            w[:-1] = 0.5 * (np.roll(cw, 1) + cw)
            w[-1] = 0.0
        else:
            w = np.full(N[a] + 1, L[a] / N[a])
        vol = w if vol is None else np.multiply.outer(vol, w)
    return vol


def mass(psi, dV):
    return float(np.sum(psi * dV))


def mass_correction(q, dV, M_target):
    """heaviside.py:118-149 (division order preserved)."""
    M_new = np.sum(q * dV)
    w = 4.0 * q * (1.0 - q)
    W = np.sum(w * dV)
    W_safe = W if W > 1e-12 else 1.0
    gate = 1.0 if W > 1e-12 else 0.0
    ratio = (M_target - M_new) / W_safe
    return np.clip(q + (gate * ratio) * w, 0.0, 1.0)


def correct_velocity(ccd, vel_star, p, rho, dt):
    """u - dt/rho * D1 p with wall-normal zeroing (velocity_corrector.py:39-89)."""
    out = []
    for ax, u in enumerate(vel_star):
        g, _ = ccd.differentiate(p, ax)
        ccd.enforce_wall_neumann(g, ax)
        out.append(u - (dt / rho) * g)
    return out


def divergence(ccd, vel):
    """sum_ax D1_ax u_ax (diagnostics.py:44-49)."""
    div = np.zeros(np.asarray(vel[0]).shape)
    for ax, u in enumerate(vel):
        d1, _ = ccd.differentiate(u, ax)
        div += d1
    return div
