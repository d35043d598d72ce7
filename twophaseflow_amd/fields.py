"""Synthetic, seeded inputs of the CCD and PPE workloads and the
post-projection quantities, on the device.

* band-limited periodic fields (CCD input): |k| <= kmax, amplitude |k|^-2,
  uniform phases, synthesised by inverse FFT on the N^d unique nodes and
  extended with the periodic image planes to the (N+1)^d node layout;
* gas spheres in liquid -> psi -> rho (tpf_fields_bubbles; reference
  simulation/initial_conditions/builder.py:100-172, heaviside.py:167-187);
* PPE right-hand side = CCD divergence of a smooth seeded velocity field;
* FP64 mass integral and velocity correction (tpf_fields_mass,
  tpf_fields_correct_velocity).
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as nat
from .boundary import axis_kinds
from .device import ptr, require_gpu, stream_handle


def band_limited_periodic(N: int, ndim: int, kmax: float = 32.0, seed: int = 2,
                          power: float = 2.0) -> np.ndarray:
    """Real sum over integer wavevectors 1 <= |k| <= kmax of |k|^-power
    cos(2 pi k.x + phase_k) on the (N+1)^ndim node grid of [0,1]^ndim
    (node N is the periodic image of node 0)."""
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(N, d=1.0 / N)
    K = np.meshgrid(*([k1] * ndim), indexing="ij")
    kk = np.sqrt(sum(k * k for k in K))
    mask = (kk >= 1.0) & (kk <= kmax)
    amp = np.zeros(kk.shape)
    amp[mask] = kk[mask] ** (-power)
    phase = rng.uniform(0.0, 2.0 * np.pi, size=kk.shape)
    spec = amp * np.exp(1j * phase)
    f = np.fft.ifftn(spec).real * float(N) ** ndim
    f /= np.max(np.abs(f))
    for a in range(ndim):
        f = np.concatenate([f, np.take(f, [0], axis=a)], axis=a)
    return np.ascontiguousarray(f)


def random_bubbles(n: int, seed: int, ndim: int = 3, rmin: float = 0.03,
                   rmax: float = 0.08):
    """n spheres, r ~ U[rmin, rmax], centres ~ U[r, 1-r]^ndim."""
    rng = np.random.default_rng(seed)
    radii = rng.uniform(rmin, rmax, size=n)
    centres = np.stack([rng.uniform(r, 1.0 - r, size=ndim) for r in radii])
    return centres, radii


def bubbles_psi_rho(N, L, centres, radii, eps: float, rho_l: float = 1.0,
                    rho_g: float = 1e-3):
    """Device psi and rho for gas spheres in liquid (torch float64, grid.shape)."""
    require_gpu()
    ndim = len(N)
    shape = tuple(int(n) + 1 for n in N)
    psi = torch.empty(shape, dtype=torch.float64, device="cuda")
    rho = torch.empty_like(psi)
    c = np.ascontiguousarray(np.asarray(centres, dtype=np.float64).reshape(-1))
    r = np.ascontiguousarray(np.asarray(radii, dtype=np.float64).reshape(-1))
    nat.check(nat.lib().tpf_fields_bubbles(
        ndim, nat.i64arr(shape), nat.darr(L), int(r.size),
        c.ctypes.data_as(C.POINTER(C.c_double)), r.ctypes.data_as(C.POINTER(C.c_double)),
        float(eps), float(rho_l), float(rho_g), ptr(psi), ptr(rho), stream_handle()))
    return psi, rho


def mass(q: torch.Tensor, N, L, bc_type="wall") -> float:
    """sum(q * dV) in FP64 with Grid.cell_volumes weights."""
    ndim = len(N)
    kinds = axis_kinds(bc_type, ndim)
    bcs = [nat.BC_PERIODIC if k == "periodic" else nat.BC_WALL for k in kinds]
    out = C.c_double()
    nat.check(nat.lib().tpf_fields_mass(ndim, nat.i64arr([n + 1 for n in N]), nat.darr(L),
                                        nat.iarr(bcs), ptr(q), C.byref(out), stream_handle()))
    return out.value


def correct_velocity(ccd, vel, p: torch.Tensor, rho: torch.Tensor, dt: float):
    """u <- u - dt/rho * D1 p with wall-normal zeroing, in place on each
    component (coupling/velocity_corrector.py:68-89)."""
    ndim = len(vel)
    kinds = axis_kinds(ccd.bc_type, ndim)
    bcs = [nat.BC_PERIODIC if k == "periodic" else nat.BC_WALL for k in kinds]
    g = torch.empty_like(p)
    for ax, u in enumerate(vel):
        ccd.differentiate_into(p, g, None, ax)
        nat.check(nat.lib().tpf_fields_correct_velocity(
            ndim, nat.i64arr(p.shape), nat.iarr(bcs), ax, float(dt), ptr(rho), ptr(g), ptr(u),
            stream_handle()))
    return vel


def divergence(ccd, vel) -> torch.Tensor:
    """sum_a D1_a u_a (simulation/diagnostics.py:44-49)."""
    div = torch.zeros_like(vel[0])
    g = torch.empty_like(vel[0])
    for ax, u in enumerate(vel):
        ccd.differentiate_into(u, g, None, ax)
        div += g
    return div


def uneven_bubble_batch(seed: int = 5, sizes=(128, 256, 512), per_size: int = 16):
    """Seeded 2-D rising-bubble cases on [0,1]^2 (wall box, bubble centred):
    per_size cases per N x N grid in `sizes`, radius ~ U[0.15, 0.3],
    density ratio rho_l/rho_g log-uniform in [10, 1000], We ~ U[1, 50].
    Returns dicts with N, radius, density_ratio, We and seed order."""
    rng = np.random.default_rng(seed)
    cases = []
    for N in sizes:
        for _ in range(per_size):
            cases.append({"N": int(N), "radius": float(rng.uniform(0.15, 0.3)),
                          "density_ratio": float(10.0 ** rng.uniform(1.0, 3.0)),
                          "We": float(rng.uniform(1.0, 50.0)), "index": len(cases)})
    return cases


def bubble_psi_2d(N: int, radius: float, centre=(0.5, 0.5), epsilon_factor: float = 1.5):
    """Host psi of a gas disc in liquid on the (N+1)^2 nodes of [0,1]^2:
    psi = 1 / (1 + exp(-(|x - c| - radius) / eps)), eps = epsilon_factor / N
    (psi -> 1 in the liquid)."""
    x = np.linspace(0.0, 1.0, N + 1)
    X, Y = np.meshgrid(x, x, indexing="ij")
    d = np.sqrt((X - centre[0]) ** 2 + (Y - centre[1]) ** 2) - radius
    return 1.0 / (1.0 + np.exp(-d / (epsilon_factor / N)))
