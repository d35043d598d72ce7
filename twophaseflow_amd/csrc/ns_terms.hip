// Momentum terms and step plumbing of the legacy two-phase time step
// (simulation/_core.py:126-166) on uniform C-order grids:
//   tpf_ns_properties        rho, mu from psi; phi = H^-1(psi)      heaviside.py:90-118, 167-187
//   tpf_ns_uccd6_convection  skew CCD convection + D2^4 hyperviscosity  uccd6_convection.py:88-128
//   tpf_ns_viscous           ccd_bulk viscous body                     viscous_spatial.py:39-73, 93-169
//   tpf_ns_predictor_rhs     AB2 + gravity + CSF - IPC gradient; Euler point of Picard CN
//   tpf_ns_predictor_final   Picard CN trapezoid                       ab2_predictor.py:110-204, picard_cn.py:44-96
//   tpf_ns_normal            n = g / max(sqrt(max(|g|^2, 1e-28)), 1e-14) closest_point_extender.py:112-131
//   tpf_ns_wall_bc           velocity boundary planes                  boundary_condition.py:36-87
//   tpf_ns_absmax / tpf_ns_ratio_max / tpf_ns_nonfinite   CFL maxima and the retry loop's finite check
// Each kernel is one pass over the nodes (HBM-bound elementwise work); the
// CCD solves come from the plan's kernels.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "reduce.hpp"
#include "tpf_common.hpp"

namespace tpf {
namespace {

constexpr int kMaxComp = 3;

// Node geometry: the ndim axes sit in the last ndim of three slots.
struct NGeom {
  int n[3];
  int periodic[3];
  int64_t stride[3];
  int64_t total;
  int ndim;
};

tpf_status make_ngeom(NGeom& g, int ndim, const int64_t* n_nodes, const int* periodic) {
  if (ndim < 1 || ndim > 3 || !n_nodes) return fail(TPF_ERR_VALUE, "ndim must be 1, 2 or 3");
  g = NGeom{};
  for (int s = 0; s < 3; ++s) g.n[s] = 1;
  for (int a = 0; a < ndim; ++a) {
    if (n_nodes[a] < 2) return fail(TPF_ERR_VALUE, "every axis needs >= 2 nodes");
    g.n[3 - ndim + a] = (int)n_nodes[a];
    g.periodic[3 - ndim + a] = periodic ? periodic[a] != 0 : 0;
  }
  g.stride[2] = 1;
  g.stride[1] = g.n[2];
  g.stride[0] = (int64_t)g.n[1] * g.n[2];
  g.total = (int64_t)g.n[0] * g.n[1] * g.n[2];
  g.ndim = ndim;
  if (g.total >= ((int64_t)1 << 31)) return fail(TPF_ERR_VALUE, "grid too large (2^31 nodes)");
  return TPF_OK;
}

__device__ inline void ncoords(const NGeom& g, int64_t idx, int (&c)[3]) {
  const int64_t plane = g.stride[0];
  c[0] = (int)(idx / plane);
  const int64_t r = idx - (int64_t)c[0] * plane;
  c[1] = (int)(r / g.n[2]);
  c[2] = (int)(r - (int64_t)c[1] * g.n[2]);
}

unsigned grid_blocks(int64_t total) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((total + kBlock - 1) / kBlock, 4096));
}

struct Ptr3 {
  const double* p[kMaxComp];
};
struct MPtr3 {
  double* p[kMaxComp];
};

// ---------------------------------------------------------------- kernels --

__global__ __launch_bounds__(kBlock) void k_properties(int64_t n, const double* __restrict__ psi,
                                                       double rl, double rg, double ml, double mg,
                                                       double eps, double* __restrict__ rho,
                                                       double* __restrict__ mu,
                                                       double* __restrict__ phi) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const double p = psi[i];
    rho[i] = rg + (rl - rg) * p;
    mu[i] = mg + (ml - mg) * p;
    if (phi) {
      const double c = fmin(fmax(p, 1e-6), 1.0 - 1e-6);
      phi[i] = eps * log(c / (1.0 - c));
    }
  }
}

// P[k] = u_k * u_j for every axis k.
__global__ __launch_bounds__(kBlock) void k_products(int64_t n, int nd, Ptr3 u, const double* uj,
                                                     double* __restrict__ P) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const double v = uj[i];
    for (int k = 0; k < nd; ++k) P[k * n + i] = u.p[k][i] * v;
  }
}

// out = sum_k [ -(u_k D1_k u_j + D1_k(u_k u_j))/2 - (sigma h_k^7)(max|u_k| D2_k^4 u_j) ],
// accumulated in the reference's order (acc starts at 0).
__global__ __launch_bounds__(kBlock) void k_uccd6_acc(int64_t n, int nd, Ptr3 u,
                                                      const double* __restrict__ D1,
                                                      const double* __restrict__ Cn,
                                                      const double* __restrict__ D8,
                                                      const double* __restrict__ umax,
                                                      double c0, double c1, double c2,
                                                      double* __restrict__ out) {
  const double coeff[3] = {c0, c1, c2};
  double um[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < nd; ++k) um[k] = umax[k];
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    double acc = 0.0;
    for (int k = 0; k < nd; ++k) {
      acc = acc - 0.5 * (u.p[k][i] * D1[k * n + i] + Cn[k * n + i]);
      acc = acc - coeff[k] * (um[k] * D8[k * n + i]);
    }
    out[i] = acc;
  }
}

// max |a_k| over the nodes: per-block partial maxima, then one finishing block.
__global__ __launch_bounds__(kBlock) void k_absmax_part(int64_t n, Ptr3 a, const double* b0,
                                                        int nk, int mode,
                                                        double* __restrict__ part) {
  __shared__ double sh[kBlock];
  for (int k = 0; k < nk; ++k) {
    // mode 0: max |a|; mode 1: max a / b0
    double m = mode == 0 ? 0.0 : -INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
      m = fmax(m, mode == 0 ? fabs(a.p[k][i]) : a.p[k][i] / b0[i]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
      __syncthreads();
    }
    if (threadIdx.x == 0) part[k * gridDim.x + blockIdx.x] = sh[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_max_fin(const double* __restrict__ part, int nblocks,
                                                    int nk, double* __restrict__ out) {
  __shared__ double sh[kBlock];
  for (int k = 0; k < nk; ++k) {
    double m = -INFINITY;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) m = fmax(m, part[k * nblocks + b]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = sh[0];
    __syncthreads();
  }
}

// out = x / s (mode 0) or out = x + y (mode 1).
__global__ __launch_bounds__(kBlock) void k_elementwise(int64_t n, int mode,
                                                        const double* __restrict__ x,
                                                        const double* __restrict__ y, double s,
                                                        double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    out[i] = mode == 0 ? x[i] / s : x[i] + y[i];
}

// Count of non-finite values over nk arrays.
__global__ __launch_bounds__(kBlock) void k_nonfinite(int64_t n, const double* const* arrs, int nk,
                                                      int* __restrict__ flag) {
  for (int k = 0; k < nk; ++k) {
    const double* a = arrs[k];
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
      if (!isfinite(a[i])) atomicAdd(flag, 1);
  }
}

// Viscous body (ccd_bulk, no interface band): for each component a,
//   out_a = sum_b Dlow_b[ mu (d_b u_a + d_a u_b) ] / (Re rho)
// with G[a][b] = d_b u_a (CCD first derivatives) and Dlow the low-order
// derivative: first-order one-sided at both ends of the axis (whatever its
// boundary), three-point inside with the coefficients (cl, cc, cr).
struct LowCoef {
  double cl[3], cc[3], cr[3], inv_h[3];
};

__global__ __launch_bounds__(kBlock) void k_viscous(NGeom g, Ptr3 G, const double* __restrict__ mu,
                                                    const double* __restrict__ rho, double Re,
                                                    LowCoef lc, MPtr3 out) {
  const int nd = g.ndim;
  const int64_t n = g.total;
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    ncoords(g, idx, c);
    for (int a = 0; a < nd; ++a) {
      double tot = 0.0;
      for (int b = 0; b < nd; ++b) {
        const int s = 3 - nd + b;
        const int64_t st = g.stride[s];
        // S(j) = mu_j (d_b u_a + d_a u_b)_j
        auto S = [&](int64_t j) {
          return mu[j] * (G.p[a][(int64_t)b * n + j] + G.p[b][(int64_t)a * n + j]);
        };
        double d;
        if (c[s] == 0) {
          d = (S(idx + st) - S(idx)) * lc.inv_h[b];
        } else if (c[s] == g.n[s] - 1) {
          d = (S(idx) - S(idx - st)) * lc.inv_h[b];
        } else {
          d = lc.cl[b] * S(idx - st) + lc.cc[b] * S(idx) + lc.cr[b] * S(idx + st);
        }
        tot += d;
      }
      out.p[a][idx] = tot / (Re * rho[idx]);
    }
  }
}

__device__ inline bool on_wall_plane(const NGeom& g, const int (&c)[3], int s) {
  return !g.periodic[s] && (c[s] == 0 || c[s] == g.n[s] - 1);
}

// rhs_c = rho ab2_c + grav_c + st_c - gp_c and the Euler point
// u_pred_c = u_c + dt (rhs_c / rho + vn_c) of the Picard CN advance.
__global__ __launch_bounds__(kBlock) void k_pred_rhs(
    NGeom g, Ptr3 u, Ptr3 conv, Ptr3 conv_prev, int ab2, const double* __restrict__ rho,
    const double* __restrict__ kappa, const double* __restrict__ dpsi,
    const double* __restrict__ dp, double inv_fr2, double We, double dt, Ptr3 vn, MPtr3 rhs,
    MPtr3 upred) {
  const int nd = g.ndim;
  const int64_t n = g.total;
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    ncoords(g, idx, c);
    const double r = rho[idx];
    const double kap = kappa[idx];
    for (int a = 0; a < nd; ++a) {
      const double cv = ab2 ? 1.5 * conv.p[a][idx] - 0.5 * conv_prev.p[a][idx] : conv.p[a][idx];
      const double grav = a == nd - 1 ? -r * inv_fr2 : 0.0;
      const double st = kap * dpsi[(int64_t)a * n + idx] / We;
      const double gp = on_wall_plane(g, c, 3 - nd + a) ? 0.0 : dp[(int64_t)a * n + idx];
      const double e = r * cv + grav + st + (-gp);
      rhs.p[a][idx] = e;
      upred.p[a][idx] = u.p[a][idx] + dt * (e / r + vn.p[a][idx]);
    }
  }
}

// u*_c = u_c + dt (rhs_c / rho + vn_c / 2 + vs_c / 2).
__global__ __launch_bounds__(kBlock) void k_pred_final(NGeom g, Ptr3 u, Ptr3 rhs,
                                                       const double* __restrict__ rho, Ptr3 vn,
                                                       Ptr3 vs, double dt, MPtr3 out) {
  const int nd = g.ndim;
  const int64_t n = g.total;
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * kBlock) {
    const double r = rho[idx];
    for (int a = 0; a < nd; ++a)
      out.p[a][idx] = u.p[a][idx] +
                      dt * (rhs.p[a][idx] / r + 0.5 * vn.p[a][idx] + 0.5 * vs.p[a][idx]);
  }
}

__global__ __launch_bounds__(kBlock) void k_normal(int64_t n, int nd, const double* __restrict__ gr,
                                                   MPtr3 out) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    double s = 0.0;
    for (int a = 0; a < nd; ++a) s += gr[(int64_t)a * n + i] * gr[(int64_t)a * n + i];
    const double nrm = fmax(sqrt(fmax(s, 1e-28)), 1e-14);
    for (int a = 0; a < nd; ++a) out.p[a][i] = gr[(int64_t)a * n + i] / nrm;
  }
}

// Wall axes: every component zero on both planes; periodic axes: the image
// plane copies the root plane (boundary_condition.py:61-87).  One thread
// per node of the boundary planes, slots in order.
__global__ __launch_bounds__(kBlock) void k_wall_bc(NGeom g, int64_t nplanes, int ncomp,
                                                    MPtr3 comp) {
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < nplanes;
       t += (int64_t)gridDim.x * kBlock) {
    int64_t u = t;
    int s = 3 - g.ndim;
    for (; s < 3; ++s) {
      const int64_t pa = g.total / g.n[s];
      if (g.periodic[s]) {
        if (u < pa) break;
        u -= pa;
      } else {
        if (u < 2 * pa) break;
        u -= 2 * pa;
      }
    }
    if (s == 3) continue;
    const int64_t pa = g.total / g.n[s];
    const bool hi = !g.periodic[s] && u >= pa;
    if (hi) u -= pa;
    const int b0 = s == 0 ? 1 : 0, b1 = s == 2 ? 1 : 2;
    int c[3];
    c[s] = (g.periodic[s] || hi) ? g.n[s] - 1 : 0;
    c[b0] = (int)(u / g.n[b1]);
    c[b1] = (int)(u - (int64_t)c[b0] * g.n[b1]);
    const int64_t idx = (int64_t)c[0] * g.stride[0] + (int64_t)c[1] * g.stride[1] + c[2];
    for (int k = 0; k < ncomp; ++k) {
      if (g.periodic[s]) {
        comp.p[k][idx] = comp.p[k][idx - (int64_t)(g.n[s] - 1) * g.stride[s]];
      } else {
        comp.p[k][idx] = 0.0;
      }
    }
  }
}

tpf_status max_reduce(int64_t n, const double* const* a, const double* b, int nk, int mode,
                      double* part, double* out_dev, hipStream_t st) {
  if (nk < 1 || nk > kMaxComp) return fail(TPF_ERR_VALUE, "1..3 arrays");
  Ptr3 pa{};
  for (int k = 0; k < nk; ++k) {
    if (!a[k]) return fail(TPF_ERR_VALUE, "null array");
    pa.p[k] = a[k];
  }
  const unsigned nb = std::min<unsigned>(grid_blocks(n), 1024);
  hipLaunchKernelGGL(k_absmax_part, dim3(nb), dim3(kBlock), 0, st, n, pa, b, nk, mode, part);
  hipLaunchKernelGGL(k_max_fin, dim3(1), dim3(kBlock), 0, st, (const double*)part, (int)nb, nk,
                     out_dev);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

}  // namespace
}  // namespace tpf

using namespace tpf;

extern "C" tpf_status tpf_ns_properties(int64_t n, const double* psi, double rho_l, double rho_g,
                                        double mu_l, double mu_g, double eps, double* rho,
                                        double* mu, double* phi, void* stream) {
  if (n <= 0 || !psi || !rho || !mu) return fail(TPF_ERR_VALUE, "null argument");
  hipLaunchKernelGGL(k_properties, dim3(grid_blocks(n)), dim3(kBlock), 0, (hipStream_t)stream, n,
                     psi, rho_l, rho_g, mu_l, mu_g, eps, rho, mu, phi);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

extern "C" int64_t tpf_ns_uccd6_work(int ndim, const int64_t* n_nodes) {
  int64_t n = 1;
  for (int a = 0; a < ndim; ++a) n *= n_nodes[a];
  return 6 * (int64_t)ndim * n + 1024 * kMaxComp + 8;
}

extern "C" tpf_status tpf_ns_uccd6_convection(tpf_ccd_plan* plan, int ndim,
                                              const int64_t* n_nodes, const double* L,
                                              const double* const* vel, double sigma,
                                              double* const* out, double* work, void* stream) {
  NGeom g;
  TPF_TRY(make_ngeom(g, ndim, n_nodes, nullptr));
  if (!plan || !L || !vel || !out || !work) return fail(TPF_ERR_VALUE, "null argument");
  if (ccd_plan_has_metric(plan)) return fail(TPF_ERR_VALUE, "uniform grids only (plan has a metric)");
  for (int a = 0; a < ndim; ++a)
    if (!vel[a] || !out[a]) return fail(TPF_ERR_VALUE, "null component");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = g.total;
  const int d = ndim;
  double* D1 = work;
  double* D2 = D1 + d * n;
  double* P = D2 + d * n;
  double* Cn = P + d * n;
  double* T = Cn + d * n;
  double* T2 = T + d * n;
  double* part = T2 + d * n;
  double* umax = part + 1024 * kMaxComp;
  const uint32_t all = (1u << d) - 1u;
  // max |u_k| for every axis k (uccd6_convection.py:124), on the device
  TPF_TRY(max_reduce(n, vel, nullptr, d, 0, part, umax, st));
  double coeff[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < d; ++k) coeff[k] = sigma * std::pow(L[k] / (double)(n_nodes[k] - 1), 7);
  Ptr3 u{};
  for (int k = 0; k < d; ++k) u.p[k] = vel[k];
  for (int j = 0; j < d; ++j) {
    // d1, d2 of u_j along every axis k in one grid
    TPF_TRY(tpf_ccd_apply(plan, vel[j], D1, D2, all, stream));
    hipLaunchKernelGGL(k_products, dim3(grid_blocks(n)), dim3(kBlock), 0, st, n, d, u, vel[j], P);
    TPF_HIP(hipGetLastError());
    const double* fs[3];
    double* o1[3];
    double* o2[3];
    // D1_k (u_k u_j) along k
    for (int k = 0; k < d; ++k) {
      fs[k] = P + k * n;
      o1[k] = Cn + k * n;
    }
    TPF_TRY(tpf_ccd_apply_fields(plan, fs, o1, nullptr, all, stream));
    // D2^4 along k: d4 = D2(d2), d6 = D2(d4), d8 = D2(d6)
    double* src = D2;
    double* bufs[2] = {T, T2};
    for (int r = 0; r < 3; ++r) {
      double* dst = bufs[r & 1];
      for (int k = 0; k < d; ++k) {
        fs[k] = src + k * n;
        o2[k] = dst + k * n;
      }
      TPF_TRY(tpf_ccd_apply_fields(plan, fs, nullptr, o2, all, stream));
      src = dst;
    }
    hipLaunchKernelGGL(k_uccd6_acc, dim3(grid_blocks(n)), dim3(kBlock), 0, st, n, d, u,
                       (const double*)D1, (const double*)Cn, (const double*)src,
                       (const double*)umax, coeff[0], coeff[1], coeff[2], out[j]);
    TPF_HIP(hipGetLastError());
  }
  return TPF_OK;
}

extern "C" int64_t tpf_ns_viscous_work(int ndim, const int64_t* n_nodes) {
  int64_t n = 1;
  for (int a = 0; a < ndim; ++a) n *= n_nodes[a];
  return (int64_t)ndim * ndim * n;
}

extern "C" tpf_status tpf_ns_viscous(tpf_ccd_plan* plan, int ndim, const int64_t* n_nodes,
                                     const double* L, const double* const* vel, const double* mu,
                                     const double* rho, double Re, double* const* out,
                                     double* work, void* stream) {
  NGeom g;
  TPF_TRY(make_ngeom(g, ndim, n_nodes, nullptr));
  if (!plan || !L || !vel || !mu || !rho || !out || !work)
    return fail(TPF_ERR_VALUE, "null argument");
  if (ccd_plan_has_metric(plan)) return fail(TPF_ERR_VALUE, "uniform grids only (plan has a metric)");
  const int64_t n = g.total;
  const uint32_t all = (1u << ndim) - 1u;
  Ptr3 G{};
  MPtr3 o{};
  LowCoef lc{};
  for (int a = 0; a < ndim; ++a) {
    if (!vel[a] || !out[a]) return fail(TPF_ERR_VALUE, "null component");
    double* Ga = work + (int64_t)a * ndim * n;
    TPF_TRY(tpf_ccd_apply(plan, vel[a], Ga, nullptr, all, stream));
    G.p[a] = Ga;
    o.p[a] = out[a];
    const double h = L[a] / (double)(n_nodes[a] - 1);
    // viscous_spatial.py:120-128 with h_l = h_r = h
    lc.cl[a] = -(h / (h * (h + h)));
    lc.cc[a] = (h - h) / (h * h);
    lc.cr[a] = h / (h * (h + h));
    lc.inv_h[a] = 1.0 / h;
  }
  hipLaunchKernelGGL(k_viscous, dim3(grid_blocks(n)), dim3(kBlock), 0, (hipStream_t)stream, g, G,
                     mu, rho, Re, lc, o);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

extern "C" tpf_status tpf_ns_predictor_rhs(int ndim, const int64_t* n_nodes, const int* periodic,
                                           const double* const* vel, const double* const* conv,
                                           const double* const* conv_prev, const double* rho,
                                           const double* kappa, const double* dpsi,
                                           const double* dp, double Fr, double We, double dt,
                                           const double* const* visc_n, double* const* rhs,
                                           double* const* u_pred, void* stream) {
  NGeom g;
  TPF_TRY(make_ngeom(g, ndim, n_nodes, periodic));
  if (!vel || !conv || !rho || !kappa || !dpsi || !dp || !visc_n || !rhs || !u_pred)
    return fail(TPF_ERR_VALUE, "null argument");
  Ptr3 u{}, cv{}, cp{}, vn{};
  MPtr3 r{}, up{};
  for (int a = 0; a < ndim; ++a) {
    if (!vel[a] || !conv[a] || !visc_n[a] || !rhs[a] || !u_pred[a] ||
        (conv_prev && !conv_prev[a]))
      return fail(TPF_ERR_VALUE, "null component");
    u.p[a] = vel[a];
    cv.p[a] = conv[a];
    cp.p[a] = conv_prev ? conv_prev[a] : conv[a];
    vn.p[a] = visc_n[a];
    r.p[a] = rhs[a];
    up.p[a] = u_pred[a];
  }
  hipLaunchKernelGGL(k_pred_rhs, dim3(grid_blocks(g.total)), dim3(kBlock), 0, (hipStream_t)stream,
                     g, u, cv, cp, conv_prev ? 1 : 0, rho, kappa, dpsi, dp, 1.0 / (Fr * Fr), We,
                     dt, vn, r, up);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

extern "C" tpf_status tpf_ns_predictor_final(int ndim, const int64_t* n_nodes, const int* periodic,
                                             const double* const* vel, const double* const* rhs,
                                             const double* rho, const double* const* visc_n,
                                             const double* const* visc_pred, double dt,
                                             double* const* u_star, void* stream) {
  NGeom g;
  TPF_TRY(make_ngeom(g, ndim, n_nodes, periodic));
  if (!vel || !rhs || !rho || !visc_n || !visc_pred || !u_star)
    return fail(TPF_ERR_VALUE, "null argument");
  Ptr3 u{}, r{}, vn{}, vs{};
  MPtr3 o{};
  for (int a = 0; a < ndim; ++a) {
    if (!vel[a] || !rhs[a] || !visc_n[a] || !visc_pred[a] || !u_star[a])
      return fail(TPF_ERR_VALUE, "null component");
    u.p[a] = vel[a];
    r.p[a] = rhs[a];
    vn.p[a] = visc_n[a];
    vs.p[a] = visc_pred[a];
    o.p[a] = u_star[a];
  }
  hipLaunchKernelGGL(k_pred_final, dim3(grid_blocks(g.total)), dim3(kBlock), 0,
                     (hipStream_t)stream, g, u, r, rho, vn, vs, dt, o);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

extern "C" tpf_status tpf_ns_normal(int ndim, int64_t n, const double* grad, double* const* nhat,
                                    void* stream) {
  if (ndim < 1 || ndim > 3 || n <= 0 || !grad || !nhat)
    return fail(TPF_ERR_VALUE, "null argument");
  MPtr3 o{};
  for (int a = 0; a < ndim; ++a) {
    if (!nhat[a]) return fail(TPF_ERR_VALUE, "null component");
    o.p[a] = nhat[a];
  }
  hipLaunchKernelGGL(k_normal, dim3(grid_blocks(n)), dim3(kBlock), 0, (hipStream_t)stream, n,
                     ndim, grad, o);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

extern "C" tpf_status tpf_ns_wall_bc(int ndim, const int64_t* n_nodes, const int* periodic,
                                     int ncomp, double* const* comps, void* stream) {
  NGeom g;
  TPF_TRY(make_ngeom(g, ndim, n_nodes, periodic));
  if (ncomp < 1 || ncomp > kMaxComp || !comps) return fail(TPF_ERR_VALUE, "1..3 components");
  MPtr3 c{};
  for (int k = 0; k < ncomp; ++k) {
    if (!comps[k]) return fail(TPF_ERR_VALUE, "null component");
    c.p[k] = comps[k];
  }
  int64_t nplanes = 0;
  for (int s = 3 - ndim; s < 3; ++s) nplanes += (g.periodic[s] ? 1 : 2) * (g.total / g.n[s]);
  hipLaunchKernelGGL(k_wall_bc, dim3(grid_blocks(nplanes)), dim3(kBlock), 0, (hipStream_t)stream,
                     g, nplanes, ncomp, c);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

// max |a_k| (mode 0) into out_host[k]; synchronous.  work >= 1024 * 3 + 3 doubles.
extern "C" tpf_status tpf_ns_absmax(int nk, const double* const* arrays, int64_t n, double* work,
                                    double* out_host, void* stream) {
  if (n <= 0 || !arrays || !work || !out_host) return fail(TPF_ERR_VALUE, "null argument");
  hipStream_t st = (hipStream_t)stream;
  TPF_TRY(max_reduce(n, arrays, nullptr, nk, 0, work, work + 1024 * kMaxComp, st));
  TPF_HIP(hipMemcpyAsync(out_host, work + 1024 * kMaxComp, sizeof(double) * nk,
                         hipMemcpyDeviceToHost, st));
  TPF_HIP(hipStreamSynchronize(st));
  return TPF_OK;
}

// max (a / b) over the nodes into *out_host; synchronous.
extern "C" tpf_status tpf_ns_ratio_max(const double* a, const double* b, int64_t n, double* work,
                                       double* out_host, void* stream) {
  if (n <= 0 || !a || !b || !work || !out_host) return fail(TPF_ERR_VALUE, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const double* arr[1] = {a};
  TPF_TRY(max_reduce(n, arr, b, 1, 1, work, work + 1024 * kMaxComp, st));
  TPF_HIP(hipMemcpyAsync(out_host, work + 1024 * kMaxComp, sizeof(double), hipMemcpyDeviceToHost,
                         st));
  TPF_HIP(hipStreamSynchronize(st));
  return TPF_OK;
}

// Number of non-finite values over nk arrays of n doubles; synchronous.
// work >= 8 + nk pointers' worth of device memory (in doubles: 8 + nk).
extern "C" tpf_status tpf_ns_nonfinite(int nk, const double* const* arrays, int64_t n,
                                       double* work, int64_t* count_host, void* stream) {
  if (nk < 1 || nk > 64 || n <= 0 || !arrays || !work || !count_host)
    return fail(TPF_ERR_VALUE, "bad argument");
  hipStream_t st = (hipStream_t)stream;
  int* flag = reinterpret_cast<int*>(work);
  const double** dptrs = reinterpret_cast<const double**>(work + 8);
  std::vector<const double*> host(arrays, arrays + nk);
  TPF_HIP(hipMemsetAsync(flag, 0, sizeof(int), st));
  TPF_HIP(hipMemcpyAsync(dptrs, host.data(), sizeof(double*) * nk, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_nonfinite, dim3(grid_blocks(n)), dim3(kBlock), 0, st, n,
                     (const double* const*)dptrs, nk, flag);
  TPF_HIP(hipGetLastError());
  int h = 0;
  TPF_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st));
  TPF_HIP(hipStreamSynchronize(st));
  *count_host = h;
  return TPF_OK;
}

extern "C" tpf_status tpf_ns_div_scalar(int64_t n, const double* x, double s, double* out,
                                        void* stream) {
  if (n <= 0 || !x || !out) return fail(TPF_ERR_VALUE, "null argument");
  hipLaunchKernelGGL(k_elementwise, dim3(grid_blocks(n)), dim3(kBlock), 0, (hipStream_t)stream, n,
                     0, x, (const double*)nullptr, s, out);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

extern "C" tpf_status tpf_ns_add(int64_t n, const double* x, const double* y, double* out,
                                 void* stream) {
  if (n <= 0 || !x || !y || !out) return fail(TPF_ERR_VALUE, "null argument");
  hipLaunchKernelGGL(k_elementwise, dim3(grid_blocks(n)), dim3(kBlock), 0, (hipStream_t)stream, n,
                     1, x, y, 0.0, out);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}
