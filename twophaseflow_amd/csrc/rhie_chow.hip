// Rhie-Chow face-velocity divergence, the PPE right-hand side of the
// projection step (spatial/rhie_chow.py:40-284): per axis the face velocity
//   u_f = (u_L + u_R)/2 - dt * 2/(rho_L + rho_R) * [(p_R - p_L)/h - (dp_L + dp_R)/2]
// with dp the CCD first derivative of p (one-sided differences on wall nodes,
// :140-150), optionally minus the balanced-force surface-tension bracket
// [kappa_f (psi_R - psi_L)/(h We) - (fs_L + fs_R)/2], fs = kappa dpsi / We
// (:152-204); face 0 is the periodic wrap (nodes N and 0) or a zero wall face
// (:206-234); the nodal divergence is (F_{k+1} - F_k)/h, node N taking
// (F_0 - F_N)/h on periodic axes and -F_N/h at a wall (:238-284).
// 1-D / 2-D: one streaming kernel, a thread per node evaluating its two faces
// per axis from p, rho, u and the stacked CCD derivatives (neighbours via
// L1/L2) and summing the axes in order; 3-D: the slab-marching kernel below.
// The face arrays never reach HBM.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "tpf_common.hpp"

namespace tpf {
namespace {

constexpr int kBlock = 256;

struct RcGeom {
  int n[3];
  int periodic[3];
  int k[3];  // axis index of the slot (stacked derivative arrays)
  int64_t total;
  double h[3];
  int bf;         // balanced-force term on
  double we;
  double dt;
};

struct RcPtrs {
  const double* u[3];
  const double* p;
  const double* rho;
  const double* dp;     // [ndim][total]
  const double* kappa;  // balanced force only
  const double* psi;
  const double* dpsi;   // [ndim][total]
};

__device__ __forceinline__ int64_t stride_of(const RcGeom& g, int a) {
  return a == 0 ? (int64_t)g.n[1] * g.n[2] : (a == 1 ? (int64_t)g.n[2] : 1);
}

// Cell pressure gradient of node m along slot a (line base b): CCD value,
// one-sided on wall nodes (rhie_chow.py:140-150).
__device__ __forceinline__ double dp_cell(const RcGeom& g, const RcPtrs& P, int a, int64_t b,
                                          int64_t st, int m, const double* dpa) {
  const int N = g.n[a] - 1;
  if (!g.periodic[a]) {
    if (m == 0) return (P.p[b + st] - P.p[b]) / g.h[a];
    if (m == N) return (P.p[b + (int64_t)N * st] - P.p[b + (int64_t)(N - 1) * st]) / g.h[a];
  }
  return dpa[b + (int64_t)m * st];
}

// Cell surface-tension force kappa dpsi / We, one-sided on wall nodes
// (rhie_chow.py:152-170).
__device__ __forceinline__ double fs_cell(const RcGeom& g, const RcPtrs& P, int a, int64_t b,
                                          int64_t st, int m, const double* dpsia) {
  const int N = g.n[a] - 1;
  const int64_t o = b + (int64_t)m * st;
  if (!g.periodic[a]) {
    if (m == 0) return P.kappa[o] * (P.psi[b + st] - P.psi[b]) / (g.h[a] * g.we);
    if (m == N)
      return P.kappa[o] * (P.psi[o] - P.psi[o - st]) / (g.h[a] * g.we);
  }
  return P.kappa[o] * dpsia[o] / g.we;
}

// Face flux between nodes mL and mR of the line with base b along slot a.
__device__ __forceinline__ double face_flux(const RcGeom& g, const RcPtrs& P, int a, int64_t b,
                                            int64_t st, int mL, int mR) {
  const double* dpa = P.dp + (int64_t)g.k[a] * g.total;
  const int64_t oL = b + (int64_t)mL * st, oR = b + (int64_t)mR * st;
  const double h = g.h[a];
  const double u_bar = 0.5 * (P.u[a][oL] + P.u[a][oR]);
  const double dp_face = (P.p[oR] - P.p[oL]) / h;
  const double dp_bar = 0.5 * (dp_cell(g, P, a, b, st, mL, dpa) + dp_cell(g, P, a, b, st, mR, dpa));
  const double inv_rho = 2.0 / (P.rho[oL] + P.rho[oR]);
  double bracket = dp_face - dp_bar;
  if (g.bf) {
    const double* dpsia = P.dpsi + (int64_t)g.k[a] * g.total;
    const double kappa_face = 0.5 * (P.kappa[oL] + P.kappa[oR]);
    const double f_face = kappa_face * (P.psi[oR] - P.psi[oL]) / (h * g.we);
    const double f_bar =
        0.5 * (fs_cell(g, P, a, b, st, mL, dpsia) + fs_cell(g, P, a, b, st, mR, dpsia));
    bracket = (dp_face - dp_bar) - (f_face - f_bar);
  }
  return u_bar - g.dt * inv_rho * bracket;
}

// Flux of face k (0..N) along slot a: face 0 is the periodic wrap or 0.
__device__ __forceinline__ double flux_k(const RcGeom& g, const RcPtrs& P, int a, int64_t b,
                                         int64_t st, int k) {
  const int N = g.n[a] - 1;
  if (k == 0) return g.periodic[a] ? face_flux(g, P, a, b, st, N, 0) : 0.0;
  return face_flux(g, P, a, b, st, k - 1, k);
}

__global__ __launch_bounds__(kBlock) void k_rc_div(RcGeom g, RcPtrs P, double* __restrict__ out) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    const uint32_t plane = (uint32_t)g.n[1] * (uint32_t)g.n[2];
    int c[3];
    c[0] = (int)((uint32_t)idx / plane);
    const uint32_t rem = (uint32_t)idx - (uint32_t)c[0] * plane;
    c[1] = (int)(rem / (uint32_t)g.n[2]);
    c[2] = (int)(rem - (uint32_t)c[1] * (uint32_t)g.n[2]);
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (g.n[a] == 1) continue;
      const int N = g.n[a] - 1;
      const int i = c[a];
      const int64_t st = stride_of(g, a);
      const int64_t b = idx - (int64_t)i * st;
      double d;
      if (i < N) d = (flux_k(g, P, a, b, st, i + 1) - flux_k(g, P, a, b, st, i)) / g.h[a];
      else if (g.periodic[a]) d = (flux_k(g, P, a, b, st, 0) - flux_k(g, P, a, b, st, N)) / g.h[a];
      else d = -flux_k(g, P, a, b, st, N) / g.h[a];
      acc += d;
    }
    out[idx] = acc;
  }
}

// 3-D: the workgroups march kSlab planes along slot 0 (one thread per node of
// a plane).  Each thread carries the slot-0 cell state of its current plane
// (u_0, p, rho, the cell pressure gradient and, with the balanced force,
// kappa, psi and the cell force) and the flux of the face below it, so every
// slot-0 face is evaluated once and every plane's values are read once; the
// in-plane slots use the per-node face evaluation above.  The arithmetic is
// the same as k_rc_div's, operation for operation.
constexpr int kSlab = 16;

struct Cell0 {
  double u, p, rho, dpc, kappa, psi, fsc;
};

__device__ __forceinline__ Cell0 cell0(const RcGeom& g, const RcPtrs& P, int64_t b, int64_t st,
                                       int m) {
  const int64_t o = b + (int64_t)m * st;
  Cell0 c;
  c.u = P.u[0][o];
  c.p = P.p[o];
  c.rho = P.rho[o];
  c.dpc = dp_cell(g, P, 0, b, st, m, P.dp + (int64_t)g.k[0] * g.total);
  c.kappa = c.psi = c.fsc = 0.0;
  if (g.bf) {
    c.kappa = P.kappa[o];
    c.psi = P.psi[o];
    c.fsc = fs_cell(g, P, 0, b, st, m, P.dpsi + (int64_t)g.k[0] * g.total);
  }
  return c;
}

// face_flux of slot 0 from the two cell states (same expression order)
__device__ __forceinline__ double face0(const RcGeom& g, const Cell0& L, const Cell0& R) {
  const double h = g.h[0];
  const double u_bar = 0.5 * (L.u + R.u);
  const double dp_face = (R.p - L.p) / h;
  const double dp_bar = 0.5 * (L.dpc + R.dpc);
  const double inv_rho = 2.0 / (L.rho + R.rho);
  double bracket = dp_face - dp_bar;
  if (g.bf) {
    const double kappa_face = 0.5 * (L.kappa + R.kappa);
    const double f_face = kappa_face * (R.psi - L.psi) / (h * g.we);
    const double f_bar = 0.5 * (L.fsc + R.fsc);
    bracket = (dp_face - dp_bar) - (f_face - f_bar);
  }
  return u_bar - g.dt * inv_rho * bracket;
}

__global__ __launch_bounds__(kBlock) void k_rc_div_slab(RcGeom g, RcPtrs P, unsigned nx,
                                                        double* __restrict__ out) {
  const unsigned bx = blockIdx.x % nx, by = blockIdx.x / nx;
  const int plane = g.n[1] * g.n[2];
  const int j = (int)(bx * kBlock + threadIdx.x);
  if (j >= plane) return;
  int c[3];
  c[1] = j / g.n[2];
  c[2] = j - c[1] * g.n[2];
  const int N0 = g.n[0] - 1;
  const bool per0 = g.periodic[0] != 0;
  const int64_t st0 = plane;
  const int s0 = (int)by * kSlab;
  const int s1 = min(N0 + 1, s0 + kSlab);
  Cell0 cur = cell0(g, P, j, st0, s0);
  // flux of the face below plane s0: flux_k(s0)
  double f_lo;
  if (s0 == 0) f_lo = per0 ? face0(g, cell0(g, P, j, st0, N0), cur) : 0.0;
  else f_lo = face0(g, cell0(g, P, j, st0, s0 - 1), cur);
  for (int i0 = s0; i0 < s1; ++i0) {
    const int64_t idx = (int64_t)i0 * plane + j;
    c[0] = i0;
    // slot 0: the face above (flux_k(i0 + 1), or the wrap / zero face at N)
    Cell0 next = cur;
    double f_up;
    if (i0 < N0) {
      next = cell0(g, P, j, st0, i0 + 1);
      f_up = face0(g, cur, next);
    } else {
      f_up = per0 ? face0(g, cur, cell0(g, P, j, st0, 0)) : 0.0;
    }
    double acc = 0.0;
    double d;
    if (i0 < N0) d = (f_up - f_lo) / g.h[0];
    else if (per0) d = (f_up - f_lo) / g.h[0];
    else d = -f_lo / g.h[0];
    acc += d;
    {  // slot 1, per node
      const int N = g.n[1] - 1;
      const int i = c[1];
      const int64_t st = stride_of(g, 1);
      const int64_t b = idx - (int64_t)i * st;
      double da;
      if (i < N) da = (flux_k(g, P, 1, b, st, i + 1) - flux_k(g, P, 1, b, st, i)) / g.h[1];
      else if (g.periodic[1]) da = (flux_k(g, P, 1, b, st, 0) - flux_k(g, P, 1, b, st, N)) / g.h[1];
      else da = -flux_k(g, P, 1, b, st, N) / g.h[1];
      acc += da;
    }
    {  // slot 2 (contiguous): the face below node i is the face above node
       // i - 1, computed by the previous lane (same row when i > 0)
      const int N = g.n[2] - 1;
      const int i = c[2];
      const int64_t b = idx - i;
      const double up = i < N ? flux_k(g, P, 2, b, 1, i + 1) : flux_k(g, P, 2, b, 1, 0);
      const double from_prev = __shfl_up(up, 1, 64);
      const int lane = threadIdx.x & 63;
      const double lo = (i > 0 && lane > 0) ? from_prev : flux_k(g, P, 2, b, 1, i);
      double da;
      if (i < N) da = (up - lo) / g.h[2];
      else if (g.periodic[2]) da = (up - lo) / g.h[2];
      else da = -lo / g.h[2];
      acc += da;
    }
    out[idx] = acc;
    f_lo = f_up;
    cur = next;
  }
}

}  // namespace
}  // namespace tpf

using namespace tpf;

extern "C" tpf_status tpf_rhie_chow_divergence(int ndim, const int64_t* n_nodes, const double* L,
                                               const int* bc, const double* const* vel,
                                               const double* p, const double* rho,
                                               const double* dp, const double* kappa,
                                               const double* psi, const double* dpsi, double we,
                                               double dt, double* out, void* stream) {
  if (!n_nodes || !L || !bc || !vel || !p || !rho || !dp || !out)
    return fail(TPF_ERR_VALUE, "null argument");
  if (ndim < 1 || ndim > 3) return fail(TPF_ERR_VALUE, "ndim must be 1, 2 or 3");
  const bool bf = kappa || psi || dpsi;
  if (bf && !(kappa && psi && dpsi && we != 0.0))
    return fail(TPF_ERR_VALUE, "the balanced-force term needs kappa, psi, dpsi and we != 0");
  RcGeom g{};
  RcPtrs P{};
  for (int a = 0; a < 3; ++a) {
    g.n[a] = 1;
    g.h[a] = 1.0;
  }
  for (int a = 0; a < ndim; ++a) {
    const int s = 3 - ndim + a;
    if (n_nodes[a] < 3 || !(L[a] > 0.0) || !vel[a])
      return fail(TPF_ERR_VALUE, "Rhie-Chow axis needs >= 3 nodes, L > 0 and a velocity");
    g.n[s] = (int)n_nodes[a];
    g.periodic[s] = bc[a] == TPF_BC_PERIODIC;
    g.k[s] = a;
    g.h[s] = L[a] / (double)(n_nodes[a] - 1);
    P.u[s] = vel[a];
  }
  g.total = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (g.total >= (int64_t)0xffffffffLL) return fail(TPF_ERR_VALUE, "grid too large (2^32 nodes)");
  g.bf = bf ? 1 : 0;
  g.we = we;
  g.dt = dt;
  P.p = p;
  P.rho = rho;
  P.dp = dp;
  P.kappa = kappa;
  P.psi = psi;
  P.dpsi = dpsi;
  if (ndim == 3) {
    const unsigned nx = (unsigned)((g.n[1] * g.n[2] + kBlock - 1) / kBlock);
    const unsigned ny = (unsigned)((g.n[0] + kSlab - 1) / kSlab);
    hipLaunchKernelGGL(k_rc_div_slab, dim3(nx * ny), dim3(kBlock), 0, (hipStream_t)stream, g, P,
                       nx, out);
    TPF_HIP(hipGetLastError());
    return TPF_OK;
  }
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((g.total + kBlock - 1) / kBlock, 2048));
  hipLaunchKernelGGL(k_rc_div, dim3((unsigned)blocks), dim3(kBlock), 0, (hipStream_t)stream, g, P,
                     out);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}
