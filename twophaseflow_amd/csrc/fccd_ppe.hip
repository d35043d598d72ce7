// FCCD matrix-free PPE: L p = D_f[(1/rho)_f G_f p] with the FCCD face
// gradient G_f p = (p_hi - p_lo)/H - (H/24)(q_hi - q_lo), q the nodal CCD
// second derivative along the face's axis, solved by restarted GMRES.
//
// Reference (paths relative to src/twophase): ppe/fccd_matrixfree.py
// (PPESolverFCCDMatrixFree: apply :378-397, operator core :399-417, weighted
// divergence :769-818, solve :505-600, pin and periodic image rows
// :300-376), ccd/fccd.py:243-277 (face_gradient), :464-502 (periodic face
// divergence), ccd/fccd_helpers.py:17-35 (uniform weights),
// ppe/fccd_matrixfree_helpers.py:35-51 (node widths), :207-248 (face
// coefficients 2/(rho_lo + rho_hi)), :251-264 (Jacobi inverse).  Built for
// the full_face boundary space and the legacy operator pairing on uniform
// grids, with the coefficient schemes and interface couplings of
// fccd_matrixfree.py:141-186:
//  * phase_density (the class default): coefficients formed from rho inside
//    the matvec;
//  * phase_separated (the pipeline default, ns_pipeline.py:195-196): face
//    coefficients precomputed per solve (zero across the density threshold
//    0.5 (rho_min + rho_max), or with affine_jump the cut-face resistance
//    1/(theta rho_lo + (1 - theta) rho_hi) of interface_stress_closure.py:
//    225-277); without affine coupling the gauge is the per-phase
//    control-volume mean (fccd_matrixfree_helpers.py:99-160; apply
//    P L P + (I - P), fccd_matrixfree.py:378-397), with jump_decomposition
//    the solve takes rhs - L J and returns p + J, J = j (1 - psi)
//    (helpers.py:281-337), with affine_jump the right-hand side gains
//    D_f[alpha_f B_Gamma(j)] (fccd_matrixfree.py:462-504,
//    interface_stress_closure.py:285-401).
//
// One matvec = one all-axes CCD launch writing q only (d1 NULL), then one
// streaming kernel that forms both face gradients of every axis at a node
// from p, q and rho (neighbours come through L1/L2) and accumulates the
// weighted divergence; the face arrays the reference materialises per axis
// never reach HBM.  Image nodes of periodic axes never feed a non-image row
// (the CCD and the face stencils use the unique nodes 0..N-1 along a
// periodic axis, and the other axes' lines through a non-image node hold
// only non-image nodes), so the reference's image synchronisation of p
// before the core is not needed: the image rows are overwritten by
// p_img - p_src afterwards.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "gmres.hpp"
#include "ppe_impl.hpp"
#include "tpf_common.hpp"

namespace tpf {
namespace {

constexpr int kBlock = 256;

// Node grid right-aligned into 3 slots (leading slots n = 1); slot 2 is the
// contiguous axis.
struct FccdGeom {
  int n[3];
  int periodic[3];
  int qk[3];          // index of the slot's axis in the q stack
  int64_t total;
  int64_t pin;
  double inv_H[3];    // 1 / H, H = L / N (face width)
  double H24[3];      // H / 24
  double inv_w[3];    // 1 / node width, interior
  double inv_w0[3];   // 1 / node width at a wall node (H / 2)
  double h_min[3];    // Jacobi: smallest face width of the axis
};

__device__ __forceinline__ void coords(const FccdGeom& g, int64_t idx, int (&c)[3]) {
  const uint32_t plane = (uint32_t)g.n[1] * (uint32_t)g.n[2];
  const uint32_t u = (uint32_t)idx;
  c[0] = (int)(u / plane);
  const uint32_t rem = u - (uint32_t)c[0] * plane;
  c[1] = (int)(rem / (uint32_t)g.n[2]);
  c[2] = (int)(rem - (uint32_t)c[1] * (uint32_t)g.n[2]);
}

__device__ __forceinline__ int64_t stride_of(const FccdGeom& g, int a) {
  return a == 0 ? (int64_t)g.n[1] * g.n[2] : (a == 1 ? (int64_t)g.n[2] : 1);
}

// Source of an image node: the first periodic axis (in axis order) on which
// the node sits at index N sets that index to 0 (fccd_matrixfree.py:300-345).
// Returns -1 for a non-image node.
__device__ __forceinline__ int64_t image_source(const FccdGeom& g, int64_t idx,
                                                const int (&c)[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (g.periodic[a] && g.n[a] > 1 && c[a] == g.n[a] - 1)
      return idx - (int64_t)c[a] * stride_of(g, a);
  return -1;
}

// Values one axis-0 plane contributes to a node's stencil.
struct Plane0 {
  double p, q, rho;
};

// Contribution of one axis to node idx (value c along it, stride st): the
// weighted divergence of the FCCD face gradients on its two faces
// (fccd_matrixfree.py:769-818; periodic: ccd/fccd.py:490-502).  lo / hi hold
// the neighbours' p and q (wrapped on periodic axes); rho_lo is the lower
// neighbour's rho, rho_hi the upper face's partner rho (never wrapped), and
// rho_lo_partner the lower face's upper rho (the node itself, or node N for
// face N-1 of a periodic axis).
__device__ __forceinline__ double axis_term(const FccdGeom& g, int a, int i, double pc, double qc,
                                            double rc, const Plane0& lo, const Plane0& hi,
                                            double rho_lo_partner) {
  const int N = g.n[a] - 1;
  const double iH = g.inv_H[a], H24 = g.H24[a];
  if (g.periodic[a]) {
    const double g_hi = (hi.p - pc) * iH - H24 * (hi.q - qc);
    const double F_hi = (2.0 / (rc + hi.rho)) * g_hi;
    const double g_lo = (pc - lo.p) * iH - H24 * (qc - lo.q);
    const double F_lo = (2.0 / (lo.rho + rho_lo_partner)) * g_lo;
    return (F_hi - F_lo) * iH;
  }
  if (i == 0) {
    const double g_hi = (hi.p - pc) * iH - H24 * (hi.q - qc);
    return (2.0 / (rc + hi.rho)) * g_hi * g.inv_w0[a];
  }
  if (i == N) {
    const double g_lo = (pc - lo.p) * iH - H24 * (qc - lo.q);
    return -((2.0 / (lo.rho + rc)) * g_lo * g.inv_w0[a]);
  }
  const double g_hi = (hi.p - pc) * iH - H24 * (hi.q - qc);
  const double g_lo = (pc - lo.p) * iH - H24 * (qc - lo.q);
  const double c_hi = 2.0 / (rc + hi.rho);
  const double c_lo = 2.0 / (lo.rho + rc);
  return (c_hi * g_hi - c_lo * g_lo) * g.inv_w[a];
}

// Neighbours of node idx along slot a (1 or 2), loaded unconditionally from
// clamped indices so that every load of a plane step issues before the first
// use (the rows involved are being read by this and the adjacent
// workgroups).  hi: upper value partner (wrapped on periodic axes) with the
// upper face's rho partner (never wrapped); lo: lower neighbour (wrapped);
// rlp: the lower face's upper rho (node N for face N-1 of a periodic axis,
// else the node itself).
struct Nbr {
  Plane0 lo, hi;
  double rlp;
};

// x[i] with a 32-bit element index (the field has < 2^29 elements): a
// uniform base plus a 32-bit byte offset, one address register per load.
__device__ __forceinline__ double at32(const double* __restrict__ x, uint32_t i) {
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(x) + (i << 3));
}

template <typename I>
__device__ __forceinline__ double at(const double* __restrict__ x, I i) {
  if constexpr (sizeof(I) == 4) return at32(x, i);
  else return x[i];
}

// The same term with precomputed face coefficients: c_hi of the face to the
// upper neighbour, c_lo of the face from the lower one (the wrapped face N-1
// at node 0 of a periodic axis).
__device__ __forceinline__ double axis_term_face(const FccdGeom& g, int a, int i, double pc,
                                                 double qc, const Plane0& lo, const Plane0& hi,
                                                 double c_lo, double c_hi) {
  const int N = g.n[a] - 1;
  const double iH = g.inv_H[a], H24 = g.H24[a];
  if (g.periodic[a]) {
    const double g_hi = (hi.p - pc) * iH - H24 * (hi.q - qc);
    const double g_lo = (pc - lo.p) * iH - H24 * (qc - lo.q);
    return (c_hi * g_hi - c_lo * g_lo) * iH;
  }
  if (i == 0) {
    const double g_hi = (hi.p - pc) * iH - H24 * (hi.q - qc);
    return c_hi * g_hi * g.inv_w0[a];
  }
  if (i == N) {
    const double g_lo = (pc - lo.p) * iH - H24 * (qc - lo.q);
    return -(c_lo * g_lo * g.inv_w0[a]);
  }
  const double g_hi = (hi.p - pc) * iH - H24 * (hi.q - qc);
  const double g_lo = (pc - lo.p) * iH - H24 * (qc - lo.q);
  return (c_hi * g_hi - c_lo * g_lo) * g.inv_w[a];
}

template <typename I>
__device__ __forceinline__ Nbr load_nbr(const FccdGeom& g, int a, int i, I idx,
                                        const double* __restrict__ p,
                                        const double* __restrict__ qa,
                                        const double* __restrict__ rho) {
  const int N = g.n[a] - 1;
  const I st = (I)stride_of(g, a);
  const bool per = g.periodic[a] != 0;
  const I hv = i < N ? ((per && i + 1 == N) ? idx - (I)i * st : idx + st) : idx;
  const I hr = i < N ? idx + st : idx;
  const I lv = i > 0 ? idx - st : (per ? idx + (I)(N - 1) * st : idx);
  const I lr = (i == 0 && per) ? idx + (I)N * st : idx;
  Nbr b;
  b.hi = Plane0{at(p, hv), at(qa, hv), at(rho, hr)};
  b.lo = Plane0{at(p, lv), at(qa, lv), at(rho, lv)};
  b.rlp = at(rho, lr);
  return b;
}

// out = L p with the pin row p[pin] and image rows p_img - p_src.
// Thread = one node of an axis-0 plane (coalesced over the plane); the
// workgroup marches kSlab planes along axis 0 keeping the previous, current
// and next planes' p, q_0 and rho in registers, so each of them is read from
// HBM once per slab (plus one halo plane on each side).  Each plane step
// issues all of its loads (next plane, both neighbours on slots 1 and 2)
// before any arithmetic: one memory round trip per step.  Workgroups are
// renumbered so each XCD gets a contiguous range of plane chunks (the axis-1
// neighbours of a chunk lie in the adjacent chunks).  S0 = 3 - ndim: the
// first active slot.
constexpr int kSlab = 16;

// FACE: face coefficients from cf (ndim x total, the face j+1/2 of an axis
// at node j) instead of 2/(rho_lo + rho_hi); the `rho` slots of the plane
// values then hold axis 0's face coefficients.
template <int S0, typename I, bool FACE>
__global__ __launch_bounds__(kBlock) void k_fccd_apply(FccdGeom g, const double* __restrict__ p,
                                                       const double* __restrict__ q,
                                                       const double* __restrict__ rho,
                                                       const double* __restrict__ cf,
                                                       double* __restrict__ out, unsigned nx) {
  constexpr bool has0 = S0 == 0, has1 = S0 <= 1;
  const unsigned nb = gridDim.x, per = nb / 8;
  const unsigned lb = blockIdx.x < per * 8 ? (blockIdx.x % 8) * per + blockIdx.x / 8 : blockIdx.x;
  const unsigned bx = lb % nx, by = lb / nx;
  const unsigned plane = (unsigned)g.n[1] * (unsigned)g.n[2];  // < 2^32 (host guard)
  const unsigned j = bx * kBlock + threadIdx.x;
  if (j >= plane) return;
  const int c1 = (int)(j / (unsigned)g.n[2]), c2 = (int)(j - (unsigned)c1 * (unsigned)g.n[2]);
  const int n0 = g.n[0], N0 = n0 - 1;
  const bool per0 = has0 && g.periodic[0] != 0;
  const int s0 = by * kSlab;
  const int s1 = min(n0, s0 + kSlab);
  const double* q0 = q + (int64_t)g.qk[0] * g.total;
  const double* q1 = q + (int64_t)g.qk[1] * g.total;
  const double* q2 = q + (int64_t)g.qk[2] * g.total;
  const double* c0_ = FACE ? cf + (int64_t)g.qk[0] * g.total : rho;
  const double* c1_ = FACE ? cf + (int64_t)g.qk[1] * g.total : rho;
  const double* c2_ = FACE ? cf + (int64_t)g.qk[2] * g.total : rho;
  // image along slot 1 or 2 (the first periodic slot wins; slot 0 is checked
  // per plane below and comes first)
  I img12 = 0;  // 0: not an image along slots 1 / 2
  if (has1 && g.periodic[1] && c1 == g.n[1] - 1) img12 = (I)c1 * g.n[2];
  else if (g.periodic[2] && c2 == g.n[2] - 1) img12 = (I)c2;
  // FACE: rho slot = axis-0 face coefficient of the value plane (the plane's
  // face to its upper neighbour)
  auto ld = [&](int vplane, int rplane) {
    const I v = (I)vplane * plane + j, r = (I)rplane * plane + j;
    return Plane0{at(p, v), at(q0, v), FACE ? at(c0_, v) : at(rho, r)};
  };
  Plane0 prev{0.0, 0.0, 0.0}, cur = ld(s0, s0), next{0.0, 0.0, 0.0};
  double rhoN0 = 0.0;  // rho of plane N (the upper partner of periodic face N-1)
  if (has0) {
    if (s0 > 0) prev = ld(s0 - 1, s0 - 1);
    else if (per0) prev = ld(N0 - 1, N0 - 1);
    if (per0 && s0 == 0) rhoN0 = at(rho, (I)N0 * plane + j);
  }
  for (int i0 = s0; i0 < s1; ++i0) {
    const I idx = (I)i0 * plane + j;
    // ---- loads
    if (has0) {
      const int vn = i0 < N0 ? ((per0 && i0 + 1 == N0) ? 0 : i0 + 1) : i0;
      next = ld(vn, i0 < N0 ? i0 + 1 : i0);
    }
    Nbr b1{}, b2;
    double q1c = 0.0, f1hi = 0.0, f1lo = 0.0;
    if (has1) {
      b1 = load_nbr(g, 1, c1, idx, p, q1, rho);
      q1c = at(q1, idx);
      if (FACE) {  // faces to the upper and from the lower neighbour (wrapped)
        const int N1 = g.n[1] - 1;
        const I st1 = (I)g.n[2];
        f1hi = c1 < N1 ? at(c1_, idx) : 0.0;
        f1lo = c1 > 0 ? at(c1_, idx - st1) : (g.periodic[1] ? at(c1_, idx + (I)(N1 - 1) * st1) : 0.0);
      }
    }
    b2 = load_nbr(g, 2, c2, idx, p, q2, rho);
    const double q2c = at(q2, idx);
    double f2hi = 0.0, f2lo = 0.0;
    if (FACE) {
      const int N2 = g.n[2] - 1;
      f2hi = c2 < N2 ? at(c2_, idx) : 0.0;
      f2lo = c2 > 0 ? at(c2_, idx - 1) : (g.periodic[2] ? at(c2_, idx + (I)(N2 - 1)) : 0.0);
    }
    const double pimg = per0 && i0 == N0 ? at(p, (I)j) : (img12 > 0 ? at(p, idx - img12) : 0.0);
    // ---- arithmetic
    double res;
    if (per0 && i0 == N0) {
      res = at(p, idx) - pimg;  // image along axis 0 (cur holds plane 0's wrapped values)
    } else if (img12 > 0) {
      res = cur.p - pimg;
    } else if ((int64_t)idx == g.pin) {
      res = cur.p;
    } else {
      const double pc = cur.p, rc = cur.rho;
      double acc = 0.0;
      if (FACE) {
        // axis 0: cur.rho = the face to plane i0 + 1, prev.rho = the face from
        // plane i0 - 1 (plane N0 - 1 on the periodic wrap)
        if (has0) acc += axis_term_face(g, 0, i0, pc, cur.q, prev, next,
                                        i0 > 0 || per0 ? prev.rho : 0.0, i0 < N0 ? cur.rho : 0.0);
        if (has1) acc += axis_term_face(g, 1, c1, pc, q1c, b1.lo, b1.hi, f1lo, f1hi);
        acc += axis_term_face(g, 2, c2, pc, q2c, b2.lo, b2.hi, f2lo, f2hi);
      } else {
        if (has0) acc += axis_term(g, 0, i0, pc, cur.q, rc, prev, next,
                                   (per0 && i0 == 0) ? rhoN0 : rc);
        if (has1) acc += axis_term(g, 1, c1, pc, q1c, rc, b1.lo, b1.hi, b1.rlp);
        acc += axis_term(g, 2, c2, pc, q2c, rc, b2.lo, b2.hi, b2.rlp);
      }
      res = acc;
    }
    __builtin_nontemporal_store(res, &out[(int64_t)idx]);
    prev = cur;
    cur = next;
  }
}

// Jacobi inverse (fccd_matrixfree_helpers.py:251-264): diag = -sum_axes
// 2/(rho h_min^2), diag[pin] = 1, |diag| <= 1e-30 -> 1.
__global__ __launch_bounds__(kBlock) void k_fccd_jacobi(FccdGeom g, const double* __restrict__ rho,
                                                        double* __restrict__ dinv) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    const double r = rho[idx];
    double diag = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (g.n[a] > 1) diag -= 2.0 / (r * g.h_min[a] * g.h_min[a]);
    if (idx == g.pin) diag = 1.0;
    dinv[idx] = 1.0 / (fabs(diag) > 1e-30 ? diag : 1.0);
  }
}

// z = r * dinv, z[pin] = 0 (fccd_matrixfree.py:559-566).
__global__ __launch_bounds__(kBlock) void k_fccd_precond(int64_t n, int64_t pin,
                                                         const double* __restrict__ dinv,
                                                         const double* __restrict__ r,
                                                         double* __restrict__ z) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    z[i] = i == pin ? 0.0 : r[i] * dinv[i];
}

// x = gauge(src): x[pin] = 0, image nodes take their sources' values from
// `src` (a copy of x: the reference's fancy assignment reads every source
// before writing).
__global__ __launch_bounds__(kBlock) void k_fccd_gauge(FccdGeom g, const double* __restrict__ src,
                                                       double* __restrict__ x) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords(g, idx, c);
    const int64_t s = image_source(g, idx, c);
    x[idx] = idx == g.pin ? 0.0 : (s >= 0 ? src[s] : src[idx]);
  }
}

// Multigrid preconditioner input: b32 = -W r, the CG-system form (-W L) z =
// -W r of the FVM operator's hierarchy (W = product of node widths, half
// cells at walls; fvm_matrixfree.py:167-177), b32[pin] = 0.
__global__ __launch_bounds__(kBlock) void k_fccd_to_cg32(FccdGeom g, const double* __restrict__ r,
                                                         float* __restrict__ b32) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords(g, idx, c);
    double W = 1.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (g.n[a] > 1) W *= (c[a] == 0 || c[a] == g.n[a] - 1) ? 0.5 / g.inv_H[a] : 1.0 / g.inv_H[a];
    b32[idx] = idx == g.pin ? 0.0f : (float)(-W * r[idx]);
  }
}

// z = z32 in FP64, z[pin] = 0.
__global__ __launch_bounds__(kBlock) void k_fccd_from32(int64_t n, int64_t pin,
                                                        const float* __restrict__ z32,
                                                        double* __restrict__ z) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    z[i] = i == pin ? 0.0 : (double)z32[i];
}

__global__ __launch_bounds__(kBlock) void k_fccd_fill(int64_t n, double v, double* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    x[i] = v;
}

// y = x - y
__global__ __launch_bounds__(kBlock) void k_fccd_xmy(int64_t n, const double* __restrict__ x,
                                                     double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    y[i] = x[i] - y[i];
}

// sum of squares, per block then one block (deterministic order)
__global__ __launch_bounds__(kBlock) void k_fccd_sumsq(int64_t n, const double* __restrict__ x,
                                                       double* __restrict__ part) {
  __shared__ double sh[kBlock];
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    v += x[i] * x[i];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ __launch_bounds__(kBlock) void k_fccd_sumsq_fin(int nb, double* __restrict__ part) {
  __shared__ double sh[kBlock];
  double v = 0.0;
  for (int i = threadIdx.x; i < nb; i += kBlock) v += part[i];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[nb] = sh[0];
}

// ---- phase-separated forms ----------------------------------------------

// Control volume of a node (core/grid.py:1146-1170 on uniform grids): H per
// node on wall axes; on periodic axes H on nodes 0..N-1 and 0 on the image
// plane.
__device__ __forceinline__ double cell_volume(const FccdGeom& g, const int (&c)[3]) {
  double v = 1.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (g.n[a] > 1) v *= (g.periodic[a] && c[a] == g.n[a] - 1) ? 0.0 : 1.0 / g.inv_H[a];
  return v;
}

// Block partials of min and max of x.
__global__ __launch_bounds__(kBlock) void k_fccd_minmax(int64_t n, const double* __restrict__ x,
                                                        double* __restrict__ part) {
  __shared__ double lo[kBlock], hi[kBlock];
  double a = INFINITY, b = -INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    a = fmin(a, x[i]);
    b = fmax(b, x[i]);
  }
  lo[threadIdx.x] = a;
  hi[threadIdx.x] = b;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + s]);
      hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = lo[0];
    part[2 * blockIdx.x + 1] = hi[0];
  }
}
__global__ __launch_bounds__(kBlock) void k_fccd_minmax_fin(int nb, const double* __restrict__ part,
                                                            double* __restrict__ out) {
  __shared__ double lo[kBlock], hi[kBlock];
  double a = INFINITY, b = -INFINITY;
  for (int i = threadIdx.x; i < nb; i += kBlock) {
    a = fmin(a, part[2 * i]);
    b = fmax(b, part[2 * i + 1]);
  }
  lo[threadIdx.x] = a;
  hi[threadIdx.x] = b;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + s]);
      hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = lo[0];
    out[1] = hi[0];
  }
}

// Face coefficients of every axis at the face's lower node (the last node of
// a line holds 0): mode 1 zero across the density threshold
// (fccd_matrixfree_helpers.py:241-248), mode 2 the affine cut-face
// resistance where psi crosses psi_thr (interface_stress_closure.py:253-277),
// 2/(rho_lo + rho_hi) elsewhere.
__global__ __launch_bounds__(kBlock) void k_fccd_face_coef(FccdGeom g, int mode,
                                                           const double* __restrict__ rho,
                                                           double rho_thr,
                                                           const double* __restrict__ psi,
                                                           double psi_thr,
                                                           double* __restrict__ cf) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords(g, idx, c);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (g.n[a] == 1) continue;
      double v = 0.0;
      if (c[a] < g.n[a] - 1) {
        const int64_t up = idx + stride_of(g, a);
        const double rl = rho[idx], rh = rho[up];
        v = 2.0 / (rl + rh);
        if (mode == 1) {
          if ((rl >= rho_thr) != (rh >= rho_thr)) v = 0.0;
        } else if (mode == 2) {
          const double pl = psi[idx], ph = psi[up];
          if ((pl < psi_thr) != (ph < psi_thr)) {
            const double theta = (psi_thr - pl) / (ph - pl);
            v = 1.0 / (theta * rl + (1.0 - theta) * rh);
          }
        }
      }
      cf[(int64_t)g.qk[a] * g.total + idx] = v;
    }
  }
}

// Per-phase control-volume sums of x (x NULL: of 1), gas = rho < thr:
// block partials (gas, liquid).
__global__ __launch_bounds__(kBlock) void k_fccd_phase_sums(FccdGeom g,
                                                            const double* __restrict__ rho,
                                                            double thr,
                                                            const double* __restrict__ x,
                                                            double* __restrict__ part) {
  __shared__ double sg[kBlock], sl[kBlock];
  double a = 0.0, b = 0.0;
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords(g, idx, c);
    const double wv = cell_volume(g, c) * (x ? x[idx] : 1.0);
    if (rho[idx] < thr) a += wv;
    else b += wv;
  }
  sg[threadIdx.x] = a;
  sl[threadIdx.x] = b;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sg[threadIdx.x] += sg[threadIdx.x + s];
      sl[threadIdx.x] += sl[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sg[0];
    part[2 * blockIdx.x + 1] = sl[0];
  }
}

// out[0..1] = per-phase sums / wsum[0..1] (wsum NULL: the sums themselves).
__global__ __launch_bounds__(kBlock) void k_fccd_phase_fin(int nb, const double* __restrict__ part,
                                                           const double* __restrict__ wsum,
                                                           double* __restrict__ out) {
  __shared__ double sg[kBlock], sl[kBlock];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nb; i += kBlock) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  sg[threadIdx.x] = a;
  sl[threadIdx.x] = b;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sg[threadIdx.x] += sg[threadIdx.x + s];
      sl[threadIdx.x] += sl[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = wsum ? sg[0] / wsum[0] : sg[0];
    out[1] = wsum ? sl[0] / wsum[1] : sl[0];
  }
}

// out = x - (gas ? m0 : m1) (fccd_matrixfree_helpers.py:137-160).
__global__ __launch_bounds__(kBlock) void k_fccd_phase_sub(int64_t n,
                                                           const double* __restrict__ rho,
                                                           double thr,
                                                           const double* __restrict__ means,
                                                           const double* x, double* out) {
  const double m0 = means[0], m1 = means[1];
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    out[i] = x[i] - (rho[i] < thr ? m0 : m1);
}

// Gauged apply, last step (fccd_matrixfree.py:386-396):
// out = (core - mean(core)) + p - p_mf; image rows p_img - p_src.
__global__ __launch_bounds__(kBlock) void k_fccd_gauge_finish(FccdGeom g,
                                                              const double* __restrict__ rho,
                                                              double thr,
                                                              const double* __restrict__ means,
                                                              const double* __restrict__ p,
                                                              const double* __restrict__ pm,
                                                              double* __restrict__ out) {
  const double m0 = means[0], m1 = means[1];
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords(g, idx, c);
    const int64_t src = image_source(g, idx, c);
    if (src >= 0) {
      out[idx] = p[idx] - p[src];
    } else {
      const double v = out[idx] - (rho[idx] < thr ? m0 : m1);
      out[idx] = v + p[idx] - pm[idx];
    }
  }
}

// J = j (1 - psi) (fccd_matrixfree_helpers.py:318-337); out = base + J.
__global__ __launch_bounds__(kBlock) void k_fccd_add_jump(int64_t n,
                                                          const double* __restrict__ jump,
                                                          const double* __restrict__ psi,
                                                          const double* base, double* out) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    out[i] = (base ? base[i] : 0.0) + jump[i] * (1.0 - psi[i]);
}

// Signed cut-face jump B_Gamma(j) of the face between lo and hi along slot a
// (interface_stress_closure.py:285-401): (gas_hi - gas_lo) j_f / H on cut
// faces, j_f = -sigma kappa at the psi cut point (kappa given), else the
// face average of the nodal jump.
__device__ __forceinline__ double face_jump(const FccdGeom& g, int a, int64_t lo, int64_t hi,
                                            const double* __restrict__ psi,
                                            const double* __restrict__ kappa,
                                            const double* __restrict__ jump, double sigma,
                                            double thr) {
  const double pl = psi[lo], ph = psi[hi];
  const bool gl = pl < thr, gh = ph < thr;
  if (gl == gh) return 0.0;
  double jf;
  if (kappa) {
    const double theta = (thr - pl) / (ph - pl);
    jf = -sigma * ((1.0 - theta) * kappa[lo] + theta * kappa[hi]);
  } else {
    jf = 0.5 * (jump[lo] + jump[hi]);
  }
  const double orient = (gh ? 1.0 : 0.0) - (gl ? 1.0 : 0.0);
  return orient * jf * g.inv_H[a];
}

// b += D_f[alpha_f B_Gamma(j)] (fccd_matrixfree.py:462-504): the face-flux
// divergence with the wall Neumann rows and the periodic wrap, as the apply's.
__global__ __launch_bounds__(kBlock) void k_fccd_affine(FccdGeom g, const double* __restrict__ cf,
                                                        const double* __restrict__ psi,
                                                        const double* __restrict__ kappa,
                                                        const double* __restrict__ jump,
                                                        double sigma, double thr,
                                                        double* __restrict__ b) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords(g, idx, c);
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (g.n[a] == 1) continue;
      const int N = g.n[a] - 1, i = c[a];
      const int64_t st = stride_of(g, a);
      const double* ca = cf + (int64_t)g.qk[a] * g.total;
      auto flux = [&](int64_t lo) {  // face lo -> lo + st
        return ca[lo] * face_jump(g, a, lo, lo + st, psi, kappa, jump, sigma, thr);
      };
      if (g.periodic[a]) {
        if (i == N) continue;  // image row: zeroed afterwards
        const double Fh = flux(idx);
        const double Fl = i > 0 ? flux(idx - st) : flux(idx + (int64_t)(N - 1) * st);
        acc += (Fh - Fl) * g.inv_H[a];
      } else if (i == 0) {
        acc += flux(idx) * g.inv_w0[a];
      } else if (i == N) {
        acc -= flux(idx - st) * g.inv_w0[a];
      } else {
        acc += (flux(idx) - flux(idx - st)) * g.inv_w[a];
      }
    }
    b[idx] += acc;
  }
}

// out = a x
__global__ __launch_bounds__(kBlock) void k_fccd_scale(int64_t n, double a,
                                                       const double* __restrict__ x,
                                                       double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    out[i] = a * x[i];
}

// b[pin] = 0 (pin >= 0) and zero image rows, in place.
__global__ __launch_bounds__(kBlock) void k_fccd_rhs_rows(FccdGeom g, int64_t pin,
                                                          double* __restrict__ b) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords(g, idx, c);
    if (idx == pin || image_source(g, idx, c) >= 0) b[idx] = 0.0;
  }
}

// image nodes take their sources' values (the sources are never images)
__global__ __launch_bounds__(kBlock) void k_fccd_sync_images(FccdGeom g, double* __restrict__ x) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords(g, idx, c);
    const int64_t s = image_source(g, idx, c);
    if (s >= 0) x[idx] = x[s];
  }
}

int grid_for(int64_t total) {
  const int64_t blocks = (total + kBlock - 1) / kBlock;
  const int64_t cap = 256 * 8;
  return (int)std::max<int64_t>(1, std::min(blocks, cap));
}

}  // namespace
}  // namespace tpf

struct tpf_fccd_ppe {
  tpf::FccdGeom g{};
  int ndim = 0;
  hipStream_t stream = nullptr;
  int grid = 1;
  tpf_ccd_plan* ccd = nullptr;
  const double* rho = nullptr;
  double* q = nullptr;       // ndim x total: CCD second derivative per axis
  double* dinv = nullptr;    // Jacobi inverse (built with the density)
  double* part = nullptr;    // norm partials (+1 result)
  double* host = nullptr;    // pinned result
  void* gmres_work = nullptr;
  // multigrid preconditioner (wall grids): the FVM operator's FP32 hierarchy
  // on an FVM handle of the same grid, created on first use
  tpf_ppe* fvm = nullptr;
  const double* fvm_rho = nullptr;
  int64_t n_nodes[3] = {0, 0, 0};
  double L[3] = {0.0, 0.0, 0.0};
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  // coefficient scheme and interface coupling (fccd_matrixfree.py:141-186)
  int scheme = TPF_FCCD_PHASE_DENSITY;
  int coupling = TPF_FCCD_COUPLING_NONE;
  // interface data (caller's device arrays): psi, and kappa (Young-Laplace
  // j = -sigma kappa, cut-point curvature) or a nodal jump
  const double* psi = nullptr;
  const double* kappa = nullptr;
  const double* jump = nullptr;   // j = p_gas - p_liquid at nodes (NULL: -sigma kappa)
  double sigma = 0.0, psi_thr = 0.5;
  // state prepared with the density
  int face_mode = 0;              // 1 zero across the threshold, 2 affine cut faces
  int mean_gauge = 0;             // per-phase mean gauge in place of the pin
  double rho_thr = 0.0;
  double* cf = nullptr;           // ndim x total face coefficients
  double* pm = nullptr;           // gauge scratch (p_mf) and jump field
  double* jfield = nullptr;       // nodal j = -sigma kappa (Young-Laplace)
  double* ph = nullptr;           // phase partials (2 x grid) + wsum[2] + means[2] + minmax[2]
};

using namespace tpf;

#define FCCD_LAUNCH(kern, ...)                                                       \
  do {                                                                               \
    hipLaunchKernelGGL(kern, dim3(s->grid), dim3(kBlock), 0, s->stream, __VA_ARGS__); \
    TPF_HIP(hipGetLastError());                                                      \
  } while (0)

extern "C" tpf_status tpf_fccd_ppe_destroy(tpf_fccd_ppe* s) {
  if (!s) return TPF_OK;
  gmres_work_free(&s->gmres_work);
  if (s->ccd) tpf_ccd_plan_destroy(s->ccd);
  if (s->fvm) tpf_ppe_destroy(s->fvm);
  if (s->ev_in) (void)hipEventDestroy(s->ev_in);
  if (s->ev_out) (void)hipEventDestroy(s->ev_out);
  double* bufs[] = {s->q, s->dinv, s->part, s->cf, s->pm, s->jfield, s->ph};
  for (double* b : bufs)
    if (b) (void)hipFree(b);
  if (s->host) (void)hipHostFree(s->host);
  delete s;
  return TPF_OK;
}

extern "C" tpf_status tpf_fccd_ppe_create(tpf_fccd_ppe** out, int ndim, const int64_t* n_nodes,
                                          const double* L, const int* bc, int64_t pin_dof,
                                          void* stream) {
  if (!out || !n_nodes || !L || !bc) return fail(TPF_ERR_VALUE, "null argument");
  *out = nullptr;
  if (ndim < 1 || ndim > 3) return fail(TPF_ERR_VALUE, "ndim must be 1, 2 or 3");
  auto* s = new tpf_fccd_ppe();
  FccdGeom& g = s->g;
  s->ndim = ndim;
  double h[3] = {1.0, 1.0, 1.0};
  for (int a = 0; a < 3; ++a) {
    g.n[a] = 1;
    g.periodic[a] = 0;
    g.qk[a] = 0;
  }
  for (int a = 0; a < ndim; ++a) {
    const int slot = 3 - ndim + a;
    if (n_nodes[a] < 3 || !(L[a] > 0.0) || (bc[a] != TPF_BC_WALL && bc[a] != TPF_BC_PERIODIC)) {
      delete s;
      return fail(TPF_ERR_VALUE, "FCCD PPE axis needs >= 3 nodes, L > 0 and a wall/periodic bc");
    }
    g.n[slot] = (int)n_nodes[a];
    s->n_nodes[a] = n_nodes[a];
    s->L[a] = L[a];
    g.periodic[slot] = bc[a] == TPF_BC_PERIODIC;
    g.qk[slot] = a;
    h[slot] = L[a] / (double)(n_nodes[a] - 1);
  }
  g.total = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (g.total >= (int64_t)0xffffffffLL) {
    delete s;
    return fail(TPF_ERR_VALUE, "FCCD PPE grid too large (2^32 nodes)");
  }
  if (pin_dof < 0 || pin_dof >= g.total) {
    delete s;
    return fail(TPF_ERR_VALUE, "pin_dof out of range");
  }
  g.pin = pin_dof;
  for (int a = 0; a < 3; ++a) {
    g.inv_H[a] = 1.0 / h[a];
    g.H24[a] = h[a] / 24.0;
    g.inv_w[a] = 1.0 / h[a];
    g.inv_w0[a] = 1.0 / (0.5 * h[a]);
    g.h_min[a] = h[a];
  }
  s->stream = (hipStream_t)stream;
  s->grid = grid_for(g.total);
  double hh[3];
  int bcs[3];
  for (int a = 0; a < ndim; ++a) {
    hh[a] = h[3 - ndim + a];
    bcs[a] = bc[a];
  }
  tpf_status st = tpf_ccd_plan_create(&s->ccd, ndim, n_nodes, hh, bcs, TPF_PREC_FP64);
  if (st != TPF_OK) {
    s->ccd = nullptr;
    tpf_fccd_ppe_destroy(s);
    return st;  // message from the CCD plan
  }
  const size_t bytes = sizeof(double) * (size_t)g.total;
  if (hipMalloc(&s->q, bytes * ndim) != hipSuccess || hipMalloc(&s->dinv, bytes) != hipSuccess ||
      hipMalloc(&s->part, sizeof(double) * (s->grid + 1)) != hipSuccess ||
      hipHostMalloc(&s->host, sizeof(double)) != hipSuccess) {
    tpf_fccd_ppe_destroy(s);
    return fail(TPF_ERR_RUNTIME, "FCCD PPE workspace allocation failed");
  }
  *out = s;
  return TPF_OK;
}

namespace {
double* ph_wsum(tpf_fccd_ppe* s) { return s->ph + 2 * s->grid; }
double* ph_means(tpf_fccd_ppe* s) { return s->ph + 2 * s->grid + 2; }
double* ph_minmax(tpf_fccd_ppe* s) { return s->ph + 2 * s->grid + 4; }

// The gauge, face coefficients and Jacobi inverse of the current density
// (fccd_matrixfree_lifecycle.py:44-117).  Synchronous when the density
// threshold is needed.
tpf_status fccd_prepare(tpf_fccd_ppe* s) {
  const int64_t n = s->g.total;
  s->face_mode = 0;
  s->mean_gauge = 0;
  const bool sep = s->scheme == TPF_FCCD_PHASE_SEPARATED;
  if (sep && (!s->ph || !s->cf || !s->pm)) {
    const size_t bytes = sizeof(double) * (size_t)n;
    if ((!s->ph && hipMalloc(&s->ph, sizeof(double) * (2 * s->grid + 8)) != hipSuccess) ||
        (!s->cf && hipMalloc(&s->cf, bytes * s->ndim) != hipSuccess) ||
        (!s->pm && hipMalloc(&s->pm, bytes) != hipSuccess))
      return fail(TPF_ERR_RUNTIME, "FCCD PPE phase workspace allocation failed");
  }
  if (sep && s->coupling == TPF_FCCD_AFFINE_JUMP) {
    if (!s->psi)
      return fail(TPF_ERR_RUNTIME, "phase_separated affine-jump face coefficient requires an "
                                   "interface-stress context");
    s->face_mode = 2;
  } else if (sep) {
    // threshold 0.5 (rho_min + rho_max) unless the density is uniform
    // (fccd_matrixfree_helpers.py:56-71)
    FCCD_LAUNCH(k_fccd_minmax, n, s->rho, s->ph);
    hipLaunchKernelGGL(k_fccd_minmax_fin, dim3(1), dim3(kBlock), 0, s->stream, s->grid,
                       (const double*)s->ph, ph_minmax(s));
    TPF_HIP(hipGetLastError());
    double mm[2];
    TPF_HIP(hipMemcpyAsync(mm, ph_minmax(s), sizeof(mm), hipMemcpyDeviceToHost, s->stream));
    TPF_HIP(hipStreamSynchronize(s->stream));
    if (std::isfinite(mm[0] + mm[1]) && std::fabs(mm[1] - mm[0]) >= 1.0e-14) {
      s->rho_thr = 0.5 * (mm[0] + mm[1]);
      s->mean_gauge = 1;
      s->face_mode = 1;
      // per-phase control-volume sums (helpers.py:99-117)
      FCCD_LAUNCH(k_fccd_phase_sums, s->g, s->rho, s->rho_thr, (const double*)nullptr, s->ph);
      hipLaunchKernelGGL(k_fccd_phase_fin, dim3(1), dim3(kBlock), 0, s->stream, s->grid,
                         (const double*)s->ph, (const double*)nullptr, ph_wsum(s));
      TPF_HIP(hipGetLastError());
    }
  }
  if (s->face_mode)
    FCCD_LAUNCH(k_fccd_face_coef, s->g, s->face_mode, s->rho, s->rho_thr, s->psi, s->psi_thr,
                s->cf);
  FccdGeom gj = s->g;
  if (s->mean_gauge) gj.pin = -1;  // no pinned row under the mean gauge
  FCCD_LAUNCH(k_fccd_jacobi, gj, s->rho, s->dinv);
  return TPF_OK;
}
}  // namespace

extern "C" tpf_status tpf_fccd_ppe_set_density(tpf_fccd_ppe* s, const double* rho) {
  if (!s || !rho) return fail(TPF_ERR_VALUE, "null argument");
  s->rho = rho;
  s->fvm_rho = nullptr;  // the multigrid hierarchy follows on next use
  return fccd_prepare(s);
}

extern "C" tpf_status tpf_fccd_ppe_set_scheme(tpf_fccd_ppe* s, int coefficient_scheme,
                                              int interface_coupling) {
  if (!s) return fail(TPF_ERR_VALUE, "null argument");
  if (coefficient_scheme != TPF_FCCD_PHASE_DENSITY &&
      coefficient_scheme != TPF_FCCD_PHASE_SEPARATED)
    return fail(TPF_ERR_VALUE,
                "FCCD PPE supports ppe_coefficient_scheme='phase_density'|'phase_separated'");
  if (interface_coupling < TPF_FCCD_COUPLING_NONE || interface_coupling > TPF_FCCD_AFFINE_JUMP)
    return fail(TPF_ERR_VALUE, "FCCD PPE supports ppe_interface_coupling_scheme="
                               "'none'|'jump_decomposition'|'affine_jump'");
  if (coefficient_scheme == TPF_FCCD_PHASE_DENSITY && interface_coupling != TPF_FCCD_COUPLING_NONE)
    return fail(TPF_ERR_VALUE, "phase_density PPE requires interface_coupling='none'");
  s->scheme = coefficient_scheme;
  s->coupling = interface_coupling;
  return s->rho ? fccd_prepare(s) : TPF_OK;
}

extern "C" tpf_status tpf_fccd_ppe_set_interface(tpf_fccd_ppe* s, const double* psi,
                                                 const double* kappa, const double* jump,
                                                 double sigma, double phase_threshold) {
  if (!s) return fail(TPF_ERR_VALUE, "null argument");
  if (psi && !kappa && !jump) return fail(TPF_ERR_VALUE, "interface data needs kappa or jump");
  s->psi = psi;
  s->kappa = psi ? kappa : nullptr;
  s->jump = psi && !kappa ? jump : nullptr;
  s->sigma = psi && kappa ? sigma : 0.0;
  s->psi_thr = phase_threshold;
  if (psi && kappa) {  // nodal j = -sigma kappa for the jump decomposition
    if (!s->jfield && hipMalloc(&s->jfield, sizeof(double) * (size_t)s->g.total) != hipSuccess)
      return fail(TPF_ERR_RUNTIME, "FCCD PPE jump allocation failed");
    hipLaunchKernelGGL(k_fccd_scale, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g.total,
                       -sigma, kappa, s->jfield);
    TPF_HIP(hipGetLastError());
  }
  return s->rho ? fccd_prepare(s) : TPF_OK;
}

namespace {
using ApplyFn = void (*)(FccdGeom, const double*, const double*, const double*, const double*,
                         double*, unsigned);
template <bool FACE>
ApplyFn pick_apply(int ndim, bool small) {
  if (small)
    return ndim == 3 ? k_fccd_apply<0, uint32_t, FACE>
                     : (ndim == 2 ? k_fccd_apply<1, uint32_t, FACE> : k_fccd_apply<2, uint32_t, FACE>);
  return ndim == 3 ? k_fccd_apply<0, int64_t, FACE>
                   : (ndim == 2 ? k_fccd_apply<1, int64_t, FACE> : k_fccd_apply<2, int64_t, FACE>);
}

// D_f[alpha_f G_f p] with the pin row p[pin] when `pinned`, image rows
// p_img - p_src (fccd_matrixfree.py:399-417).
tpf_status fccd_core(tpf_fccd_ppe* s, const double* p, double* out, bool pinned) {
  const uint32_t mask = (1u << s->ndim) - 1u;
  TPF_TRY(tpf_ccd_apply(s->ccd, p, nullptr, s->q, mask, s->stream));
  const int plane = s->g.n[1] * s->g.n[2];
  const unsigned nx = (unsigned)((plane + kBlock - 1) / kBlock);
  const unsigned ny = (unsigned)((s->g.n[0] + kSlab - 1) / kSlab);
  const bool small = s->g.total < ((int64_t)1 << 29);
  ApplyFn kern = s->face_mode ? pick_apply<true>(s->ndim, small) : pick_apply<false>(s->ndim, small);
  FccdGeom g = s->g;
  if (!pinned) g.pin = -1;
  hipLaunchKernelGGL(kern, dim3(nx * ny), dim3(kBlock), 0, s->stream, g, p,
                     (const double*)s->q, s->rho, (const double*)s->cf, out, nx);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

// per-phase control-volume means of x into ph_means (device only)
tpf_status phase_means(tpf_fccd_ppe* s, const double* x) {
  FCCD_LAUNCH(k_fccd_phase_sums, s->g, s->rho, s->rho_thr, x, s->ph);
  hipLaunchKernelGGL(k_fccd_phase_fin, dim3(1), dim3(kBlock), 0, s->stream, s->grid,
                     (const double*)s->ph, (const double*)ph_wsum(s), ph_means(s));
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

// out = x - <x>_phase (out may be x)
tpf_status project_means(tpf_fccd_ppe* s, const double* x, double* out) {
  TPF_TRY(phase_means(s, x));
  FCCD_LAUNCH(k_fccd_phase_sub, s->g.total, s->rho, s->rho_thr, (const double*)ph_means(s), x,
              out);
  return TPF_OK;
}

// The solver operator: the pinned core, or P L P p + (I - P) p under the
// per-phase mean gauge (fccd_matrixfree.py:378-397).
tpf_status fccd_apply(tpf_fccd_ppe* s, const double* p, double* out) {
  if (!s->mean_gauge) return fccd_core(s, p, out, true);
  TPF_TRY(project_means(s, p, s->pm));
  TPF_TRY(fccd_core(s, s->pm, out, false));
  TPF_TRY(phase_means(s, out));
  FCCD_LAUNCH(k_fccd_gauge_finish, s->g, s->rho, s->rho_thr, (const double*)ph_means(s), p,
              (const double*)s->pm, out);
  return TPF_OK;
}

bool jump_active(const tpf_fccd_ppe* s) {  // fccd_matrixfree_helpers.py:281-292
  return s->scheme == TPF_FCCD_PHASE_SEPARATED && s->coupling == TPF_FCCD_JUMP_DECOMPOSITION &&
         s->psi && s->kappa && s->sigma > 0.0;
}

// b = the right-hand side GMRES iterates on (fccd_matrixfree.py:506-517):
// rhs - L J, + D_f[alpha_f B_Gamma(j)], the phase-mean projection or the pin,
// zero image rows.  `scratch` holds one n-vector.
tpf_status fccd_system_rhs(tpf_fccd_ppe* s, const double* rhs, double* b, double* scratch) {
  const int64_t n = s->g.total;
  if (jump_active(s)) {
    FCCD_LAUNCH(k_fccd_add_jump, n, (const double*)s->jfield, s->psi, (const double*)nullptr,
                scratch);
    TPF_TRY(fccd_core(s, scratch, b, false));
    FCCD_LAUNCH(k_fccd_xmy, n, rhs, b);
  } else {
    TPF_HIP(hipMemcpyAsync(b, rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, s->stream));
  }
  if (s->scheme == TPF_FCCD_PHASE_SEPARATED && s->coupling == TPF_FCCD_AFFINE_JUMP) {
    if (!s->psi)
      return fail(TPF_ERR_RUNTIME, "phase_separated affine-jump FCCD PPE requires an "
                                   "interface-stress context before applying affine RHS");
    FCCD_LAUNCH(k_fccd_affine, s->g, (const double*)s->cf, s->psi, s->kappa, s->jump, s->sigma,
                s->psi_thr, b);
  }
  if (s->mean_gauge) TPF_TRY(project_means(s, b, b));
  FCCD_LAUNCH(k_fccd_rhs_rows, s->g, s->mean_gauge ? (int64_t)-1 : s->g.pin, b);
  return TPF_OK;
}

// Multigrid cycle parameters of the preconditioner: W(2,2), W recursion on
// the top two levels, 20 coarse sweeps (the FVM multigrid-FCG defaults).
constexpr int kMgNu = 2, kMgGamma = 2, kMgCoarse = 20, kMgWLevels = 2;

tpf_status fccd_mg_ready(tpf_fccd_ppe* s) {
  for (int a = 0; a < 3; ++a)
    if (s->g.n[a] > 1 && s->g.periodic[a])
      return fail(TPF_ERR_VALUE, "the multigrid preconditioner needs wall axes");
  if (!s->fvm) {
    TPF_TRY(tpf_ppe_create(&s->fvm, s->ndim, s->n_nodes, s->L, s->g.pin, s->stream));
    TPF_HIP(hipEventCreateWithFlags(&s->ev_in, hipEventDisableTiming));
    TPF_HIP(hipEventCreateWithFlags(&s->ev_out, hipEventDisableTiming));
  }
  if (s->fvm_rho != s->rho) {
    TPF_TRY(tpf_ppe_set_density(s->fvm, s->rho));
    s->fvm_rho = s->rho;
  }
  return mg_ensure(s->fvm, 12);  // synchronous
}

// z = M(r): one FP32 cycle of the FVM hierarchy on -W r, on the FVM handle's
// stream, ordered after and before the FCCD stream by events.
tpf_status fccd_mg_precond(tpf_fccd_ppe* s, const double* r, double* z) {
  hipStream_t fs = s->fvm->stream;
  TPF_HIP(hipEventRecord(s->ev_in, s->stream));
  TPF_HIP(hipStreamWaitEvent(fs, s->ev_in, 0));
  hipLaunchKernelGGL(k_fccd_to_cg32, dim3(s->grid), dim3(kBlock), 0, fs, s->g, r,
                     mg_fine_b(s->fvm));
  TPF_HIP(hipGetLastError());
  TPF_TRY(mg_cycle(s->fvm, fs, kMgNu, kMgGamma, kMgCoarse, kMgWLevels));
  hipLaunchKernelGGL(k_fccd_from32, dim3(s->grid), dim3(kBlock), 0, fs, s->g.total, s->g.pin,
                     (const float*)mg_fine_x(s->fvm), z);
  TPF_HIP(hipGetLastError());
  TPF_HIP(hipEventRecord(s->ev_out, fs));
  TPF_HIP(hipStreamWaitEvent(s->stream, s->ev_out, 0));
  return TPF_OK;
}

tpf_status fccd_norm2(tpf_fccd_ppe* s, const double* x, double& out) {
  FCCD_LAUNCH(k_fccd_sumsq, s->g.total, x, s->part);
  hipLaunchKernelGGL(k_fccd_sumsq_fin, dim3(1), dim3(kBlock), 0, s->stream, s->grid, s->part);
  TPF_HIP(hipGetLastError());
  TPF_HIP(hipMemcpyAsync(s->host, s->part + s->grid, sizeof(double), hipMemcpyDeviceToHost,
                         s->stream));
  TPF_HIP(hipStreamSynchronize(s->stream));
  out = std::sqrt(*s->host);
  return TPF_OK;
}
}  // namespace

extern "C" tpf_status tpf_fccd_ppe_apply(tpf_fccd_ppe* s, const double* p, double* out) {
  if (!s || !p || !out) return fail(TPF_ERR_VALUE, "null argument");
  if (p == out) return fail(TPF_ERR_VALUE, "p and out must be distinct arrays");
  if (!s->rho) return fail(TPF_ERR_RUNTIME, "prepare_operator(rho) must be called before apply(p)");
  return fccd_apply(s, p, out);
}

extern "C" tpf_status tpf_fccd_ppe_solve_gmres(tpf_fccd_ppe* s, const double* rhs, double* x,
                                               const tpf_gmres_opts* o, tpf_solve_report* rep) {
  if (!s || !rhs || !x || !o) return fail(TPF_ERR_VALUE, "null argument");
  if (!s->rho) return fail(TPF_ERR_RUNTIME, "FCCD PPE density not set");
  if (o->precond < 0 || o->precond > 2)
    return fail(TPF_ERR_VALUE,
                "FCCD PPE supports preconditioner 0 (none), 1 (Jacobi) or 2 (multigrid)");
  if (o->precond == 2 && s->mean_gauge)
    return fail(TPF_ERR_VALUE, "the multigrid preconditioner takes the pinned gauge "
                               "(phase_density or affine_jump coupling)");
  if (o->precond == 2) TPF_TRY(fccd_mg_ready(s));
  if (!(o->tol > 0.0) || o->maxiter < 1 || o->restart < 1)
    return fail(TPF_ERR_VALUE, "GMRES needs tol > 0, maxiter >= 1, restart >= 1");
  const int64_t n = s->g.total;
  const int m = (int)std::min<int64_t>(o->restart, n);  // SciPy: min(restart, n)
  TPF_TRY(gmres_work(&s->gmres_work, n, m, s->grid, 1));
  double* b = gmres_rhs(s->gmres_work);
  double* tmp = gmres_scratch(s->gmres_work);
  TPF_TRY(fccd_system_rhs(s, rhs, b, tmp));
  if (o->use_x0) {  // sync images, then pin 0 or the phase means (fccd_matrixfree.py:540-548)
    if (s->mean_gauge) {
      FCCD_LAUNCH(k_fccd_sync_images, s->g, x);
      TPF_TRY(project_means(s, x, x));
    } else {
      TPF_HIP(hipMemcpyAsync(tmp, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s->stream));
      FCCD_LAUNCH(k_fccd_gauge, s->g, (const double*)tmp, x);
    }
  } else {
    FCCD_LAUNCH(k_fccd_fill, n, 0.0, x);
  }
  const int64_t zpin = s->mean_gauge ? (int64_t)-1 : s->g.pin;
  const bool jac = o->precond == 1, mg = o->precond == 2;
  const dim3 grid(s->grid);
  tpf_solve_report r{};
  tpf_status st = gmres_core(
      &s->gmres_work, s->stream, (unsigned)s->grid, n, m,
      [&](const double* in, double* out) { return fccd_apply(s, in, out); },
      [&](const double* in, double* out) -> tpf_status {
        if (mg) return fccd_mg_precond(s, in, out);
        if (jac) {  // z = r D^-1, then the pin or the phase means (fccd_matrixfree.py:562-568)
          hipLaunchKernelGGL(k_fccd_precond, grid, dim3(kBlock), 0, s->stream, n, zpin,
                             (const double*)s->dinv, in, out);
          TPF_HIP(hipGetLastError());
          if (s->mean_gauge) return project_means(s, out, out);
        } else {
          TPF_HIP(hipMemcpyAsync(out, in, sizeof(double) * n, hipMemcpyDeviceToDevice,
                                 s->stream));
        }
        return TPF_OK;
      },
      b, x, o->tol, o->maxiter, r);
  if (st != TPF_OK) return st;
  // sol[pin] = 0 or the phase means, images from sources; then the true
  // residual on that field (fccd_matrixfree.py:584-596)
  if (s->mean_gauge) {
    TPF_TRY(project_means(s, x, x));
    FCCD_LAUNCH(k_fccd_sync_images, s->g, x);
  } else {
    TPF_HIP(hipMemcpyAsync(tmp, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s->stream));
    FCCD_LAUNCH(k_fccd_gauge, s->g, (const double*)tmp, x);
  }
  st = fccd_apply(s, x, tmp);
  if (st == TPF_OK) {
    FCCD_LAUNCH(k_fccd_xmy, n, (const double*)b, tmp);
    double rn = 0.0, bn = 0.0;
    st = fccd_norm2(s, tmp, rn);
    if (st == TPF_OK) st = fccd_norm2(s, b, bn);
    r.true_relres = bn > 0.0 ? rn / bn : rn;
  }
  if (st == TPF_OK && jump_active(s)) {  // p + J, images synced (fccd_matrixfree.py:607-609)
    FCCD_LAUNCH(k_fccd_add_jump, n, (const double*)s->jfield, s->psi, (const double*)x, x);
    FCCD_LAUNCH(k_fccd_sync_images, s->g, x);
  }
  TPF_HIP(hipStreamSynchronize(s->stream));
  if (st != TPF_OK) return st;
  if (rep) *rep = r;
  if (!r.converged)
    return fail(TPF_ERR_RUNTIME, "PPESolverFCCDMatrixFree did not converge: info=" +
                                     std::to_string(o->maxiter) + ", relative_l2=" +
                                     std::to_string(r.true_relres) +
                                     ", maxiter=" + std::to_string(o->maxiter));
  return TPF_OK;
}


extern "C" tpf_status tpf_fccd_ppe_system_rhs(tpf_fccd_ppe* s, const double* rhs, double* b) {
  if (!s || !rhs || !b) return fail(TPF_ERR_VALUE, "null argument");
  if (rhs == b) return fail(TPF_ERR_VALUE, "rhs and b must be distinct arrays");
  if (!s->rho) return fail(TPF_ERR_RUNTIME, "FCCD PPE density not set");
  double* scratch = nullptr;
  if (jump_active(s)) {
    if (!s->gmres_work) TPF_TRY(gmres_work(&s->gmres_work, s->g.total, 1, s->grid, 1));
    scratch = gmres_scratch(s->gmres_work);
  }
  return fccd_system_rhs(s, rhs, b, scratch);
}

extern "C" tpf_status tpf_fccd_ppe_face_coeff(tpf_fccd_ppe* s, int axis, double* out) {
  if (!s || !out) return fail(TPF_ERR_VALUE, "null argument");
  if (axis < 0 || axis >= s->ndim) return fail(TPF_ERR_VALUE, "axis out of range");
  if (!s->rho) return fail(TPF_ERR_RUNTIME, "FCCD PPE density not set");
  if (!s->face_mode) {  // coefficients formed inside the matvec: build them once here
    if (!s->cf && hipMalloc(&s->cf, sizeof(double) * (size_t)s->g.total * s->ndim) != hipSuccess)
      return fail(TPF_ERR_RUNTIME, "FCCD PPE coefficient allocation failed");
    FCCD_LAUNCH(k_fccd_face_coef, s->g, 0, s->rho, 0.0, (const double*)nullptr, 0.5, s->cf);
  }
  TPF_HIP(hipMemcpyAsync(out, s->cf + (size_t)axis * s->g.total, sizeof(double) * s->g.total,
                         hipMemcpyDeviceToDevice, s->stream));
  return TPF_OK;
}

extern "C" tpf_status tpf_fccd_ppe_interface_jump(tpf_fccd_ppe* s, const double* p, double* out) {
  if (!s || !p || !out) return fail(TPF_ERR_VALUE, "null argument");
  if (jump_active(s)) {
    FCCD_LAUNCH(k_fccd_add_jump, s->g.total, (const double*)s->jfield, s->psi, p, out);
  } else if (p != out) {
    TPF_HIP(hipMemcpyAsync(out, p, sizeof(double) * s->g.total, hipMemcpyDeviceToDevice,
                           s->stream));
  }
  return TPF_OK;
}

extern "C" tpf_status tpf_fccd_ppe_gauge_info(tpf_fccd_ppe* s, int* mean_gauge,
                                              double* threshold) {
  if (!s || !mean_gauge || !threshold) return fail(TPF_ERR_VALUE, "null argument");
  *mean_gauge = s->mean_gauge;
  *threshold = s->mean_gauge ? s->rho_thr : std::nan("");
  return TPF_OK;
}
