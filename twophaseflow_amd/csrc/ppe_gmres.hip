// GMRES form of the FVM PPE with the shifted line preconditioner.
//
// Reference: ppe/fvm_matrixfree.py:150-260 (GMRES branch), :302-330
// (apply_line_preconditioner / apply_jacobi_preconditioner),
// fvm_matrixfree_helpers.py:47-101 (line coefficients, shift
// 2/(c_tau rho h_min^2), Jacobi diagonal), linalg_backend.py:77-115 and
// 270-340 (exact Thomas on the CPU backend, PCR capped at `max_stages` on the
// GPU backend) and gmres_helpers.py:16-42 (SciPy gmres, atol = 0,
// rtol = tol).  The system is L p = rhs with rhs[pin] = 0 and the pin row
// replaced by identity (fvm_matrixfree.py:163-166, 297-299).
//
// GMRES(m) follows SciPy 1.15's left-preconditioned restart loop
// (scipy/sparse/linalg/_isolve/iterative.py, gmres): inner stop on the
// preconditioned residual estimate <= ptol (adapted per cycle), outer stop on
// the true ||rhs - L x|| <= tol ||rhs||, `maxiter` counts restart cycles.
// Orthogonalisation is classical Gram-Schmidt applied twice (CGS2: one fused
// multi-dot and one multi-axpy kernel per pass, chunks of 16 basis vectors
// spread over grid.y) instead of SciPy's modified Gram-Schmidt, so iteration
// histories agree to rounding, not bit for bit.  The Hessenberg column, the
// Givens rotations and the inner stop test of every Arnoldi step run in one
// single-thread kernel on the device (k_arnoldi), so the host launches the
// steps ahead and reads the stop flag once every kCheckSteps steps; steps
// issued past the stop leave the Givens state alone.  The host takes back
// the Hessenberg matrix once per restart cycle for the triangular solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "gmres.hpp"
#include "ppe_impl.hpp"
#include "reduce.hpp"

namespace tpf {
namespace {

constexpr int kChunk = 16;          // basis vectors per multi-dot block row
// Fused PCR: lines of up to kPcrFusedMaxN nodes, kPcrTileNodes nodes per
// workgroup tile (8 doubles each in LDS: 32 KB per coefficient set at 1024).
constexpr int kPcrFusedMaxN = 2048;
constexpr int kPcrTileNodes = 512;
constexpr int kDot = kChunk + 1;    // + <w, w>
constexpr int kCheckSteps = 8;      // Arnoldi steps launched between host reads of the stop flag
// Multi-dot workgroups per block row: at most 512 on small problems (fewer
// partials to finish: 65^3 GMRES step 0.226 -> 0.190 ms), one per 8192
// unknowns above that (with the cap, a 257^3 multi-dot row of 16 basis
// vectors streamed at 4.2 TB/s in 357 us).
inline unsigned dot_blocks(unsigned grid, int64_t n) {
  const int64_t want = std::max<int64_t>(512, n / 8192);
  return (unsigned)std::min<int64_t>(grid, want);
}

// Device side of one restart cycle's Hessenberg / Givens work.
struct GivensDev {
  double ptol;     // inner tolerance of the cycle (set by the host)
  double presid;   // |last rotated residual| = preconditioned residual estimate
  double inv_h1;   // scale of the newest basis vector (1 after a breakdown)
  int stop;        // the inner loop has ended
  int col_end;     // the column it ended on
  int breakdown;   // h_{col+1,col} <= eps h0: exact-solution indicator
  int inner;       // Arnoldi steps of the cycle
};

__device__ inline void coords3(const PpeGeom& g, int64_t idx, int (&c)[3]) {
  const uint32_t plane = (uint32_t)(g.n1 * g.n2);
  const uint32_t u = (uint32_t)idx;
  c[0] = (int)(u / plane);
  const uint32_t rem = u - (uint32_t)c[0] * plane;
  c[1] = (int)(rem / (uint32_t)g.n2);
  c[2] = (int)(rem - (uint32_t)c[1] * (uint32_t)g.n2);
}

__host__ __device__ inline int slot_n(const PpeGeom& g, int a) {
  return a == 0 ? g.n0 : (a == 1 ? g.n1 : g.n2);
}
__host__ __device__ inline int64_t slot_stride(const PpeGeom& g, int a) {
  return a == 0 ? (int64_t)g.n1 * g.n2 : (a == 1 ? (int64_t)g.n2 : 1);
}

// Face weight of node i along slot a: (1/h) / dv_i, dv = h/2 at the walls.
__device__ inline double face_w(const PpeGeom& g, int a, int i) {
  const int n = slot_n(g, a);
  const double dv = (i == 0 || i == n - 1) ? 0.5 * g.h[a] : g.h[a];
  return g.ih[a] / dv;
}

// Unweighted L p (fvm_matrixfree.py:281-300): per axis (1/dv_i) sum over the
// two faces of 2/(rho_i + rho_nbr)/h (p_nbr - p_i); pin row = identity.
__global__ __launch_bounds__(kBlock) void k_L(PpeGeom g, const double* __restrict__ p,
                                              const double* __restrict__ rho,
                                              double* __restrict__ out) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords3(g, idx, c);
    const double pc = p[idx], rc = rho[idx];
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int n = slot_n(g, a);
      if (n == 1) continue;
      const int64_t st = slot_stride(g, a);
      const double w = face_w(g, a, c[a]);
      if (c[a] > 0) acc += w * (2.0 / (rc + rho[idx - st])) * (p[idx - st] - pc);
      if (c[a] < n - 1) acc += w * (2.0 / (rc + rho[idx + st])) * (p[idx + st] - pc);
    }
    out[idx] = idx == g.pin ? pc : acc;
  }
}

// Jacobi of the GMRES form (fvm_matrixfree_helpers.py:92-100,
// fvm_matrixfree.py:325-330): z = r / sum_axes(main), diag[pin] = 1,
// |diag| <= 1e-30 -> 1, then z[pin] = 0.
__global__ __launch_bounds__(kBlock) void k_jacobi_L(PpeGeom g, const double* __restrict__ rho,
                                                     const double* __restrict__ r,
                                                     double* __restrict__ z) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords3(g, idx, c);
    const double rc = rho[idx];
    double diag = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int n = slot_n(g, a);
      if (n == 1) continue;
      const int64_t st = slot_stride(g, a);
      const double w = face_w(g, a, c[a]);
      double lo = 0.0, up = 0.0;
      if (c[a] > 0) lo = w * (2.0 / (rc + rho[idx - st]));
      if (c[a] < n - 1) up = w * (2.0 / (rc + rho[idx + st]));
      diag += -(lo + up);
    }
    if (idx == g.pin) diag = 1.0;
    if (!(fabs(diag) > 1e-30)) diag = 1.0;
    z[idx] = idx == g.pin ? 0.0 : r[idx] / diag;
  }
}

// Shifted line system of node i on one line along slot a
// (fvm_matrixfree_helpers.py:47-90): lo z_{i-1} + di z_i + up z_{i+1}, with
// lo = -lower, up = -upper, di = shift - main = shift + lower + upper.
__device__ inline void line_coef(const PpeGeom& g, const double* __restrict__ rho, int a,
                                 int64_t o, int64_t st, int n, int i, double c_tau,
                                 double h_min2, double& lo, double& di, double& up) {
  const double ri = rho[o];
  const double w = face_w(g, a, i);
  const double lower = i > 0 ? w * (2.0 / (ri + rho[o - st])) : 0.0;
  const double upper = i < n - 1 ? w * (2.0 / (ri + rho[o + st])) : 0.0;
  const double shift = 2.0 / (c_tau * ri * h_min2);
  lo = -lower;
  up = -upper;
  di = shift - (-(lower + upper));
}

struct LineArgs {
  int a;          // slot
  double c_tau;
  double h_min2;  // h_min^2
  double scale;   // 1 / ndim
  int first = 0;  // k_pcr_fused: out = 0 + term (the zero fill folded in)
  int last = 0;   // k_pcr_fused: out[pin] = 0 afterwards (folded in)
};

// Exact Thomas per line (linalg_backend.py:323-339), one thread per line;
// consecutive threads take neighbouring lines, so the sweeps are coalesced
// for the two strided slots.  cp / dp are per-node scratch.
__global__ __launch_bounds__(kBlock) void k_line_thomas(PpeGeom g, LineArgs la,
                                                        const double* __restrict__ rho,
                                                        const double* __restrict__ r,
                                                        double* __restrict__ out,
                                                        double* __restrict__ cp,
                                                        double* __restrict__ dp) {
  const int a = la.a;
  const int n = slot_n(g, a);
  const int64_t st = slot_stride(g, a);
  const int64_t nlines = g.total / n;
  for (int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x; l < nlines;
       l += (int64_t)gridDim.x * kBlock) {
    int64_t base;
    if (a == 0) base = l;
    else if (a == 1) base = (l / g.n2) * ((int64_t)g.n1 * g.n2) + (l % g.n2);
    else base = l * n;
    double cprev = 0.0, dprev = 0.0;
    for (int i = 0; i < n; ++i) {
      const int64_t o = base + i * st;
      double lo, di, up;
      line_coef(g, rho, a, o, st, n, i, la.c_tau, la.h_min2, lo, di, up);
      const double ri = o == g.pin ? 0.0 : r[o];  // r_work[pin] = 0
      const double den = i == 0 ? di : di - lo * cprev;
      const double cc = i < n - 1 ? up / den : 0.0;
      const double dd = (i == 0 ? ri : ri - lo * dprev) / den;
      cp[o] = cc;
      dp[o] = dd;
      cprev = cc;
      dprev = dd;
    }
    double x = 0.0;
    for (int i = n - 1; i >= 0; --i) {
      const int64_t o = base + i * st;
      x = i == n - 1 ? dp[o] : dp[o] - cp[o] * x;
      out[o] += la.scale * x;
    }
  }
}

// PCR (linalg_backend.py:77-115): coefficient arrays over the whole field,
// one launch per stage, neighbours at +-stride along slot a.
__global__ __launch_bounds__(kBlock) void k_pcr_init(PpeGeom g, LineArgs la,
                                                     const double* __restrict__ rho,
                                                     const double* __restrict__ r,
                                                     double* __restrict__ lo,
                                                     double* __restrict__ di,
                                                     double* __restrict__ up,
                                                     double* __restrict__ rw) {
  const int a = la.a;
  const int n = slot_n(g, a);
  const int64_t st = slot_stride(g, a);
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords3(g, idx, c);
    double l, d, u;
    line_coef(g, rho, a, idx, st, n, c[a], la.c_tau, la.h_min2, l, d, u);
    lo[idx] = l;
    di[idx] = d;
    up[idx] = u;
    rw[idx] = idx == g.pin ? 0.0 : r[idx];
  }
}

__global__ __launch_bounds__(kBlock) void k_pcr_stage(PpeGeom g, int a, int stride,
                                                      const double* __restrict__ lo,
                                                      const double* __restrict__ di,
                                                      const double* __restrict__ up,
                                                      const double* __restrict__ rw,
                                                      double* __restrict__ lo2,
                                                      double* __restrict__ di2,
                                                      double* __restrict__ up2,
                                                      double* __restrict__ rw2) {
  const int n = slot_n(g, a);
  const int64_t st = slot_stride(g, a);
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords3(g, idx, c);
    const int i = c[a];
    double alpha = 0.0, beta = 0.0;
    double uL = 0.0, rL = 0.0, lR = 0.0, rR = 0.0, lL = 0.0, uR = 0.0;
    if (i >= stride) {
      const int64_t o = idx - stride * st;
      const double dL = di[o];
      alpha = -lo[idx] / (fabs(dL) > 0.0 ? dL : 1.0);
      lL = lo[o];
      uL = up[o];
      rL = rw[o];
    }
    if (i < n - stride) {
      const int64_t o = idx + stride * st;
      const double dR = di[o];
      beta = -up[idx] / (fabs(dR) > 0.0 ? dR : 1.0);
      lR = lo[o];
      uR = up[o];
      rR = rw[o];
    }
    lo2[idx] = alpha * lL;
    di2[idx] = di[idx] + alpha * uL + beta * lR;
    up2[idx] = beta * uR;
    rw2[idx] = rw[idx] + alpha * rL + beta * rR;
  }
}

// All PCR stages of one slot in one launch: a workgroup holds `L` whole
// lines (L * n nodes) of lo / di / up / rw in LDS, twice (ping-pong), forms
// the shifted line coefficients, runs the stages there with a barrier between
// them and accumulates scale * rw / di.  The arithmetic per element is that of
// k_pcr_init / k_pcr_stage / k_pcr_finish.  Threads run over lines on the
// strided slots and over nodes on the contiguous slot (coalesced loads).
__global__ __launch_bounds__(kBlock) void k_pcr_fused(PpeGeom g, LineArgs la, int L, int stages,
                                                      const double* __restrict__ rho,
                                                      const double* __restrict__ r,
                                                      double* __restrict__ out) {
  extern __shared__ double sh[];
  const int a = la.a;
  const int n = slot_n(g, a);
  const int64_t st = slot_stride(g, a);
  const int64_t nlines = g.total / n;
  const int64_t l0 = (int64_t)blockIdx.x * L;
  const int nl = (int)(nlines - l0 < L ? nlines - l0 : L);
  const int cnt = nl * n;
  const int tile = L * n;
  double* A[4] = {sh, sh + tile, sh + 2 * tile, sh + 3 * tile};
  double* B[4] = {sh + 4 * tile, sh + 5 * tile, sh + 6 * tile, sh + 7 * tile};
  auto elem = [&](int e, int& li, int& i) {
    if (a == 2) { li = e / n; i = e - li * n; }
    else { i = e / nl; li = e - i * nl; }
  };
  auto base_of = [&](int64_t l) -> int64_t {
    if (a == 0) return l;
    if (a == 1) return (l / g.n2) * ((int64_t)g.n1 * g.n2) + (l % g.n2);
    return l * n;
  };
  for (int e = threadIdx.x; e < cnt; e += kBlock) {
    int li, i;
    elem(e, li, i);
    const int64_t o = base_of(l0 + li) + (int64_t)i * st;
    double lo, di, up;
    line_coef(g, rho, a, o, st, n, i, la.c_tau, la.h_min2, lo, di, up);
    const int s = li * n + i;
    A[0][s] = lo;
    A[1][s] = di;
    A[2][s] = up;
    A[3][s] = o == g.pin ? 0.0 : r[o];
  }
  __syncthreads();
  for (int k = 0, stride = 1; k < stages; ++k, stride *= 2) {
    for (int e = threadIdx.x; e < cnt; e += kBlock) {
      int li, i;
      elem(e, li, i);
      const int s = li * n + i;
      double alpha = 0.0, beta = 0.0;
      double uL = 0.0, rL = 0.0, lR = 0.0, rR = 0.0, lL = 0.0, uR = 0.0;
      if (i >= stride) {
        const int q = s - stride;
        const double dL = A[1][q];
        alpha = -A[0][s] / (fabs(dL) > 0.0 ? dL : 1.0);
        lL = A[0][q];
        uL = A[2][q];
        rL = A[3][q];
      }
      if (i < n - stride) {
        const int q = s + stride;
        const double dR = A[1][q];
        beta = -A[2][s] / (fabs(dR) > 0.0 ? dR : 1.0);
        lR = A[0][q];
        uR = A[2][q];
        rR = A[3][q];
      }
      B[0][s] = alpha * lL;
      B[1][s] = A[1][s] + alpha * uL + beta * lR;
      B[2][s] = beta * uR;
      B[3][s] = A[3][s] + alpha * rL + beta * rR;
    }
    __syncthreads();
    for (int q = 0; q < 4; ++q) {
      double* t = A[q];
      A[q] = B[q];
      B[q] = t;
    }
  }
  for (int e = threadIdx.x; e < cnt; e += kBlock) {
    int li, i;
    elem(e, li, i);
    const int s = li * n + i;
    const int64_t o = base_of(l0 + li) + (int64_t)i * st;
    const double v = la.scale * (A[3][s] / A[1][s]);
    if (la.last && o == g.pin) out[o] = 0.0;
    else out[o] = (la.first ? 0.0 : out[o]) + v;
  }
}

__global__ __launch_bounds__(kBlock) void k_pcr_finish(int64_t n, double scale,
                                                       const double* __restrict__ di,
                                                       const double* __restrict__ rw,
                                                       double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    out[i] += scale * (rw[i] / di[i]);
}

// ---- vector kernels ----------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_fill(int64_t n, double v, double* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    x[i] = v;
}
__global__ void k_set_pin(double* __restrict__ x, int64_t pin) {
  if (blockIdx.x == 0 && threadIdx.x == 0) x[pin] = 0.0;
}
// y = a x + b y
__global__ __launch_bounds__(kBlock) void k_axpby(int64_t n, double a,
                                                  const double* __restrict__ x, double b,
                                                  double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    y[i] = a * x[i] + b * y[i];
}
// y = x - y
__global__ __launch_bounds__(kBlock) void k_xmy(int64_t n, const double* __restrict__ x,
                                                double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    y[i] = x[i] - y[i];
}

// Block row blockIdx.y: partial sums of <V_k, w> for the basis vectors
// k = 16*blockIdx.y + j (j < 16, k < kcount) and of <w, w> (slot 16).
__global__ __launch_bounds__(kBlock) void k_multidot(int64_t n, const double* __restrict__ V,
                                                     int kcount, const double* __restrict__ w,
                                                     double* partials) {
  __shared__ double sh[kDot * (kBlock / 64)];
  const int k0 = blockIdx.y * kChunk;
  const int kc = min(kChunk, kcount - k0);
  double v[kDot];
#pragma unroll
  for (int k = 0; k < kDot; ++k) v[k] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const double wi = w[i];
#pragma unroll
    for (int k = 0; k < kChunk; ++k)
      if (k < kc) v[k] += V[(int64_t)(k0 + k) * n + i] * wi;
    v[kChunk] += wi * wi;
  }
  block_partials<kDot>(v, partials + (size_t)blockIdx.y * gridDim.x * kDot, sh);
}
__global__ __launch_bounds__(kBlock) void k_multidot_fin(const double* partials, int nblocks,
                                                         double* out) {
  __shared__ double sh[kDot * (kBlock / 64)];
  double v[kDot];
  finish_sum<kDot>(v, partials + (size_t)blockIdx.y * nblocks * kDot, nblocks, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < kDot; ++k) out[blockIdx.y * kDot + k] = v[k];
}
// w += sign * sum_{k < kc} d_k V_k with d_k = dots[(k / 16) * 17 + k % 16],
// the layout k_multidot_fin writes (the Gram-Schmidt coefficients stay on the
// device).
__global__ __launch_bounds__(kBlock) void k_multiaxpy_dots(int64_t n, const double* __restrict__ V,
                                                           int kc,
                                                           const double* __restrict__ dots,
                                                           double sign, double* __restrict__ w) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    double acc = 0.0;
    for (int k = 0; k < kc; ++k) acc += dots[(k / kChunk) * kDot + k % kChunk] * V[(int64_t)k * n + i];
    w[i] += sign * acc;
  }
}

// w += sign * sum_{k < kc} coef[k] V_k
__global__ __launch_bounds__(kBlock) void k_multiaxpy(int64_t n, const double* __restrict__ V,
                                                      int kc, const double* __restrict__ coef,
                                                      double sign, double* __restrict__ w) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    double acc = 0.0;
    for (int k = 0; k < kc; ++k) acc += coef[k] * V[(int64_t)k * n + i];
    w[i] += sign * acc;
  }
}

// Plane rotation with c f + s g = r, -s f + c g = 0 (LAPACK 3.10 dlartg
// conventions: c >= 0, r carries the sign of f).
__host__ __device__ inline void lartg(double f, double g, double& cs, double& sn, double& r) {
  if (g == 0.0) {
    cs = 1.0;
    sn = 0.0;
    r = f;
  } else if (f == 0.0) {
    cs = 0.0;
    sn = g > 0.0 ? 1.0 : -1.0;
    r = fabs(g);
  } else {
    const double d = hypot(f, g);
    cs = fabs(f) / d;
    r = f > 0.0 ? d : -d;
    sn = g / r;
  }
}

// Arnoldi step `col` after CGS2: column col of the Hessenberg matrix from the
// three dot-product sets (two Gram-Schmidt passes, then <w,w>), the
// breakdown test, the earlier rotations, the new rotation, the rotated
// right-hand side S and the inner stop test (presid <= ptol or breakdown).
// The column and the rotations are staged in LDS by the workgroup; one
// thread runs the rotation chain.  Returns at once when an earlier step
// stopped the cycle.
constexpr int kArnoldiMax = 1024;  // restart cap of the LDS staging
__global__ __launch_bounds__(256) void k_arnoldi(int col, int ld, const double* __restrict__ dots,
                                                 int set, double* __restrict__ h,
                                                 double* __restrict__ S,
                                                 double* __restrict__ gc,
                                                 double* __restrict__ gs, GivensDev* st) {
  __shared__ double hc[kArnoldiMax + 1], cs_[kArnoldiMax], sn_[kArnoldiMax];
  __shared__ int stop;
  if (threadIdx.x == 0) stop = st->stop;
  __syncthreads();
  if (stop) return;
  const double* d0 = dots;
  const double* d1 = dots + set;
  for (int k = threadIdx.x; k <= col; k += blockDim.x) {
    const int at = (k / kChunk) * kDot + k % kChunk;
    hc[k] = d0[at] + d1[at];
    if (k < col) {
      cs_[k] = gc[k];
      sn_[k] = gs[k];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double eps = 2.220446049250313e-16;
    const double h0 = sqrt(d0[kChunk]);
    const double h1 = sqrt(dots[2 * set + kChunk]);
    int brk = 0;
    hc[col + 1] = h1;
    if (h1 <= eps * h0) {
      hc[col + 1] = 0.0;
      brk = 1;
      st->inv_h1 = 1.0;
    } else {
      st->inv_h1 = 1.0 / h1;
    }
    double carry = hc[0];
    for (int k = 0; k < col; ++k) {
      const double n0 = carry, n1 = hc[k + 1];
      hc[k] = cs_[k] * n0 + sn_[k] * n1;
      carry = -sn_[k] * n0 + cs_[k] * n1;
    }
    double cs, sn, mag;
    lartg(carry, hc[col + 1], cs, sn, mag);
    gc[col] = cs;
    gs[col] = sn;
    hc[col] = mag;
    hc[col + 1] = 0.0;
    const double tmp = -sn * S[col];
    S[col] = cs * S[col];
    S[col + 1] = tmp;
    const double presid = fabs(tmp);
    st->presid = presid;
    st->inner += 1;
    if (presid <= st->ptol || brk) {
      st->stop = 1;
      st->col_end = col;
      st->breakdown = brk;
    }
  }
  __syncthreads();
  double* hcol = h + (size_t)col * ld;
  for (int k = threadIdx.x; k <= col + 1; k += blockDim.x) hcol[k] = hc[k];
}

// v *= inv_h1 of step `col` (skipped for steps issued after the stop).
__global__ __launch_bounds__(kBlock) void k_scale_basis(int64_t n, int col,
                                                        const GivensDev* __restrict__ st,
                                                        double* __restrict__ v) {
  if (st->stop && st->col_end < col) return;
  const double a = st->inv_h1;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    v[i] = 0.0 * v[i] + a * v[i];
}

// ---- host side -----------------------------------------------------------
struct GmresWork {
  int m = 0;             // restart the basis was sized for
  int chunks = 0;        // ceil((m + 1) / 16)
  int64_t n = 0;         // vector length
  int nscratch = 0;      // extra n-vectors in s8
  double* V = nullptr;   // (m + 1) x n basis, row-major
  double* w = nullptr;   // residual
  double* t = nullptr;   // A v
  double* b = nullptr;   // right-hand side as solved
  double* s8 = nullptr;  // nscratch n: Thomas factors or PCR ping-pong arrays
  double* coef = nullptr;
  double* part = nullptr;
  double* dots = nullptr;
  double* host_dn = nullptr;  // pinned: chunks * 17 results
  double* host_up = nullptr;  // pinned: m + 1 coefficients
  double* hdev = nullptr;     // Hessenberg m x (m + 1), S (m + 1), cosines, sines (m each)
  double* host_h = nullptr;   // pinned: Hessenberg and S of one cycle
  GivensDev* gdev = nullptr;
  GivensDev* ghost = nullptr;  // pinned
};

void free_work(GmresWork* gw) {
  double* bufs[] = {gw->V, gw->w, gw->t, gw->b, gw->s8, gw->coef, gw->part, gw->dots, gw->hdev};
  for (double* p : bufs)
    if (p) (void)hipFree(p);
  if (gw->gdev) (void)hipFree(gw->gdev);
  if (gw->host_dn) (void)hipHostFree(gw->host_dn);
  if (gw->host_up) (void)hipHostFree(gw->host_up);
  if (gw->host_h) (void)hipHostFree(gw->host_h);
  if (gw->ghost) (void)hipHostFree(gw->ghost);
  delete gw;
}

// The basis and scratch of one handle, grown on demand (kept between solves).
tpf_status ensure_work(void** slot, int64_t n, int m, int grid_blocks, int nscratch) {
  GmresWork* gw = static_cast<GmresWork*>(*slot);
  if (gw && gw->m >= m && gw->n == n && gw->nscratch >= nscratch) return TPF_OK;
  if (gw) free_work(gw);
  gw = new GmresWork();
  *slot = gw;
  gw->m = m;
  gw->n = n;
  gw->nscratch = nscratch;
  gw->chunks = (m + 1 + kChunk - 1) / kChunk;
  TPF_HIP(hipMalloc(&gw->V, sizeof(double) * n * (size_t)(m + 1)));
  TPF_HIP(hipMalloc(&gw->w, sizeof(double) * n));
  TPF_HIP(hipMalloc(&gw->t, sizeof(double) * n));
  TPF_HIP(hipMalloc(&gw->b, sizeof(double) * n));
  if (nscratch > 0) TPF_HIP(hipMalloc(&gw->s8, sizeof(double) * n * nscratch));
  TPF_HIP(hipMalloc(&gw->coef, sizeof(double) * (size_t)(m + 1)));
  TPF_HIP(hipMalloc(&gw->part, sizeof(double) * kDot * (size_t)grid_blocks * gw->chunks));
  // three result sets: the two Gram-Schmidt passes and the norm of one step
  TPF_HIP(hipMalloc(&gw->dots, sizeof(double) * 3 * kDot * (size_t)gw->chunks));
  TPF_HIP(hipHostMalloc(&gw->host_dn, sizeof(double) * 3 * kDot * (size_t)gw->chunks));
  TPF_HIP(hipHostMalloc(&gw->host_up, sizeof(double) * (size_t)(m + 1)));
  const size_t hsz = (size_t)m * (m + 1) + (size_t)(m + 1) + 2 * (size_t)m;
  TPF_HIP(hipMalloc(&gw->hdev, sizeof(double) * hsz));
  TPF_HIP(hipHostMalloc(&gw->host_h, sizeof(double) * ((size_t)m * (m + 1) + (size_t)(m + 1))));
  TPF_HIP(hipMalloc(&gw->gdev, sizeof(GivensDev)));
  TPF_HIP(hipHostMalloc(&gw->ghost, sizeof(GivensDev)));
  return TPF_OK;
}

// Stream, grid and work of the Krylov vector kernels.
struct Krylov {
  GmresWork* gw;
  hipStream_t st;
  dim3 grid;
  int64_t n;
};

struct Ctx {
  tpf_ppe* s;
  GmresWork* gw;
  hipStream_t st;
  dim3 grid;
  int64_t n;
  Krylov kr() const { return Krylov{gw, st, grid, n}; }
  int precond;
  int stages;
  double c_tau, h_min2;
};

#define TPF_LAUNCH(kern, grid, ...)                                               \
  do {                                                                            \
    hipLaunchKernelGGL(kern, grid, dim3(kBlock), 0, c.st, __VA_ARGS__);           \
    TPF_HIP(hipGetLastError());                                                   \
  } while (0)

tpf_status apply_L(Ctx& c, const double* x, double* y) {
  TPF_LAUNCH(k_L, c.grid, c.s->g, x, c.s->rho, y);
  return TPF_OK;
}

// z = M^-1 r (fvm_matrixfree.py:302-330); z and r must differ.
tpf_status psolve(Ctx& c, const double* r, double* z) {
  const PpeGeom& g = c.s->g;
  if (c.precond == 0) {
    TPF_HIP(hipMemcpyAsync(z, r, sizeof(double) * c.n, hipMemcpyDeviceToDevice, c.st));
    return TPF_OK;
  }
  if (c.precond == 1) {
    TPF_LAUNCH(k_jacobi_L, c.grid, g, c.s->rho, r, z);
    return TPF_OK;
  }
  const int ndim = (g.n0 > 1) + (g.n1 > 1) + (g.n2 > 1);
  // every axis on the fused PCR kernel: the zero fill and the pin reset ride
  // in its first and last launches
  bool fused_all = c.stages > 0;
  int a_first = -1, a_last = -1;
  for (int a = 0; a < 3; ++a) {
    if (slot_n(g, a) == 1) continue;
    fused_all = fused_all && slot_n(g, a) <= kPcrFusedMaxN;
    if (a_first < 0) a_first = a;
    a_last = a;
  }
  if (!fused_all) TPF_LAUNCH(k_fill, c.grid, c.n, 0.0, z);
  double* s8 = c.gw->s8;
  for (int a = 0; a < 3; ++a) {
    const int na = slot_n(g, a);
    if (na == 1) continue;
    LineArgs la{a, c.c_tau, c.h_min2, 1.0 / ndim};
    if (fused_all) {
      la.first = a == a_first;
      la.last = a == a_last;
    }
    if (c.stages <= 0) {
      const int64_t nlines = c.n / na;
      const int64_t lg = std::min<int64_t>((nlines + kBlock - 1) / kBlock, 1 << 20);
      TPF_LAUNCH(k_line_thomas, dim3((unsigned)lg), g, la, c.s->rho, r, z, s8, s8 + c.n);
      continue;
    }
    // linalg_backend.py:147-150: min(ceil(log2 n), max(1, max_stages)) stages
    int full = 0;
    while ((1 << full) < na) ++full;
    full = std::max(1, full);
    const int ns = std::min(full, std::max(1, c.stages));
    if (na <= kPcrFusedMaxN) {  // whole lines in LDS, one launch per slot
      const int L = std::max(1, kPcrTileNodes / na);
      const int64_t nlines = c.n / na;
      const size_t lds = sizeof(double) * 8 * (size_t)L * na;
      if (lds > 65536)
        TPF_HIP(hipFuncSetAttribute((const void*)k_pcr_fused,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      hipLaunchKernelGGL(k_pcr_fused, dim3((unsigned)((nlines + L - 1) / L)), dim3(kBlock), lds,
                         c.st, g, la, L, ns, c.s->rho, r, z);
      TPF_HIP(hipGetLastError());
      continue;
    }
    double* A[4] = {s8, s8 + c.n, s8 + 2 * c.n, s8 + 3 * c.n};
    double* B[4] = {s8 + 4 * c.n, s8 + 5 * c.n, s8 + 6 * c.n, s8 + 7 * c.n};
    TPF_LAUNCH(k_pcr_init, c.grid, g, la, c.s->rho, r, A[0], A[1], A[2], A[3]);
    for (int k = 0, stride = 1; k < ns; ++k, stride *= 2) {
      TPF_LAUNCH(k_pcr_stage, c.grid, g, a, stride, A[0], A[1], A[2], A[3], B[0], B[1], B[2],
                 B[3]);
      for (int q = 0; q < 4; ++q) std::swap(A[q], B[q]);
    }
    TPF_LAUNCH(k_pcr_finish, c.grid, c.n, la.scale, A[1], A[3], z);
  }
  if (!fused_all) TPF_LAUNCH(k_set_pin, dim3(1), z, g.pin);
  return TPF_OK;
}

// out[j*17 + k] = <V_{16j+k}, w> for 16j+k < kcount, out[j*17 + 16] = <w,w>.
// Synchronises the stream: every upload of host_up issued before this call
// has completed when it returns, so host_up may be rewritten afterwards.
tpf_status multidot(const Krylov& c, const double* V, int kcount, const double* w,
                    double* out) {
  const int chunks = std::max(1, (kcount + kChunk - 1) / kChunk);
  const unsigned gx = dot_blocks(c.grid.x, c.n);
  hipLaunchKernelGGL(k_multidot, dim3(gx, chunks), dim3(kBlock), 0, c.st, c.n, V, kcount,
                     w, c.gw->part);
  TPF_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_multidot_fin, dim3(1, chunks), dim3(kBlock), 0, c.st,
                     (const double*)c.gw->part, (int)gx, c.gw->dots);
  TPF_HIP(hipGetLastError());
  TPF_HIP(hipMemcpyAsync(c.gw->host_dn, c.gw->dots, sizeof(double) * kDot * chunks,
                         hipMemcpyDeviceToHost, c.st));
  TPF_HIP(hipStreamSynchronize(c.st));
  std::copy(c.gw->host_dn, c.gw->host_dn + kDot * chunks, out);
  return TPF_OK;
}

// Result set `set` (0..2) of the step's dot products: out[set][j*17 + k] as
// multidot; no copy, no synchronisation.
tpf_status multidot_dev(const Krylov& c, const double* V, int kcount, const double* w, int set) {
  const int chunks = std::max(1, (kcount + kChunk - 1) / kChunk);
  double* out = c.gw->dots + (size_t)set * kDot * c.gw->chunks;
  const unsigned gx = dot_blocks(c.grid.x, c.n);
  hipLaunchKernelGGL(k_multidot, dim3(gx, chunks), dim3(kBlock), 0, c.st, c.n, V, kcount,
                     w, c.gw->part);
  TPF_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_multidot_fin, dim3(1, chunks), dim3(kBlock), 0, c.st,
                     (const double*)c.gw->part, (int)gx, out);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

tpf_status norm2(const Krylov& c, const double* w, double& out) {
  double d[kDot];
  TPF_TRY(multidot(c, w, 0, w, d));
  out = std::sqrt(d[kChunk]);
  return TPF_OK;
}

// w += sign * sum_k coef[k] V_k.  Asynchronous: the stream synchronisation
// in the next multidot / norm2 (every caller issues one before the next
// multiaxpy) completes the upload before host_up is written again.
tpf_status multiaxpy(const Krylov& c, const double* V, int kc, const double* coef,
                     double sign, double* w) {
  std::copy(coef, coef + kc, c.gw->host_up);
  TPF_HIP(hipMemcpyAsync(c.gw->coef, c.gw->host_up, sizeof(double) * kc, hipMemcpyHostToDevice,
                         c.st));
  TPF_LAUNCH(k_multiaxpy, c.grid, c.n, V, kc, (const double*)c.gw->coef, sign, w);
  return TPF_OK;
}

tpf_status check_opts(tpf_ppe* s, const tpf_gmres_opts* o) {
  if (!s || !o) return fail(TPF_ERR_VALUE, "null argument");
  if (!s->rho) return fail(TPF_ERR_RUNTIME, "PPE density not set (call tpf_ppe_set_density)");
  if (o->precond < 0 || o->precond > 2)
    return fail(TPF_ERR_VALUE, "GMRES preconditioner must be 0 (none), 1 (Jacobi), 2 (line)");
  if (o->precond == 2 && !(o->c_tau > 0.0))
    return fail(TPF_ERR_VALUE, "line preconditioner needs c_tau > 0");
  return TPF_OK;
}

tpf_status make_ctx(tpf_ppe* s, const tpf_gmres_opts* o, int m, Ctx& c) {
  TPF_TRY(ensure_work(&s->gmres_work, s->g.total, m, s->grid, 8));
  hipEvent_t ev;
  TPF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  TPF_HIP(hipEventRecord(ev, s->user_stream));
  TPF_HIP(hipStreamWaitEvent(s->stream, ev, 0));
  TPF_HIP(hipEventDestroy(ev));
  const PpeGeom& g = s->g;
  double h_min = 0.0;
  for (int a = 0; a < 3; ++a)
    if (slot_n(g, a) > 1) h_min = h_min == 0.0 ? g.h[a] : std::min(h_min, g.h[a]);
  c = Ctx{s, static_cast<GmresWork*>(s->gmres_work), s->stream, dim3((unsigned)s->grid),
          g.total, o->precond, o->pcr_stages, o->c_tau, h_min * h_min};
  return TPF_OK;
}

}  // namespace

// Restarted left-preconditioned GMRES(m) on `b` (SciPy 1.15's loop, see the
// file header); x holds the initial guess on entry.  Fills r; r.converged = 0
// when the cycle budget ends above tol (the caller reports the error).
tpf_status gmres_core(void** work, hipStream_t st, unsigned grid, int64_t n, int m,
                      const GmresOp& apply, const GmresOp& psolve, const double* b, double* x,
                      double tol, int64_t maxiter, tpf_solve_report& r) {
  GmresWork* gw = static_cast<GmresWork*>(*work);
  if (m < 1 || m > kArnoldiMax) return fail(TPF_ERR_VALUE, "GMRES restart must be in [1, 1024]");
  if (!gw || gw->m < m || gw->n != n) return fail(TPF_ERR_RUNTIME, "GMRES work not sized");
  const Krylov c{gw, st, dim3(grid), n};
  double* V = gw->V;
  r = tpf_solve_report{};
  double bnrm2 = 0.0;
  TPF_TRY(norm2(c, b, bnrm2));
  if (bnrm2 == 0.0) {  // SciPy returns b itself (zero)
    TPF_LAUNCH(k_fill, c.grid, n, 0.0, x);
    r.converged = 1;
    return TPF_OK;
  }
  const double atol = tol * bnrm2;
  const double eps = 2.220446049250313e-16;
  // Mb_nrm2 = ||M b|| seeds the inner tolerance
  double Mb = 0.0;
  TPF_TRY(psolve(b, gw->w));
  TPF_TRY(norm2(c, gw->w, Mb));
  double ptol_max = 1.0;
  double ptol = Mb * std::min(ptol_max, atol / bnrm2);
  double presid = 0.0;
  // r = b - A x
  TPF_TRY(apply(x, gw->w));
  TPF_LAUNCH(k_xmy, c.grid, n, b, gw->w);
  double rnorm = 0.0;
  TPF_TRY(norm2(c, gw->w, rnorm));
  int64_t inner = 0, cycles = 0;
  const size_t ld = (size_t)m + 1;
  std::vector<double> y(ld);
  double* hd = gw->hdev;                  // Hessenberg, column-major with leading dim m + 1
  double* Sd = hd + (size_t)m * ld;       // rotated right-hand side
  double* gcd = Sd + ld;                  // rotation cosines
  double* gsd = gcd + m;                  // rotation sines
  GivensDev* gd = gw->gdev;
  const bool early = rnorm < atol;  // SciPy's first-cycle check
  while (!early && cycles < maxiter) {
    ++cycles;
    TPF_TRY(psolve(gw->w, V));  // v0 = M r / ||M r||
    double t0 = 0.0;
    TPF_TRY(norm2(c, V, t0));   // synchronises: host_up and ghost may be rewritten
    TPF_LAUNCH(k_axpby, c.grid, n, 0.0, (const double*)V, 1.0 / t0, V);
    // S = (t0, 0, ...), fresh Givens state with this cycle's ptol
    std::fill(gw->host_up, gw->host_up + ld, 0.0);
    gw->host_up[0] = t0;
    TPF_HIP(hipMemcpyAsync(Sd, gw->host_up, sizeof(double) * ld, hipMemcpyHostToDevice, c.st));
    *gw->ghost = GivensDev{ptol, 0.0, 1.0, 0, 0, 0, 0};
    TPF_HIP(hipMemcpyAsync(gd, gw->ghost, sizeof(GivensDev), hipMemcpyHostToDevice, c.st));
    const size_t set = (size_t)kDot * gw->chunks;
    int col = 0;
    bool stopped = false;
    for (col = 0; col < m; ++col) {
      double* vc = V + (int64_t)col * n;
      double* vn = V + (int64_t)(col + 1) * n;
      TPF_TRY(apply(vc, gw->t));
      TPF_TRY(psolve(gw->t, vn));  // w, built in place in v_{col+1}
      // CGS2 with the coefficients kept on the device: two dot / subtract
      // passes and the norm, then the Givens step on the device
      for (int pass = 0; pass < 2; ++pass) {
        TPF_TRY(multidot_dev(c, V, col + 1, vn, pass));
        TPF_LAUNCH(k_multiaxpy_dots, c.grid, n, (const double*)V, col + 1,
                   (const double*)(gw->dots + pass * set), -1.0, vn);
      }
      TPF_TRY(multidot_dev(c, vn, 0, vn, 2));
      hipLaunchKernelGGL(k_arnoldi, dim3(1), dim3(256), 0, c.st, col, (int)ld,
                         (const double*)gw->dots, (int)set, hd, Sd, gcd, gsd, gd);
      TPF_HIP(hipGetLastError());
      TPF_LAUNCH(k_scale_basis, c.grid, n, col, (const GivensDev*)gd, vn);
      if ((col + 1) % kCheckSteps == 0 || col == m - 1) {
        TPF_HIP(hipMemcpyAsync(gw->ghost, gd, sizeof(GivensDev), hipMemcpyDeviceToHost, c.st));
        TPF_HIP(hipStreamSynchronize(c.st));
        if (gw->ghost->stop) {
          stopped = true;
          break;
        }
      }
    }
    const GivensDev gstate = *gw->ghost;
    col = stopped ? gstate.col_end : m - 1;
    const bool breakdown = stopped && gstate.breakdown;
    presid = gstate.presid;
    inner += gstate.inner;
    // the cycle's Hessenberg columns and S back to the host
    const size_t hcount = (size_t)(col + 1) * ld;
    TPF_HIP(hipMemcpyAsync(gw->host_h, hd, sizeof(double) * hcount, hipMemcpyDeviceToHost, c.st));
    TPF_HIP(hipMemcpyAsync(gw->host_h + (size_t)m * ld, Sd, sizeof(double) * ld,
                           hipMemcpyDeviceToHost, c.st));
    TPF_HIP(hipStreamSynchronize(c.st));
    const double* h = gw->host_h;
    const double* S = gw->host_h + (size_t)m * ld;
    // upper-triangular solve h[:col+1, :col+1]^T y = S[:col+1] (SciPy's loop)
    for (int k = 0; k <= col; ++k) y[k] = S[k];
    if (h[(size_t)col * ld + col] == 0.0) y[col] = 0.0;
    for (int k = col; k > 0; --k) {
      if (y[k] != 0.0) {
        y[k] /= h[(size_t)k * ld + k];
        const double t = y[k];
        for (int j = 0; j < k; ++j) y[j] -= t * h[(size_t)k * ld + j];
      }
    }
    if (y[0] != 0.0) y[0] /= h[0];
    TPF_TRY(multiaxpy(c, V, col + 1, y.data(), 1.0, x));
    TPF_TRY(apply(x, gw->w));
    TPF_LAUNCH(k_xmy, c.grid, n, b, gw->w);
    TPF_TRY(norm2(c, gw->w, rnorm));
    if (rnorm <= atol || breakdown) break;
    if (presid <= ptol) ptol_max = std::max(eps, 0.25 * ptol_max);
    else ptol_max = std::min(1.0, 1.5 * ptol_max);
    ptol = presid * std::min(ptol_max, atol / rnorm);
  }
  r.iterations = inner;
  r.relres = rnorm / bnrm2;
  r.true_relres = rnorm / bnrm2;
  r.converged = (early || rnorm <= atol) ? 1 : 0;
  r.refinements = (int)cycles;
  return TPF_OK;
}

tpf_status gmres_work(void** slot, int64_t n, int m, int grid_blocks, int nscratch) {
  return ensure_work(slot, n, m, grid_blocks, nscratch);
}

void gmres_work_free(void** slot) {
  if (*slot) free_work(static_cast<GmresWork*>(*slot));
  *slot = nullptr;
}

double* gmres_rhs(void* work) { return static_cast<GmresWork*>(work)->b; }

double* gmres_scratch(void* work) { return static_cast<GmresWork*>(work)->s8; }

}  // namespace tpf

using namespace tpf;

void tpf_gmres_release(tpf_ppe* s) { gmres_work_free(&s->gmres_work); }

extern "C" tpf_status tpf_ppe_gmres_precond(tpf_ppe* s, const double* r, double* z,
                                            const tpf_gmres_opts* o) {
  TPF_TRY(check_opts(s, o));
  if (!r || !z || r == z) return fail(TPF_ERR_VALUE, "r and z must be distinct device arrays");
  Ctx c;
  TPF_TRY(make_ctx(s, o, 1, c));
  TPF_TRY(psolve(c, r, z));
  TPF_HIP(hipStreamSynchronize(c.st));
  return TPF_OK;
}

extern "C" tpf_status tpf_ppe_solve_gmres(tpf_ppe* s, const double* rhs, double* x,
                                          const tpf_gmres_opts* o, tpf_solve_report* rep) {
  TPF_TRY(check_opts(s, o));
  if (!rhs || !x) return fail(TPF_ERR_VALUE, "null argument");
  if (!(o->tol > 0.0) || o->maxiter < 1 || o->restart < 1)
    return fail(TPF_ERR_VALUE, "GMRES needs tol > 0, maxiter >= 1, restart >= 1");
  const int m = (int)std::min<int64_t>(o->restart, s->g.total);  // SciPy: min(restart, n)
  Ctx c;
  TPF_TRY(make_ctx(s, o, m, c));
  const int64_t n = c.n;
  const int64_t pin = s->g.pin;
  double* b = c.gw->b;
  TPF_HIP(hipMemcpyAsync(b, rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, c.st));
  TPF_LAUNCH(k_set_pin, dim3(1), b, pin);
  if (o->use_x0) TPF_LAUNCH(k_set_pin, dim3(1), x, pin);
  else TPF_LAUNCH(k_fill, c.grid, n, 0.0, x);
  tpf_solve_report r{};
  TPF_TRY(gmres_core(&s->gmres_work, c.st, c.grid.x, n, m,
                     [&](const double* in, double* out) { return apply_L(c, in, out); },
                     [&](const double* in, double* out) { return psolve(c, in, out); }, b, x,
                     o->tol, o->maxiter, r));
  TPF_LAUNCH(k_set_pin, dim3(1), x, pin);  // sol[pin] = 0
  TPF_HIP(hipStreamSynchronize(c.st));
  if (rep) *rep = r;
  if (!r.converged)
    return fail(TPF_ERR_RUNTIME, "PPE GMRES did not converge cleanly (info=" +
                                     std::to_string(o->maxiter) +
                                     ", ||rhs - L p||/||rhs|| = " + std::to_string(r.relres) + ")");
  return TPF_OK;
}
