// Variable-density FVM PPE for gfx950: matrix-free operator, Jacobi-PCG in
// the reference's CG system form, iteration chunks replayed as a hipGraph.
//
// Reference: ppe/fvm_matrixfree.py:46-363 and fvm_matrixfree_helpers.py:
// 23-103.  The CG system is (-W L) p = -W rhs (fvm_matrixfree.py:163-202)
// with W the tensor-product control volume (half cells on wall nodes) and
// the pin row/column replaced by identity.  Row i of -W L is
//     sum_a w_a(i)/h_a * sum_{nbr along a} 2/(rho_i+rho_nbr) (p_i - p_nbr),
// w_a(i) = prod_{b != a} dv_b(i_b): the face coefficient is the harmonic
// mean of 1/rho (fvm_matrixfree_helpers.py:61) and is recomputed from rho
// on every application, so a matvec moves 24 B/DOF (p, rho in; Ap out).
//
// PCG per iteration (three streaming kernels, 13 FP64 words per DOF):
//   K_p : p = Dinv r + beta p ;  x += alpha_prev p_old   (deferred x update)
//   K_q : q = A p ;  <p,q>          -> last block: alpha
//   K_r : r -= alpha q ; <r,Dinv r>, <r,r>  -> last block: beta, stop test
// Block partial sums are combined by the last block to finish, in block
// order, so every reduction is deterministic.  Kernels return immediately
// once the device stop flag is set, which lets a captured chunk of
// iterations run past convergence at no cost.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "ppe_impl.hpp"
#include "reduce.hpp"

namespace tpf {
namespace {

__device__ inline void coords(const PpeGeom& g, int64_t idx, int& i0, int& i1, int& i2) {
  const uint32_t plane = (uint32_t)(g.n1 * g.n2);
  const uint32_t u = (uint32_t)idx;
  i0 = (int)(u / plane);
  const uint32_t rem = u - (uint32_t)i0 * plane;
  i1 = (int)(rem / (uint32_t)g.n2);
  i2 = (int)(rem - (uint32_t)i1 * (uint32_t)g.n2);
}

__device__ inline double dv(const PpeGeom& g, int a, int i) {
  const int n = a == 0 ? g.n0 : (a == 1 ? g.n1 : g.n2);
  if (n == 1) return 1.0;
  return (i == 0 || i == n - 1) ? 0.5 * g.h[a] : g.h[a];
}

// 1/x from the hardware reciprocal refined by Newton steps (x > 0 here:
// sums of two densities); avoids the long IEEE division sequence.
__device__ inline double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}
__device__ inline float recip(float x) {
  float r = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}

// (A p)_i of the CG system (pin handled by the caller) and its diagonal,
// evaluated in T with rho of the same precision.
template <typename T>
__device__ inline void stencil(const PpeGeom& g, const T* __restrict__ p,
                               const T* __restrict__ rho, int64_t idx, int i0, int i1, int i2,
                               T& ap, T& diag) {
  const T r0 = rho[idx];
  const T pc = p[idx];
  const T dv0 = (T)dv(g, 0, i0), dv1 = (T)dv(g, 1, i1), dv2 = (T)dv(g, 2, i2);
  const T two = (T)2;
  T acc = 0, dg = 0;
  if (g.n2 > 1) {  // slot 2 (contiguous)
    const T w = dv0 * dv1 * (T)g.ih[2];
    if (i2 > 0) {
      const T c = w * two * recip(r0 + rho[idx - 1]);
      acc += c * (pc - p[idx - 1]);
      dg += c;
    }
    if (i2 < g.n2 - 1) {
      const T c = w * two * recip(r0 + rho[idx + 1]);
      acc += c * (pc - p[idx + 1]);
      dg += c;
    }
  }
  if (g.n1 > 1) {
    const int64_t s = g.n2;
    const T w = dv0 * dv2 * (T)g.ih[1];
    if (i1 > 0) {
      const T c = w * two * recip(r0 + rho[idx - s]);
      acc += c * (pc - p[idx - s]);
      dg += c;
    }
    if (i1 < g.n1 - 1) {
      const T c = w * two * recip(r0 + rho[idx + s]);
      acc += c * (pc - p[idx + s]);
      dg += c;
    }
  }
  if (g.n0 > 1) {
    const int64_t s = (int64_t)g.n1 * g.n2;
    const T w = dv1 * dv2 * (T)g.ih[0];
    if (i0 > 0) {
      const T c = w * two * recip(r0 + rho[idx - s]);
      acc += c * (pc - p[idx - s]);
      dg += c;
    }
    if (i0 < g.n0 - 1) {
      const T c = w * two * recip(r0 + rho[idx + s]);
      acc += c * (pc - p[idx + s]);
      dg += c;
    }
  }
  ap = acc;
  diag = dg;
}

// ---- kernels ---------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_diag(PpeGeom g, const double* __restrict__ rho,
                                                 double* __restrict__ dinv) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int i0, i1, i2;
    coords(g, idx, i0, i1, i2);
    double ap, dg;
    stencil<double>(g, rho, rho, idx, i0, i1, i2, ap, dg);
    // fvm_matrixfree_helpers.py:100-103: diag[pin] = 1, guarded reciprocal
    if (idx == g.pin || !(fabs(dg) > 1e-30)) dg = 1.0;
    dinv[idx] = 1.0 / dg;
  }
}

// Reference apply (unweighted, fvm_matrixfree.py:281-300): out = L p, out[pin] = p[pin].
__global__ __launch_bounds__(kBlock) void k_apply_L(PpeGeom g, const double* __restrict__ p,
                                                    const double* __restrict__ rho,
                                                    double* __restrict__ out) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int i0, i1, i2;
    coords(g, idx, i0, i1, i2);
    double ap, dg;
    stencil<double>(g, p, rho, idx, i0, i1, i2, ap, dg);
    const double W = dv(g, 0, i0) * dv(g, 1, i1) * dv(g, 2, i2);
    out[idx] = idx == g.pin ? p[idx] : -ap / W;
  }
}

// b = -W rhs (b[pin] = 0); r = b - A x; scalars: rr, rz, bb.
__global__ __launch_bounds__(kBlock) void k_init(PpeGeom g, const double* __restrict__ rhs,
                                                 const double* __restrict__ rho,
                                                 const double* __restrict__ dinv,
                                                 double* __restrict__ x, double* __restrict__ b,
                                                 double* __restrict__ r, double* __restrict__ p,
                                                 PcgState* st, double* partials, int zero_x) {
  __shared__ double sh[3 * kBlock / 64];
  double v[3] = {0.0, 0.0, 0.0};
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int i0, i1, i2;
    coords(g, idx, i0, i1, i2);
    const double W = dv(g, 0, i0) * dv(g, 1, i1) * dv(g, 2, i2);
    const double bi = idx == g.pin ? 0.0 : -W * rhs[idx];
    double ap = 0.0, dg;
    if (zero_x) {
      x[idx] = 0.0;
    } else {  // the caller zeroed x[pin] beforehand
      stencil<double>(g, x, rho, idx, i0, i1, i2, ap, dg);
      if (idx == g.pin) ap = x[idx];
    }
    const double ri = bi - ap;
    b[idx] = bi;
    r[idx] = ri;
    p[idx] = 0.0;
    v[0] += ri * ri;
    v[1] += ri * dinv[idx] * ri;
    v[2] += bi * bi;
  }
  block_partials<3>(v, partials, sh);
}

__global__ __launch_bounds__(kBlock) void k_fin_init(PcgState* st, const double* partials,
                                                     int nblocks) {
  __shared__ double sh[3 * kBlock / 64];
  double v[3];
  finish_sum<3>(v, partials, nblocks, sh);
  if (threadIdx.x == 0) {
    st->rr = v[0];
    st->rz = v[1];
    st->bb = v[2];
    st->beta = 0.0;
    st->alpha = 0.0;
    st->iter = 0;
    st->done = (v[0] <= st->tol2 * v[2]) ? 1 : 0;
  }
}

// p = Dinv r + beta p ; x += alpha p_old
template <typename T>
__global__ __launch_bounds__(kBlock) void k_update_p(PpeGeom g, const T* __restrict__ r,
                                                     const T* __restrict__ dinv,
                                                     T* __restrict__ p, T* __restrict__ x,
                                                     const PcgState* st) {
  if (st->done) return;
  const T beta = (T)st->beta, alpha = (T)st->alpha;
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    const T po = p[idx];
    x[idx] += alpha * po;
    p[idx] = dinv[idx] * r[idx] + beta * po;
  }
}

// q = A p ; <p, q> -> alpha = rz / <p,q>   (reductions in FP64)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_matvec(PpeGeom g, const T* __restrict__ p,
                                                   const T* __restrict__ rho,
                                                   T* __restrict__ q, PcgState* st,
                                                   double* partials) {
  if (st->done) return;
  __shared__ double sh[kBlock / 64];
  double v[1] = {0.0};
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int i0, i1, i2;
    coords(g, idx, i0, i1, i2);
    T ap, dg;
    stencil<T>(g, p, rho, idx, i0, i1, i2, ap, dg);
    if (idx == g.pin) ap = p[idx];
    q[idx] = ap;
    v[0] += (double)p[idx] * (double)ap;
  }
  block_partials<1>(v, partials, sh);
}

// alpha = rz / <p, q>
__global__ __launch_bounds__(kBlock) void k_fin_alpha(PcgState* st, const double* partials,
                                                      int nblocks) {
  if (st->done) return;
  __shared__ double sh[kBlock / 64];
  double v[1];
  finish_sum<1>(v, partials, nblocks, sh);
  if (threadIdx.x == 0) {
    st->pq = v[0];
    st->alpha = st->rz / v[0];
  }
}

// r -= alpha q ; <r, Dinv r>, <r, r> -> beta, stop test  (reductions in FP64)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_update_r(PpeGeom g, const T* __restrict__ q,
                                                     const T* __restrict__ dinv,
                                                     T* __restrict__ r, PcgState* st,
                                                     double* partials) {
  if (st->done) return;
  __shared__ double sh[2 * kBlock / 64];
  const T alpha = (T)st->alpha;
  double v[2] = {0.0, 0.0};
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    const T ri = r[idx] - alpha * q[idx];
    r[idx] = ri;
    const double rd = (double)ri;
    v[0] += rd * (double)dinv[idx] * rd;
    v[1] += rd * rd;
  }
  block_partials<2>(v, partials, sh);
}

// beta = rz_new / rz, stop test on the recursively updated ||r||
__global__ __launch_bounds__(kBlock) void k_fin_beta(PcgState* st, const double* partials,
                                                     int nblocks) {
  if (st->done) return;
  __shared__ double sh[2 * kBlock / 64];
  double v[2];
  finish_sum<2>(v, partials, nblocks, sh);
  if (threadIdx.x == 0) {
    st->beta = v[0] / st->rz;
    st->rz = v[0];
    st->rr = v[1];
    st->iter += 1;
    if (v[1] <= st->tol2 * st->bb || st->iter >= st->maxiter) st->done = 1;
  }
}

// x += alpha p (the update deferred past the last iteration)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_finish_x(PpeGeom g, const T* __restrict__ p,
                                                     T* __restrict__ x, const PcgState* st) {
  const T alpha = (T)st->alpha;
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    x[idx] += alpha * p[idx];
    if (idx == g.pin) x[idx] = (T)0;
  }
}

// Mixed mode: FP64 residual of the CG system r = b - A x, handed to the FP32
// inner solve (r32 = r, x32 = p32 = 0); <r,r>, <r32,Dinv r32>, <b,b> in FP64.
__global__ __launch_bounds__(kBlock) void k_resid_mixed(
    PpeGeom g, const double* __restrict__ rhs, const double* __restrict__ rho,
    const double* __restrict__ x, const float* __restrict__ dinv32, float* __restrict__ r32,
    float* __restrict__ x32, float* __restrict__ p32, PcgState* st, double* partials) {
  __shared__ double sh[3 * kBlock / 64];
  double v[3] = {0.0, 0.0, 0.0};
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int i0, i1, i2;
    coords(g, idx, i0, i1, i2);
    const double W = dv(g, 0, i0) * dv(g, 1, i1) * dv(g, 2, i2);
    const double bi = idx == g.pin ? 0.0 : -W * rhs[idx];
    double ap, dg;
    stencil<double>(g, x, rho, idx, i0, i1, i2, ap, dg);
    if (idx == g.pin) ap = x[idx];
    const double ri = bi - ap;
    const float rf = (float)ri;
    r32[idx] = rf;
    x32[idx] = 0.0f;
    p32[idx] = 0.0f;
    v[0] += ri * ri;
    v[1] += (double)rf * (double)dinv32[idx] * (double)rf;
    v[2] += bi * bi;
  }
  block_partials<3>(v, partials, sh);
}

__global__ __launch_bounds__(kBlock) void k_fin_mixed(PcgState* st, const double* partials,
                                                      int nblocks) {
  __shared__ double sh[3 * kBlock / 64];
  double v[3];
  finish_sum<3>(v, partials, nblocks, sh);
  if (threadIdx.x == 0) {
    st->true_rr = v[0];
    st->true_bb = v[2];
    st->rz = v[1];
  }
}

// out = A p on the CG system (pin row identity), in T
template <typename T>
__global__ __launch_bounds__(kBlock) void k_system_apply(PpeGeom g, const T* __restrict__ p,
                                                         const T* __restrict__ rho,
                                                         T* __restrict__ out) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int i0, i1, i2;
    coords(g, idx, i0, i1, i2);
    T ap, dg;
    stencil<T>(g, p, rho, idx, i0, i1, i2, ap, dg);
    out[idx] = idx == g.pin ? p[idx] : ap;
  }
}

// b = -W rhs, b[pin] = 0 (fvm_matrixfree.py:163-177)
__global__ __launch_bounds__(kBlock) void k_system_rhs(PpeGeom g, const double* __restrict__ rhs,
                                                       double* __restrict__ b) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int i0, i1, i2;
    coords(g, idx, i0, i1, i2);
    const double W = dv(g, 0, i0) * dv(g, 1, i1) * dv(g, 2, i2);
    b[idx] = idx == g.pin ? 0.0 : -W * rhs[idx];
  }
}

__global__ void k_set_one(double* __restrict__ x, int64_t idx, double v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) x[idx] = v;
}

__global__ __launch_bounds__(kBlock) void k_fill_zero(int64_t n, double* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    x[i] = 0.0;
}

// x64 += x32
__global__ __launch_bounds__(kBlock) void k_accumulate(PpeGeom g, const float* __restrict__ x32,
                                                       double* __restrict__ x) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock)
    x[idx] += (double)x32[idx];
}

// FP32 copies of rho and the Jacobi diagonal for the inner solve.
__global__ __launch_bounds__(kBlock) void k_to_f32(int64_t n, const double* __restrict__ a,
                                                   float* __restrict__ af,
                                                   const double* __restrict__ b2,
                                                   float* __restrict__ bf) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    af[i] = (float)a[i];
    bf[i] = (float)b2[i];
  }
}

// FP64 residual of the CG system: returns <r,r> and <b,b>.
__global__ __launch_bounds__(kBlock) void k_true_res(PpeGeom g, const double* __restrict__ rhs,
                                                     const double* __restrict__ rho,
                                                     const double* __restrict__ x,
                                                     PcgState* st, double* partials) {
  __shared__ double sh[2 * kBlock / 64];
  double v[2] = {0.0, 0.0};
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int i0, i1, i2;
    coords(g, idx, i0, i1, i2);
    const double W = dv(g, 0, i0) * dv(g, 1, i1) * dv(g, 2, i2);
    const double bi = idx == g.pin ? 0.0 : -W * rhs[idx];
    double ap, dg;
    stencil<double>(g, x, rho, idx, i0, i1, i2, ap, dg);
    if (idx == g.pin) ap = x[idx];
    const double ri = bi - ap;
    v[0] += ri * ri;
    v[1] += bi * bi;
  }
  block_partials<2>(v, partials, sh);
}

__global__ __launch_bounds__(kBlock) void k_fin_true(PcgState* st, const double* partials,
                                                     int nblocks) {
  __shared__ double sh[2 * kBlock / 64];
  double v[2];
  finish_sum<2>(v, partials, nblocks, sh);
  if (threadIdx.x == 0) {
    st->true_rr = v[0];
    st->true_bb = v[1];
  }
}

// ---- multigrid-preconditioned flexible CG (FCG(1), Notay) on the CG system
// Device scalars of one solve; every kernel below is followed by a
// single-block finish kernel (deterministic reductions).
struct FcgState {
  double rr, bb, tol2;
  double zq;      // <z_k, q_{k-1}>
  double pq;      // <p_k, q_k>
  double pq_old;  // <p_{k-1}, q_{k-1}>
  double pr;      // <p_k, r_k>
  double alpha, beta;
  int64_t iter;
};

// b = -W rhs (b[pin] = 0); r = b - A x (x = 0 when zero_x); r32 = r.
__global__ __launch_bounds__(kBlock) void k_fcg_init(PpeGeom g, const double* __restrict__ rhs,
                                                     const double* __restrict__ rho,
                                                     double* __restrict__ x,
                                                     double* __restrict__ r,
                                                     float* __restrict__ r32, int zero_x,
                                                     double* partials) {
  __shared__ double sh[2 * kBlock / 64];
  double v[2] = {0.0, 0.0};
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int i0, i1, i2;
    coords(g, idx, i0, i1, i2);
    const double W = dv(g, 0, i0) * dv(g, 1, i1) * dv(g, 2, i2);
    const double bi = idx == g.pin ? 0.0 : -W * rhs[idx];
    double ap = 0.0, dg;
    if (!zero_x) {  // x[pin] = 0 beforehand
      stencil<double>(g, x, rho, idx, i0, i1, i2, ap, dg);
      if (idx == g.pin) ap = x[idx];
    }
    const double ri = bi - ap;
    r[idx] = ri;
    r32[idx] = (float)ri;
    v[0] += ri * ri;
    v[1] += bi * bi;
  }
  block_partials<2>(v, partials, sh);
}

__global__ __launch_bounds__(kBlock) void k_fcg_fin_init(FcgState* st, const double* partials,
                                                         int nblocks) {
  __shared__ double sh[2 * kBlock / 64];
  double v[2];
  finish_sum<2>(v, partials, nblocks, sh);
  if (threadIdx.x == 0) {
    st->rr = v[0];
    st->bb = v[1];
    st->iter = 0;
    st->beta = 0.0;
  }
}

// <z, q_old>
__global__ __launch_bounds__(kBlock) void k_fcg_zq(int64_t n, const float* __restrict__ z,
                                                   const double* __restrict__ q,
                                                   double* partials) {
  __shared__ double sh[kBlock / 64];
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    v[0] += (double)z[i] * q[i];
  block_partials<1>(v, partials, sh);
}

__global__ __launch_bounds__(kBlock) void k_fcg_fin_beta(FcgState* st, const double* partials,
                                                         int nblocks) {
  __shared__ double sh[kBlock / 64];
  double v[1];
  finish_sum<1>(v, partials, nblocks, sh);
  if (threadIdx.x == 0) {
    st->zq = v[0];
    st->beta = st->iter > 0 ? v[0] / st->pq_old : 0.0;
  }
}

// p = z - beta p ; q = A p is formed by the next kernel
__global__ __launch_bounds__(kBlock) void k_fcg_p(int64_t n, const float* __restrict__ z,
                                                  double* __restrict__ p, const FcgState* st) {
  const double beta = st->beta;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    p[i] = (double)z[i] - (beta != 0.0 ? beta * p[i] : 0.0);
}

// q = A p (pin row identity); partials <p,q>, <p,r>
__global__ __launch_bounds__(kBlock) void k_fcg_matvec(PpeGeom g, const double* __restrict__ p,
                                                       const double* __restrict__ rho,
                                                       const double* __restrict__ r,
                                                       double* __restrict__ q,
                                                       double* partials) {
  __shared__ double sh[2 * kBlock / 64];
  double v[2] = {0.0, 0.0};
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int i0, i1, i2;
    coords(g, idx, i0, i1, i2);
    double ap, dg;
    stencil<double>(g, p, rho, idx, i0, i1, i2, ap, dg);
    const double pi = p[idx];
    if (idx == g.pin) ap = pi;
    q[idx] = ap;
    v[0] += pi * ap;
    v[1] += pi * r[idx];
  }
  block_partials<2>(v, partials, sh);
}

__global__ __launch_bounds__(kBlock) void k_fcg_fin_alpha(FcgState* st, const double* partials,
                                                          int nblocks) {
  __shared__ double sh[2 * kBlock / 64];
  double v[2];
  finish_sum<2>(v, partials, nblocks, sh);
  if (threadIdx.x == 0) {
    st->pq = v[0];
    st->pr = v[1];
    st->alpha = v[1] / v[0];
    st->pq_old = v[0];
  }
}

// x += alpha p ; r -= alpha q ; r32 = r ; partial <r,r>
__global__ __launch_bounds__(kBlock) void k_fcg_update(int64_t n, const double* __restrict__ p,
                                                       const double* __restrict__ q,
                                                       double* __restrict__ x,
                                                       double* __restrict__ r,
                                                       float* __restrict__ r32,
                                                       const FcgState* st, double* partials) {
  __shared__ double sh[kBlock / 64];
  const double alpha = st->alpha;
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    r32[i] = (float)ri;
    v[0] += ri * ri;
  }
  block_partials<1>(v, partials, sh);
}

__global__ __launch_bounds__(kBlock) void k_fcg_fin_r(FcgState* st, const double* partials,
                                                      int nblocks) {
  __shared__ double sh[kBlock / 64];
  double v[1];
  finish_sum<1>(v, partials, nblocks, sh);
  if (threadIdx.x == 0) {
    st->rr = v[0];
    st->iter += 1;
  }
}

int grid_for(int64_t total) {
  int64_t blocks = (total + kBlock - 1) / kBlock;
  const int64_t cap = 256 * 8;  // 8 blocks of 256 per CU; grid-stride loops do the rest
  return (int)(blocks < cap ? (blocks < 1 ? 1 : blocks) : cap);
}

}  // namespace
}  // namespace tpf

using namespace tpf;

extern "C" tpf_status tpf_ppe_create(tpf_ppe** out, int ndim, const int64_t* n_nodes,
                                     const double* L, int64_t pin_dof, void* stream) {
  if (!out || !n_nodes || !L) return fail(TPF_ERR_VALUE, "null argument");
  *out = nullptr;
  if (ndim < 1 || ndim > 3) return fail(TPF_ERR_VALUE, "ndim must be 1, 2 or 3");
  auto* s = new tpf_ppe();
  PpeGeom& g = s->g;
  int n[3] = {1, 1, 1};
  double h[3] = {1.0, 1.0, 1.0};
  for (int a = 0; a < ndim; ++a) {
    const int slot = 3 - ndim + a;  // right-align so the last axis is contiguous
    if (n_nodes[a] < 2 || !(L[a] > 0.0)) {
      delete s;
      return fail(TPF_ERR_VALUE, "PPE axis needs >= 2 nodes and L > 0");
    }
    n[slot] = (int)n_nodes[a];
    h[slot] = L[a] / (double)(n_nodes[a] - 1);
  }
  g.n0 = n[0];
  g.n1 = n[1];
  g.n2 = n[2];
  g.total = (int64_t)n[0] * n[1] * n[2];
  if (g.total >= (int64_t)0xffffffffLL) {
    delete s;
    return fail(TPF_ERR_VALUE, "PPE grid too large (2^32 nodes)");
  }
  for (int a = 0; a < 3; ++a) {
    g.h[a] = h[a];
    g.ih[a] = 1.0 / h[a];
  }
  if (pin_dof < 0 || pin_dof >= g.total) {
    delete s;
    return fail(TPF_ERR_VALUE, "pin_dof out of range");
  }
  g.pin = pin_dof;
  s->user_stream = (hipStream_t)stream;
  s->grid = grid_for(g.total);
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    return fail(TPF_ERR_RUNTIME, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  const size_t bytes = sizeof(double) * (size_t)g.total;
  double** bufs[] = {&s->dinv, &s->b, &s->r, &s->p, &s->q};
  for (double** bp : bufs) {
    e = hipMalloc(bp, bytes);
    if (e != hipSuccess) {
      tpf_ppe_destroy(s);
      return fail(TPF_ERR_RUNTIME, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
  }
  if (hipMalloc(&s->partials, sizeof(double) * 4 * s->grid) != hipSuccess ||
      hipMalloc(&s->st, sizeof(PcgState)) != hipSuccess ||
      hipHostMalloc(&s->st_host, sizeof(PcgState)) != hipSuccess) {
    tpf_ppe_destroy(s);
    return fail(TPF_ERR_RUNTIME, "PPE workspace allocation failed");
  }
  PcgState zero{};
  if (hipMemcpy(s->st, &zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess) {
    tpf_ppe_destroy(s);
    return fail(TPF_ERR_RUNTIME, "PPE state init failed");
  }
  *out = s;
  return TPF_OK;
}

extern "C" tpf_status tpf_ppe_destroy(tpf_ppe* s) {
  if (!s) return TPF_OK;
  tpf_gmres_release(s);
  tpf_mg_release(s);
  if (s->fcg_st) (void)hipFree(s->fcg_st);
  if (s->fcg_host) (void)hipHostFree(s->fcg_host);
  if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
  if (s->graph32_exec) (void)hipGraphExecDestroy(s->graph32_exec);
  float* fb[] = {s->r32, s->p32, s->q32, s->x32, s->dinv32, s->rho32};
  for (float* b : fb)
    if (b) (void)hipFree(b);
  if (s->st_in) (void)hipFree(s->st_in);
  if (s->st_in_host) (void)hipHostFree(s->st_in_host);
  double* bufs[] = {s->dinv, s->b, s->r, s->p, s->q, s->partials};
  for (double* b : bufs)
    if (b) (void)hipFree(b);
  if (s->st) (void)hipFree(s->st);
  if (s->st_host) (void)hipHostFree(s->st_host);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return TPF_OK;
}

namespace tpf {
// Orders s->stream after the caller's stream.
static tpf_status join_user_stream(tpf_ppe* s) {
  hipEvent_t ev;
  TPF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  TPF_HIP(hipEventRecord(ev, s->user_stream));
  TPF_HIP(hipStreamWaitEvent(s->stream, ev, 0));
  TPF_HIP(hipEventDestroy(ev));
  return TPF_OK;
}
}  // namespace tpf

extern "C" tpf_status tpf_ppe_set_density(tpf_ppe* s, const double* rho) {
  if (!s || !rho) return fail(TPF_ERR_VALUE, "null argument");
  s->rho = rho;
  s->rho32_src = nullptr;  // derived FP32 copies and the multigrid hierarchy are rebuilt
  mg_invalidate(s);
  TPF_TRY(join_user_stream(s));
  hipLaunchKernelGGL(k_diag, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g, rho, s->dinv);
  TPF_HIP(hipGetLastError());
  TPF_HIP(hipStreamSynchronize(s->stream));
  return TPF_OK;
}

extern "C" tpf_status tpf_ppe_apply(tpf_ppe* s, const double* p, double* out) {
  if (!s || !p || !out) return fail(TPF_ERR_VALUE, "null argument");
  if (!s->rho) return fail(TPF_ERR_RUNTIME, "PPE density not set (call tpf_ppe_set_density)");
  TPF_TRY(join_user_stream(s));
  hipLaunchKernelGGL(k_apply_L, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g, p, s->rho, out);
  TPF_HIP(hipGetLastError());
  TPF_HIP(hipStreamSynchronize(s->stream));
  return TPF_OK;
}

namespace tpf {
static tpf_status read_state(tpf_ppe* s) {
  TPF_HIP(hipMemcpyAsync(s->st_host, s->st, sizeof(PcgState), hipMemcpyDeviceToHost, s->stream));
  TPF_HIP(hipStreamSynchronize(s->stream));
  return TPF_OK;
}

// Records `chunk` PCG iterations into a graph (once per handle and chunk).
static tpf_status ensure_graph(tpf_ppe* s, double* x, int chunk) {
  if (s->graph_exec && s->graph_x == x && s->graph_chunk == chunk && s->graph_rho == s->rho)
    return TPF_OK;
  if (s->graph_exec) {
    TPF_HIP(hipGraphExecDestroy(s->graph_exec));
    s->graph_exec = nullptr;
  }
  hipGraph_t graph;
  TPF_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  for (int it = 0; it < chunk; ++it) {
    hipLaunchKernelGGL(k_update_p<double>, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g,
                       (const double*)s->r, (const double*)s->dinv, s->p, x, s->st);
    hipLaunchKernelGGL(k_matvec<double>, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g,
                       (const double*)s->p, s->rho, s->q, s->st, s->partials);
    hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(kBlock), 0, s->stream, s->st,
                       (const double*)s->partials, s->grid);
    hipLaunchKernelGGL(k_update_r<double>, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g,
                       (const double*)s->q, (const double*)s->dinv, s->r, s->st, s->partials);
    hipLaunchKernelGGL(k_fin_beta, dim3(1), dim3(kBlock), 0, s->stream, s->st,
                       (const double*)s->partials, s->grid);
  }
  TPF_HIP(hipStreamEndCapture(s->stream, &graph));
  hipError_t e = hipGraphInstantiate(&s->graph_exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess)
    return fail(TPF_ERR_RUNTIME, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  s->graph_x = x;
  s->graph_chunk = chunk;
  s->graph_rho = s->rho;
  return TPF_OK;
}
}  // namespace tpf

extern "C" tpf_status tpf_ppe_true_relres(tpf_ppe* s, const double* rhs, const double* x,
                                          double* relres) {
  if (!s || !rhs || !x || !relres) return fail(TPF_ERR_VALUE, "null argument");
  if (!s->rho) return fail(TPF_ERR_RUNTIME, "PPE density not set");
  TPF_TRY(join_user_stream(s));
  hipLaunchKernelGGL(k_true_res, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g, rhs, s->rho, x,
                     s->st, s->partials);
  hipLaunchKernelGGL(k_fin_true, dim3(1), dim3(kBlock), 0, s->stream, s->st,
                     (const double*)s->partials, s->grid);
  TPF_HIP(hipGetLastError());
  TPF_TRY(read_state(s));
  const double bb = s->st_host->true_bb;
  *relres = bb > 0.0 ? std::sqrt(s->st_host->true_rr / bb) : std::sqrt(s->st_host->true_rr);
  return TPF_OK;
}

extern "C" tpf_status tpf_ppe_solve(tpf_ppe* s, const double* rhs, double* x,
                                    const tpf_solve_opts* opts, tpf_solve_report* rep) {
  if (!s || !rhs || !x || !opts) return fail(TPF_ERR_VALUE, "null argument");
  if (!s->rho) return fail(TPF_ERR_RUNTIME, "PPE density not set (call tpf_ppe_set_density)");
  if (opts->precond != 1 && opts->precond != 0)
    return fail(TPF_ERR_VALUE, "CG supports preconditioner 'jacobi' or 'none'");
  if (opts->mixed) return tpf_ppe_solve_mixed(s, rhs, x, opts, rep);
  if (opts->precond == 0) return fail(TPF_ERR_VALUE, "unpreconditioned CG is not built");
  if (!(opts->tol > 0.0) || opts->maxiter < 1) return fail(TPF_ERR_VALUE, "bad tol/maxiter");
  TPF_TRY(join_user_stream(s));
  if (opts->use_x0) {  // x0[pin] = 0 (fvm_matrixfree.py:185-186)
    hipLaunchKernelGGL(k_set_one, dim3(1), dim3(64), 0, s->stream, x, s->g.pin, 0.0);
    TPF_HIP(hipGetLastError());
  }
  PcgState init{};
  init.tol2 = opts->tol * opts->tol;
  // Residual replacement: when the recursive ||r|| has met tol but the FP64
  // true residual has not, restart CG from the current x (at most 4 times).
  int64_t iters = 0;
  double rr = 0.0, bb = 0.0, tr = 0.0;
  const int chunk = opts->check_every > 0 ? opts->check_every : 32;
  for (int restart = 0;; ++restart) {
    const int zero_x = (restart == 0 && !opts->use_x0) ? 1 : 0;
    init.maxiter = opts->maxiter - iters;
    *s->st_host = init;  // pinned staging for the async copy
    TPF_HIP(hipMemcpyAsync(s->st, s->st_host, sizeof(init), hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_init, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g, rhs, s->rho,
                       s->dinv, x, s->b, s->r, s->p, s->st, s->partials, zero_x);
    hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(kBlock), 0, s->stream, s->st,
                       (const double*)s->partials, s->grid);
    TPF_HIP(hipGetLastError());
    TPF_TRY(ensure_graph(s, x, chunk));
    TPF_TRY(read_state(s));
    if (s->st_host->bb == 0.0) {  // b = 0: SciPy's cg returns b itself, whatever x0
      hipLaunchKernelGGL(k_fill_zero, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g.total, x);
      TPF_HIP(hipGetLastError());
      TPF_HIP(hipStreamSynchronize(s->stream));
      rr = bb = tr = 0.0;
      break;
    }
    while (!s->st_host->done) {
      TPF_HIP(hipGraphLaunch(s->graph_exec, s->stream));
      TPF_TRY(read_state(s));
    }
    hipLaunchKernelGGL(k_finish_x<double>, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g,
                       (const double*)s->p, x, s->st);
    TPF_HIP(hipGetLastError());
    iters += s->st_host->iter;
    rr = s->st_host->rr;
    bb = s->st_host->bb;
    TPF_TRY(tpf_ppe_true_relres(s, rhs, x, &tr));
    const bool rec_ok = bb > 0 ? std::sqrt(rr / bb) <= opts->tol : true;
    if (tr <= opts->tol || !rec_ok || iters >= opts->maxiter || restart >= 4) break;
  }
  tpf_solve_report r{};
  r.iterations = iters;
  r.relres = bb > 0 ? std::sqrt(rr / bb) : 0.0;
  r.true_relres = tr;
  r.converged = tr <= opts->tol ? 1 : 0;
  r.refinements = 0;
  if (rep) *rep = r;
  if (r.relres > opts->tol)
    return fail(TPF_ERR_RUNTIME, "PPE CG did not converge cleanly (maxiter " +
                                     std::to_string(opts->maxiter) + " reached, ||r||/||b|| = " +
                                     std::to_string(r.relres) + ")");
  return TPF_OK;
}

// ---- mixed precision: FP32 inner PCG under FP64 iterative refinement -------
//
// Outer loop (FP64): r = b - A x on the reference's CG system, stop when
// ||r||/||b|| <= tol.  Inner loop (FP32 vectors, FP32 rho, FP64 scalars):
// PCG on A d = r to a relative reduction kInnerTol, then x += d.
namespace tpf {
namespace {
constexpr double kInnerTol = 1e-4;
constexpr int kMaxOuter = 40;

tpf_status ensure_mixed(tpf_ppe* s) {
  const size_t n = (size_t)s->g.total;
  if (!s->r32) {
    float** fb[] = {&s->r32, &s->p32, &s->q32, &s->x32, &s->dinv32, &s->rho32};
    for (float** b : fb) TPF_HIP(hipMalloc(b, n * sizeof(float)));
    TPF_HIP(hipMalloc(&s->st_in, sizeof(PcgState)));
    TPF_HIP(hipHostMalloc(&s->st_in_host, sizeof(PcgState)));
    PcgState zero{};
    TPF_HIP(hipMemcpy(s->st_in, &zero, sizeof(zero), hipMemcpyHostToDevice));
  }
  if (s->rho32_src != s->rho) {
    hipLaunchKernelGGL(k_to_f32, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g.total, s->rho,
                       s->rho32, (const double*)s->dinv, s->dinv32);
    TPF_HIP(hipGetLastError());
    s->rho32_src = s->rho;
  }
  return TPF_OK;
}

tpf_status ensure_graph32(tpf_ppe* s, int chunk) {
  if (s->graph32_exec && s->graph32_chunk == chunk && s->graph32_rho == s->rho) return TPF_OK;
  if (s->graph32_exec) {
    TPF_HIP(hipGraphExecDestroy(s->graph32_exec));
    s->graph32_exec = nullptr;
  }
  hipGraph_t graph;
  TPF_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  for (int it = 0; it < chunk; ++it) {
    hipLaunchKernelGGL(k_update_p<float>, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g,
                       (const float*)s->r32, (const float*)s->dinv32, s->p32, s->x32, s->st_in);
    hipLaunchKernelGGL(k_matvec<float>, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g,
                       (const float*)s->p32, (const float*)s->rho32, s->q32, s->st_in,
                       s->partials);
    hipLaunchKernelGGL(k_fin_alpha, dim3(1), dim3(kBlock), 0, s->stream, s->st_in,
                       (const double*)s->partials, s->grid);
    hipLaunchKernelGGL(k_update_r<float>, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g,
                       (const float*)s->q32, (const float*)s->dinv32, s->r32, s->st_in,
                       s->partials);
    hipLaunchKernelGGL(k_fin_beta, dim3(1), dim3(kBlock), 0, s->stream, s->st_in,
                       (const double*)s->partials, s->grid);
  }
  TPF_HIP(hipStreamEndCapture(s->stream, &graph));
  hipError_t e = hipGraphInstantiate(&s->graph32_exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess)
    return fail(TPF_ERR_RUNTIME, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  s->graph32_chunk = chunk;
  s->graph32_rho = s->rho;
  return TPF_OK;
}

tpf_status read_inner(tpf_ppe* s) {
  TPF_HIP(hipMemcpyAsync(s->st_in_host, s->st_in, sizeof(PcgState), hipMemcpyDeviceToHost,
                         s->stream));
  TPF_HIP(hipStreamSynchronize(s->stream));
  return TPF_OK;
}
}  // namespace
}  // namespace tpf

tpf_status tpf_ppe_solve_mixed(tpf_ppe* s, const double* rhs, double* x,
                               const tpf_solve_opts* opts, tpf_solve_report* rep) {
  if (opts->precond != 1) return fail(TPF_ERR_VALUE, "mixed solve needs the Jacobi preconditioner");
  if (!(opts->tol > 0.0) || opts->maxiter < 1) return fail(TPF_ERR_VALUE, "bad tol/maxiter");
  TPF_TRY(join_user_stream(s));
  TPF_TRY(ensure_mixed(s));
  if (opts->use_x0) {  // x0[pin] = 0 (fvm_matrixfree.py:185-186)
    hipLaunchKernelGGL(k_set_one, dim3(1), dim3(64), 0, s->stream, x, s->g.pin, 0.0);
    TPF_HIP(hipGetLastError());
  } else {
    hipLaunchKernelGGL(k_fill_zero, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g.total, x);
    TPF_HIP(hipGetLastError());
  }
  const int chunk = opts->check_every > 0 ? opts->check_every : 32;
  TPF_TRY(ensure_graph32(s, chunk));
  int64_t total_inner = 0;
  int outer = 0;
  double relres = 0.0;
  for (;;) {
    hipLaunchKernelGGL(k_resid_mixed, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g, rhs,
                       s->rho, (const double*)x, (const float*)s->dinv32, s->r32, s->x32, s->p32,
                       s->st_in, s->partials);
    hipLaunchKernelGGL(k_fin_mixed, dim3(1), dim3(kBlock), 0, s->stream, s->st_in,
                       (const double*)s->partials, s->grid);
    TPF_HIP(hipGetLastError());
    TPF_TRY(read_inner(s));
    PcgState st = *s->st_in_host;
    const double bb = st.true_bb;
    if (bb == 0.0) {  // b = 0: SciPy's cg returns b itself, whatever x0
      hipLaunchKernelGGL(k_fill_zero, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g.total, x);
      TPF_HIP(hipGetLastError());
      relres = 0.0;
      break;
    }
    relres = bb > 0.0 ? std::sqrt(st.true_rr / bb) : std::sqrt(st.true_rr);
    if (relres <= opts->tol || outer >= kMaxOuter || total_inner >= opts->maxiter) break;
    st.rr = st.true_rr;
    st.bb = st.true_rr;  // inner stop: ||r_in|| <= kInnerTol ||r_outer||
    const double itol = opts->inner_tol > 0.0 ? opts->inner_tol : kInnerTol;
    st.tol2 = itol * itol;
    st.alpha = st.beta = 0.0;
    st.iter = 0;
    st.maxiter = opts->maxiter - total_inner;
    st.done = 0;
    st.counter = 0;
    *s->st_in_host = st;  // pinned staging for the async copy
    TPF_HIP(hipMemcpyAsync(s->st_in, s->st_in_host, sizeof(st), hipMemcpyHostToDevice, s->stream));
    do {
      TPF_HIP(hipGraphLaunch(s->graph32_exec, s->stream));
      TPF_TRY(read_inner(s));
    } while (!s->st_in_host->done);
    hipLaunchKernelGGL(k_finish_x<float>, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g,
                       (const float*)s->p32, s->x32, s->st_in);
    hipLaunchKernelGGL(k_accumulate, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g,
                       (const float*)s->x32, x);
    TPF_HIP(hipGetLastError());
    total_inner += s->st_in_host->iter;
    ++outer;
  }
  TPF_HIP(hipStreamSynchronize(s->stream));
  tpf_solve_report r{};
  r.iterations = total_inner;
  r.relres = relres;
  r.true_relres = relres;
  r.converged = relres <= opts->tol ? 1 : 0;
  r.refinements = outer;
  if (rep) *rep = r;
  if (!r.converged)
    return fail(TPF_ERR_RUNTIME, "mixed-precision PPE did not converge (||r||/||b|| = " +
                                     std::to_string(relres) + " after " +
                                     std::to_string(total_inner) + " inner iterations)");
  return TPF_OK;
}

extern "C" tpf_status tpf_ppe_system_apply(tpf_ppe* s, const double* p, double* out) {
  if (!s || !p || !out) return fail(TPF_ERR_VALUE, "null argument");
  if (!s->rho) return fail(TPF_ERR_RUNTIME, "PPE density not set");
  // asynchronous on the caller's stream (the one given at create time)
  hipLaunchKernelGGL(k_system_apply<double>, dim3(s->grid), dim3(kBlock), 0, s->user_stream,
                     s->g, p, s->rho, out);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

extern "C" tpf_status tpf_ppe_prepare_f32(tpf_ppe* s) {
  if (!s) return fail(TPF_ERR_VALUE, "null argument");
  if (!s->rho) return fail(TPF_ERR_RUNTIME, "PPE density not set");
  TPF_TRY(join_user_stream(s));
  TPF_TRY(ensure_mixed(s));
  TPF_HIP(hipStreamSynchronize(s->stream));
  return TPF_OK;
}

extern "C" tpf_status tpf_ppe_jacobi_inv(tpf_ppe* s, double* out) {
  if (!s || !out) return fail(TPF_ERR_VALUE, "null argument");
  if (!s->rho) return fail(TPF_ERR_RUNTIME, "PPE density not set");
  TPF_HIP(hipMemcpyAsync(out, s->dinv, sizeof(double) * s->g.total, hipMemcpyDeviceToDevice,
                         s->user_stream));
  return TPF_OK;
}

extern "C" tpf_status tpf_ppe_system_rhs(tpf_ppe* s, const double* rhs, double* b) {
  if (!s || !rhs || !b) return fail(TPF_ERR_VALUE, "null argument");
  hipLaunchKernelGGL(k_system_rhs, dim3(s->grid), dim3(kBlock), 0, s->user_stream, s->g, rhs, b);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

extern "C" tpf_status tpf_ppe_system_apply_f32(tpf_ppe* s, const float* p, float* out) {
  if (!s || !p || !out) return fail(TPF_ERR_VALUE, "null argument");
  if (!s->rho) return fail(TPF_ERR_RUNTIME, "PPE density not set");
  if (!s->rho32 || s->rho32_src != s->rho)
    return fail(TPF_ERR_RUNTIME, "FP32 operator copies not built (call tpf_ppe_prepare_f32)");
  hipLaunchKernelGGL(k_system_apply<float>, dim3(s->grid), dim3(kBlock), 0, s->user_stream, s->g,
                     p, (const float*)s->rho32, out);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

// ---- tpf_ppe_solve_mg -------------------------------------------------------
namespace tpf {
struct Deflation {  // learned coarse correction after each cycle (k = 0: none)
  const float* W;
  int k, level;
};
static tpf_status solve_mg(tpf_ppe* s, const double* rhs, double* x, const tpf_mg_opts* o,
                           tpf_solve_report* rep, const Deflation& defl,
                           const double* einv = nullptr);
}  // namespace tpf

extern "C" tpf_status tpf_ppe_solve_mg(tpf_ppe* s, const double* rhs, double* x,
                                       const tpf_mg_opts* o, tpf_solve_report* rep) {
  return tpf::solve_mg(s, rhs, x, o, rep, tpf::Deflation{nullptr, 0, 0});
}

extern "C" tpf_status tpf_ppe_solve_mg_deflated(tpf_ppe* s, const double* rhs, double* x,
                                                const tpf_mg_opts* o, const float* W, int k,
                                                int level, const double* einv,
                                                tpf_solve_report* rep) {
  if (!W || !einv) return fail(TPF_ERR_VALUE, "null deflation argument");
  return tpf::solve_mg(s, rhs, x, o, rep, tpf::Deflation{W, k, level}, einv);
}

namespace tpf {
static tpf_status solve_mg(tpf_ppe* s, const double* rhs, double* x, const tpf_mg_opts* o,
                           tpf_solve_report* rep, const Deflation& defl, const double* einv) {
  if (!s || !rhs || !x || !o) return fail(TPF_ERR_VALUE, "null argument");
  if (!s->rho) return fail(TPF_ERR_RUNTIME, "PPE density not set (call tpf_ppe_set_density)");
  if (!(o->tol > 0.0) || o->maxiter < 1) return fail(TPF_ERR_VALUE, "bad tol/maxiter");
  if (o->gamma < 1 || o->gamma > 2 || o->nu < 1 || o->coarse_sweeps < 1)
    return fail(TPF_ERR_VALUE, "bad multigrid cycle options (gamma 1|2, nu >= 1, sweeps >= 1)");
  const int max_levels = o->max_levels > 0 ? o->max_levels : 12;
  TPF_TRY(mg_ensure(s, max_levels));  // on the caller's stream, synchronous
  if (defl.k > 0) TPF_TRY(mg_deflate_setup(s, defl.k, defl.level, einv));
  TPF_TRY(join_user_stream(s));
  hipStream_t st = s->stream;
  // workspace: r in s->r, p in s->p, q in s->q (FP64); FP32 residual and
  // preconditioned residual live in the hierarchy's level-0 b and x.
  if (!s->fcg_st) {
    TPF_HIP(hipMalloc(&s->fcg_st, sizeof(FcgState)));
    TPF_HIP(hipHostMalloc(&s->fcg_host, sizeof(FcgState)));
  }
  FcgState* dst = static_cast<FcgState*>(s->fcg_st);
  FcgState* hst = static_cast<FcgState*>(s->fcg_host);
  float* r32 = mg_fine_b(s);
  const float* z32 = mg_fine_x(s);
  const int64_t n = s->g.total;
  const dim3 grid(s->grid), blk(kBlock);
  const double tol2 = o->tol * o->tol;
  auto read = [&]() -> tpf_status {
    TPF_HIP(hipMemcpyAsync(hst, dst, sizeof(FcgState), hipMemcpyDeviceToHost, st));
    TPF_HIP(hipStreamSynchronize(st));
    return TPF_OK;
  };
  if (o->use_x0) hipLaunchKernelGGL(k_set_one, dim3(1), dim3(64), 0, st, x, s->g.pin, 0.0);
  int64_t iters = 0;
  double rr = 0.0, bb = 0.0, tr = 0.0;
  int restarts = 0;
  bool conv = false;
  // Residual replacement: when the recursive ||r|| meets tol but the FP64
  // true residual does not, FCG restarts from x (at most 4 times).
  for (;; ++restarts) {
    const int zero_x = (restarts == 0 && !o->use_x0) ? 1 : 0;
    if (zero_x) hipLaunchKernelGGL(k_fill_zero, grid, blk, 0, st, n, x);
    hipLaunchKernelGGL(k_fcg_init, grid, blk, 0, st, s->g, rhs, s->rho, x, s->r, r32, 0,
                       s->partials);
    hipLaunchKernelGGL(k_fcg_fin_init, dim3(1), blk, 0, st, dst, (const double*)s->partials,
                       s->grid);
    TPF_HIP(hipGetLastError());
    TPF_TRY(read());
    rr = hst->rr;
    bb = hst->bb;
    if (bb == 0.0) {
      hipLaunchKernelGGL(k_fill_zero, grid, blk, 0, st, n, x);
      TPF_HIP(hipStreamSynchronize(st));
      conv = true;
      tr = 0.0;
      break;
    }
    bool hit = rr <= tol2 * bb;
    while (!hit && iters < o->maxiter) {
      TPF_TRY(mg_cycle(s, st, o->nu, o->gamma, o->coarse_sweeps, o->w_levels));
      if (defl.k > 0) TPF_TRY(mg_deflate(s, st, defl.W, defl.k, defl.level));
      hipLaunchKernelGGL(k_fcg_zq, grid, blk, 0, st, n, z32, (const double*)s->q, s->partials);
      hipLaunchKernelGGL(k_fcg_fin_beta, dim3(1), blk, 0, st, dst, (const double*)s->partials,
                         s->grid);
      hipLaunchKernelGGL(k_fcg_p, grid, blk, 0, st, n, z32, s->p, (const FcgState*)dst);
      hipLaunchKernelGGL(k_fcg_matvec, grid, blk, 0, st, s->g, (const double*)s->p, s->rho,
                         (const double*)s->r, s->q, s->partials);
      hipLaunchKernelGGL(k_fcg_fin_alpha, dim3(1), blk, 0, st, dst, (const double*)s->partials,
                         s->grid);
      hipLaunchKernelGGL(k_fcg_update, grid, blk, 0, st, n, (const double*)s->p,
                         (const double*)s->q, x, s->r, r32, (const FcgState*)dst, s->partials);
      hipLaunchKernelGGL(k_fcg_fin_r, dim3(1), blk, 0, st, dst, (const double*)s->partials,
                         s->grid);
      TPF_HIP(hipGetLastError());
      TPF_TRY(read());
      ++iters;
      rr = hst->rr;
      hit = rr <= tol2 * bb;
    }
    // FP64 true residual of the CG system judges the result
    hipLaunchKernelGGL(k_set_one, dim3(1), dim3(64), 0, st, x, s->g.pin, 0.0);
    hipLaunchKernelGGL(k_true_res, grid, blk, 0, st, s->g, rhs, s->rho, (const double*)x, s->st,
                       s->partials);
    hipLaunchKernelGGL(k_fin_true, dim3(1), blk, 0, st, s->st, (const double*)s->partials,
                       s->grid);
    TPF_HIP(hipGetLastError());
    TPF_TRY(read_state(s));
    tr = s->st_host->true_bb > 0.0 ? std::sqrt(s->st_host->true_rr / s->st_host->true_bb) : 0.0;
    conv = tr <= o->tol;
    if (conv || iters >= o->maxiter || restarts >= 4) break;
  }
  if (rep) {
    rep->iterations = iters;
    rep->relres = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
    rep->true_relres = tr;
    rep->converged = conv ? 1 : 0;
    rep->refinements = restarts;
  }
  if (!conv)
    return fail(TPF_ERR_RUNTIME, "multigrid FCG did not reach tol in maxiter iterations");
  return TPF_OK;
}
}  // namespace tpf

// ---- many multigrid-FCG solves advanced together ----------------------------
//
// tpf_ppe_solve_mg on every handle, with the solves interleaved: each round
// enqueues the next step of every unfinished problem on that problem's own
// stream (restart set-up, one FCG iteration, or the FP64 true residual) with
// an asynchronous copy of its scalars, then waits on the streams and updates
// each problem's state machine.  A problem runs exactly the kernels of its
// single solve in the same order on its own stream, so its result equals
// tpf_ppe_solve_mg bit for bit; the streams of different problems overlap,
// where single solves wait on the host once per iteration each.
extern "C" tpf_status tpf_ppe_solve_mg_many(tpf_ppe* const* handles, int n,
                                            const double* const* rhs, double* const* x,
                                            const tpf_mg_opts* o, tpf_solve_report* reps) {
  using namespace tpf;
  if (!handles || n < 0 || !rhs || !x || !o || !reps)
    return fail(TPF_ERR_VALUE, "null argument");
  if (!(o->tol > 0.0) || o->maxiter < 1) return fail(TPF_ERR_VALUE, "bad tol/maxiter");
  if (o->gamma < 1 || o->gamma > 2 || o->nu < 1 || o->coarse_sweeps < 1)
    return fail(TPF_ERR_VALUE, "bad multigrid cycle options (gamma 1|2, nu >= 1, sweeps >= 1)");
  enum Phase { kInit, kIter, kTrue, kDone };
  struct Prob {
    tpf_ppe* s = nullptr;
    FcgState* dst = nullptr;
    FcgState* hst = nullptr;
    int64_t iters = 0;
    double rr = 0.0, bb = 0.0, tr = 0.0;
    int restarts = 0;
    bool conv = false, pending = false;
    Phase phase = kDone;
  };
  std::vector<Prob> P((size_t)n);
  const int max_levels = o->max_levels > 0 ? o->max_levels : 12;
  const double tol2 = o->tol * o->tol;
  for (int i = 0; i < n; ++i) {
    reps[i] = tpf_solve_report{};
    tpf_ppe* s = handles[i];
    if (!s || !rhs[i] || !x[i])
      return fail(TPF_ERR_VALUE, "null handle, rhs or solution for problem " + std::to_string(i));
    for (int j = 0; j < i; ++j)
      if (handles[j] == s) return fail(TPF_ERR_VALUE, "a handle appears twice in one call");
    if (!s->rho) return fail(TPF_ERR_RUNTIME, "PPE density not set (call tpf_ppe_set_density)");
    TPF_TRY(mg_ensure(s, max_levels));
    TPF_TRY(join_user_stream(s));
    if (!s->fcg_st) {
      TPF_HIP(hipMalloc(&s->fcg_st, sizeof(FcgState)));
      TPF_HIP(hipHostMalloc(&s->fcg_host, sizeof(FcgState)));
    }
    Prob& q = P[(size_t)i];
    q.s = s;
    q.dst = static_cast<FcgState*>(s->fcg_st);
    q.hst = static_cast<FcgState*>(s->fcg_host);
    q.phase = kInit;
    if (o->use_x0) hipLaunchKernelGGL(k_set_one, dim3(1), dim3(64), 0, s->stream, x[i], s->g.pin, 0.0);
  }
  for (;;) {
    bool any = false;
    for (int i = 0; i < n; ++i) {  // enqueue every problem's next step
      Prob& q = P[(size_t)i];
      if (q.phase == kDone) continue;
      tpf_ppe* s = q.s;
      hipStream_t st = s->stream;
      const int64_t nn = s->g.total;
      const dim3 grid(s->grid), blk(kBlock);
      float* r32 = mg_fine_b(s);
      const float* z32 = mg_fine_x(s);
      if (q.phase == kInit) {
        const int zero_x = (q.restarts == 0 && !o->use_x0) ? 1 : 0;
        if (zero_x) hipLaunchKernelGGL(k_fill_zero, grid, blk, 0, st, nn, x[i]);
        hipLaunchKernelGGL(k_fcg_init, grid, blk, 0, st, s->g, rhs[i], s->rho, x[i], s->r, r32,
                           0, s->partials);
        hipLaunchKernelGGL(k_fcg_fin_init, dim3(1), blk, 0, st, q.dst,
                           (const double*)s->partials, s->grid);
        TPF_HIP(hipMemcpyAsync(q.hst, q.dst, sizeof(FcgState), hipMemcpyDeviceToHost, st));
      } else if (q.phase == kIter) {
        TPF_TRY(mg_cycle(s, st, o->nu, o->gamma, o->coarse_sweeps, o->w_levels));
        hipLaunchKernelGGL(k_fcg_zq, grid, blk, 0, st, nn, z32, (const double*)s->q, s->partials);
        hipLaunchKernelGGL(k_fcg_fin_beta, dim3(1), blk, 0, st, q.dst,
                           (const double*)s->partials, s->grid);
        hipLaunchKernelGGL(k_fcg_p, grid, blk, 0, st, nn, z32, s->p, (const FcgState*)q.dst);
        hipLaunchKernelGGL(k_fcg_matvec, grid, blk, 0, st, s->g, (const double*)s->p, s->rho,
                           (const double*)s->r, s->q, s->partials);
        hipLaunchKernelGGL(k_fcg_fin_alpha, dim3(1), blk, 0, st, q.dst,
                           (const double*)s->partials, s->grid);
        hipLaunchKernelGGL(k_fcg_update, grid, blk, 0, st, nn, (const double*)s->p,
                           (const double*)s->q, x[i], s->r, r32, (const FcgState*)q.dst,
                           s->partials);
        hipLaunchKernelGGL(k_fcg_fin_r, dim3(1), blk, 0, st, q.dst, (const double*)s->partials,
                           s->grid);
        TPF_HIP(hipMemcpyAsync(q.hst, q.dst, sizeof(FcgState), hipMemcpyDeviceToHost, st));
      } else {  // kTrue: the FP64 true residual of the CG system judges the result
        hipLaunchKernelGGL(k_set_one, dim3(1), dim3(64), 0, st, x[i], s->g.pin, 0.0);
        hipLaunchKernelGGL(k_true_res, grid, blk, 0, st, s->g, rhs[i], s->rho,
                           (const double*)x[i], s->st, s->partials);
        hipLaunchKernelGGL(k_fin_true, dim3(1), blk, 0, st, s->st, (const double*)s->partials,
                           s->grid);
        TPF_HIP(hipMemcpyAsync(s->st_host, s->st, sizeof(PcgState), hipMemcpyDeviceToHost, st));
      }
      TPF_HIP(hipGetLastError());
      q.pending = true;
      any = true;
    }
    if (!any) break;
    for (int i = 0; i < n; ++i) {  // wait and advance every state machine
      Prob& q = P[(size_t)i];
      if (!q.pending) continue;
      q.pending = false;
      tpf_ppe* s = q.s;
      TPF_HIP(hipStreamSynchronize(s->stream));
      if (q.phase == kInit) {
        q.rr = q.hst->rr;
        q.bb = q.hst->bb;
        if (q.bb == 0.0) {  // zero right-hand side: x = 0
          hipLaunchKernelGGL(k_fill_zero, dim3(s->grid), dim3(kBlock), 0, s->stream, s->g.total,
                             x[i]);
          TPF_HIP(hipGetLastError());
          q.conv = true;
          q.tr = 0.0;
          q.phase = kDone;
          continue;
        }
        q.phase = (q.rr <= tol2 * q.bb || q.iters >= o->maxiter) ? kTrue : kIter;
      } else if (q.phase == kIter) {
        ++q.iters;
        q.rr = q.hst->rr;
        if (q.rr <= tol2 * q.bb || q.iters >= o->maxiter) q.phase = kTrue;
      } else {
        q.tr = s->st_host->true_bb > 0.0 ? std::sqrt(s->st_host->true_rr / s->st_host->true_bb)
                                         : 0.0;
        q.conv = q.tr <= o->tol;
        if (q.conv || q.iters >= o->maxiter || q.restarts >= 4) {
          q.phase = kDone;
        } else {
          ++q.restarts;
          q.phase = kInit;
        }
      }
    }
  }
  std::string failed;
  for (int i = 0; i < n; ++i) {
    const Prob& q = P[(size_t)i];
    TPF_HIP(hipStreamSynchronize(q.s->stream));  // zero right-hand sides' fills
    tpf_solve_report& r = reps[i];
    r.iterations = q.iters;
    r.relres = q.bb > 0.0 ? std::sqrt(q.rr / q.bb) : 0.0;
    r.true_relres = q.tr;
    r.converged = q.conv ? 1 : 0;
    r.refinements = q.restarts;
    if (!q.conv) failed += (failed.empty() ? "" : ", ") + std::to_string(i);
  }
  if (!failed.empty())
    return fail(TPF_ERR_RUNTIME, "multigrid FCG did not reach tol for problem(s) " + failed +
                                     " (maxiter " + std::to_string(o->maxiter) + ")");
  return TPF_OK;
}

// ---- batched Jacobi-PCG: independent problems iterated together ------------
//
// Every kernel of one batched iteration covers all problems: workgroup blk
// belongs to the problem whose block range [begin[i], begin[i+1]) holds it,
// and runs that problem's grid-stride loop with the problem's own workgroup
// count (the count tpf_ppe_solve uses), writing block partials into the
// problem's buffer; the finishing kernels run one workgroup per problem.
// Each problem's arithmetic and reduction order are those of tpf_ppe_solve,
// so its result is bit-identical to a single solve.  Chunks of iterations
// are captured once as a hipGraph (the descriptor table is read from device
// memory at run time) and replayed until every problem has converged.
namespace tpf {
namespace {

struct PcgProb {
  PpeGeom g;
  const double* rho;
  const double* dinv;
  const double* rhs;
  double *x, *b, *r, *p, *q, *partials;
  PcgState* st;
  int grid;
  int active;  // takes part in the current (re)start
  int zero_x;
};

// Flags gathered for the host after each chunk.
struct PcgFlags {
  double rr, bb, true_rr, true_bb;
  int64_t iter;
  int done, pad;
};

__device__ inline int prob_of(const unsigned* __restrict__ begin, int n, unsigned blk) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {  // last i with begin[i] <= blk
    const int mid = (lo + hi + 1) >> 1;
    if (begin[mid] <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int K>
__device__ inline void block_partials_at(double (&v)[K], double* partials, double* sh,
                                         unsigned blk) {
  block_sum<K>(v, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) partials[blk * K + k] = v[k];
}

#define PCG_BATCH_PROLOGUE                                 \
  const int pi = prob_of(begin, n, blockIdx.x);            \
  const PcgProb& P = tab[pi];                              \
  const unsigned blk = blockIdx.x - begin[pi];             \
  const int64_t nthreads = (int64_t)P.grid * kBlock;       \
  const PpeGeom& g = P.g;

__global__ __launch_bounds__(kBlock) void kb_init(const PcgProb* __restrict__ tab,
                                                  const unsigned* __restrict__ begin, int n) {
  PCG_BATCH_PROLOGUE
  if (!P.active) return;
  __shared__ double sh[3 * kBlock / 64];
  double v[3] = {0.0, 0.0, 0.0};
  for (int64_t idx = (int64_t)blk * kBlock + threadIdx.x; idx < g.total; idx += nthreads) {
    int i0, i1, i2;
    coords(g, idx, i0, i1, i2);
    const double W = dv(g, 0, i0) * dv(g, 1, i1) * dv(g, 2, i2);
    const double bi = idx == g.pin ? 0.0 : -W * P.rhs[idx];
    double ap = 0.0, dg;
    if (P.zero_x) {
      P.x[idx] = 0.0;
    } else {
      stencil<double>(g, P.x, P.rho, idx, i0, i1, i2, ap, dg);
      if (idx == g.pin) ap = P.x[idx];
    }
    const double ri = bi - ap;
    P.b[idx] = bi;
    P.r[idx] = ri;
    P.p[idx] = 0.0;
    v[0] += ri * ri;
    v[1] += ri * P.dinv[idx] * ri;
    v[2] += bi * bi;
  }
  block_partials_at<3>(v, P.partials, sh, blk);
}

__global__ __launch_bounds__(kBlock) void kb_fin_init(const PcgProb* __restrict__ tab) {
  const PcgProb& P = tab[blockIdx.x];
  if (!P.active) return;
  __shared__ double sh[3 * kBlock / 64];
  double v[3];
  finish_sum<3>(v, P.partials, P.grid, sh);
  if (threadIdx.x == 0) {
    PcgState* st = P.st;
    st->rr = v[0];
    st->rz = v[1];
    st->bb = v[2];
    st->beta = 0.0;
    st->alpha = 0.0;
    st->iter = 0;
    st->done = (v[0] <= st->tol2 * v[2] || v[2] == 0.0) ? 1 : 0;  // b = 0: x = 0
  }
}

__global__ __launch_bounds__(kBlock) void kb_update_p(const PcgProb* __restrict__ tab,
                                                      const unsigned* __restrict__ begin, int n) {
  PCG_BATCH_PROLOGUE
  const PcgState* st = P.st;
  if (st->done) return;
  const double beta = st->beta, alpha = st->alpha;
  for (int64_t idx = (int64_t)blk * kBlock + threadIdx.x; idx < g.total; idx += nthreads) {
    const double po = P.p[idx];
    P.x[idx] += alpha * po;
    P.p[idx] = P.dinv[idx] * P.r[idx] + beta * po;
  }
}

__global__ __launch_bounds__(kBlock) void kb_matvec(const PcgProb* __restrict__ tab,
                                                    const unsigned* __restrict__ begin, int n) {
  PCG_BATCH_PROLOGUE
  if (P.st->done) return;
  __shared__ double sh[kBlock / 64];
  double v[1] = {0.0};
  for (int64_t idx = (int64_t)blk * kBlock + threadIdx.x; idx < g.total; idx += nthreads) {
    int i0, i1, i2;
    coords(g, idx, i0, i1, i2);
    double ap, dg;
    stencil<double>(g, P.p, P.rho, idx, i0, i1, i2, ap, dg);
    if (idx == g.pin) ap = P.p[idx];
    P.q[idx] = ap;
    v[0] += P.p[idx] * ap;
  }
  block_partials_at<1>(v, P.partials, sh, blk);
}

__global__ __launch_bounds__(kBlock) void kb_fin_alpha(const PcgProb* __restrict__ tab) {
  const PcgProb& P = tab[blockIdx.x];
  PcgState* st = P.st;
  if (st->done) return;
  __shared__ double sh[kBlock / 64];
  double v[1];
  finish_sum<1>(v, P.partials, P.grid, sh);
  if (threadIdx.x == 0) {
    st->pq = v[0];
    st->alpha = st->rz / v[0];
  }
}

__global__ __launch_bounds__(kBlock) void kb_update_r(const PcgProb* __restrict__ tab,
                                                      const unsigned* __restrict__ begin, int n) {
  PCG_BATCH_PROLOGUE
  const PcgState* st = P.st;
  if (st->done) return;
  __shared__ double sh[2 * kBlock / 64];
  const double alpha = st->alpha;
  double v[2] = {0.0, 0.0};
  for (int64_t idx = (int64_t)blk * kBlock + threadIdx.x; idx < g.total; idx += nthreads) {
    const double ri = P.r[idx] - alpha * P.q[idx];
    P.r[idx] = ri;
    v[0] += ri * P.dinv[idx] * ri;
    v[1] += ri * ri;
  }
  block_partials_at<2>(v, P.partials, sh, blk);
}

__global__ __launch_bounds__(kBlock) void kb_fin_beta(const PcgProb* __restrict__ tab) {
  const PcgProb& P = tab[blockIdx.x];
  PcgState* st = P.st;
  if (st->done) return;
  __shared__ double sh[2 * kBlock / 64];
  double v[2];
  finish_sum<2>(v, P.partials, P.grid, sh);
  if (threadIdx.x == 0) {
    st->beta = v[0] / st->rz;
    st->rz = v[0];
    st->rr = v[1];
    st->iter += 1;
    if (v[1] <= st->tol2 * st->bb || st->iter >= st->maxiter) st->done = 1;
  }
}

__global__ __launch_bounds__(kBlock) void kb_finish_x(const PcgProb* __restrict__ tab,
                                                      const unsigned* __restrict__ begin, int n) {
  PCG_BATCH_PROLOGUE
  if (!P.active) return;
  const double alpha = P.st->alpha;
  const bool zero_b = P.st->bb == 0.0;  // SciPy's cg returns b itself when b = 0
  for (int64_t idx = (int64_t)blk * kBlock + threadIdx.x; idx < g.total; idx += nthreads) {
    P.x[idx] += alpha * P.p[idx];
    if (idx == g.pin || zero_b) P.x[idx] = 0.0;
  }
}

__global__ __launch_bounds__(kBlock) void kb_true_res(const PcgProb* __restrict__ tab,
                                                      const unsigned* __restrict__ begin, int n) {
  PCG_BATCH_PROLOGUE
  if (!P.active) return;
  __shared__ double sh[2 * kBlock / 64];
  double v[2] = {0.0, 0.0};
  for (int64_t idx = (int64_t)blk * kBlock + threadIdx.x; idx < g.total; idx += nthreads) {
    int i0, i1, i2;
    coords(g, idx, i0, i1, i2);
    const double W = dv(g, 0, i0) * dv(g, 1, i1) * dv(g, 2, i2);
    const double bi = idx == g.pin ? 0.0 : -W * P.rhs[idx];
    double ap, dg;
    stencil<double>(g, P.x, P.rho, idx, i0, i1, i2, ap, dg);
    if (idx == g.pin) ap = P.x[idx];
    const double ri = bi - ap;
    v[0] += ri * ri;
    v[1] += bi * bi;
  }
  block_partials_at<2>(v, P.partials, sh, blk);
}

__global__ __launch_bounds__(kBlock) void kb_fin_true(const PcgProb* __restrict__ tab) {
  const PcgProb& P = tab[blockIdx.x];
  if (!P.active) return;
  __shared__ double sh[2 * kBlock / 64];
  double v[2];
  finish_sum<2>(v, P.partials, P.grid, sh);
  if (threadIdx.x == 0) {
    P.st->true_rr = v[0];
    P.st->true_bb = v[1];
  }
}

__global__ void kb_pin_zero(const PcgProb* __restrict__ tab, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n && tab[i].active) tab[i].x[tab[i].g.pin] = 0.0;
}

__global__ void kb_gather(const PcgProb* __restrict__ tab, int n, PcgFlags* __restrict__ out) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  const PcgState* st = tab[i].st;
  PcgFlags f;
  f.rr = st->rr;
  f.bb = st->bb;
  f.true_rr = st->true_rr;
  f.true_bb = st->true_bb;
  f.iter = st->iter;
  f.done = st->done;
  f.pad = 0;
  out[i] = f;
}

#undef PCG_BATCH_PROLOGUE

}  // namespace
}  // namespace tpf

struct tpf_ppe_batch {
  int n = 0;
  std::vector<tpf_ppe*> h;
  hipStream_t stream = nullptr;
  tpf::PcgProb* d_tab = nullptr;
  tpf::PcgProb* h_tab = nullptr;      // pinned staging
  unsigned* d_begin = nullptr;
  tpf::PcgFlags* d_flags = nullptr;
  tpf::PcgFlags* h_flags = nullptr;   // pinned
  unsigned blocks = 0;
  hipStream_t user_stream = nullptr;
  hipGraphExec_t graph = nullptr;
  int graph_chunk = 0;
};

namespace tpf {
namespace {

tpf_status batch_gather(tpf_ppe_batch* b) {
  hipLaunchKernelGGL(kb_gather, dim3((b->n + 63) / 64), dim3(64), 0, b->stream,
                     (const PcgProb*)b->d_tab, b->n, b->d_flags);
  TPF_HIP(hipGetLastError());
  TPF_HIP(hipMemcpyAsync(b->h_flags, b->d_flags, sizeof(PcgFlags) * b->n, hipMemcpyDeviceToHost,
                         b->stream));
  TPF_HIP(hipStreamSynchronize(b->stream));
  return TPF_OK;
}

tpf_status batch_upload(tpf_ppe_batch* b) {
  TPF_HIP(hipMemcpyAsync(b->d_tab, b->h_tab, sizeof(PcgProb) * b->n, hipMemcpyHostToDevice,
                         b->stream));
  return TPF_OK;
}

tpf_status batch_graph(tpf_ppe_batch* b, int chunk) {
  if (b->graph && b->graph_chunk == chunk) return TPF_OK;
  if (b->graph) {
    TPF_HIP(hipGraphExecDestroy(b->graph));
    b->graph = nullptr;
  }
  const PcgProb* tab = b->d_tab;
  const unsigned* beg = b->d_begin;
  const int n = b->n;
  hipGraph_t graph;
  TPF_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  for (int it = 0; it < chunk; ++it) {
    hipLaunchKernelGGL(kb_update_p, dim3(b->blocks), dim3(kBlock), 0, b->stream, tab, beg, n);
    hipLaunchKernelGGL(kb_matvec, dim3(b->blocks), dim3(kBlock), 0, b->stream, tab, beg, n);
    hipLaunchKernelGGL(kb_fin_alpha, dim3(n), dim3(kBlock), 0, b->stream, tab);
    hipLaunchKernelGGL(kb_update_r, dim3(b->blocks), dim3(kBlock), 0, b->stream, tab, beg, n);
    hipLaunchKernelGGL(kb_fin_beta, dim3(n), dim3(kBlock), 0, b->stream, tab);
  }
  TPF_HIP(hipStreamEndCapture(b->stream, &graph));
  hipError_t e = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess)
    return fail(TPF_ERR_RUNTIME, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  b->graph_chunk = chunk;
  return TPF_OK;
}

}  // namespace
}  // namespace tpf

extern "C" tpf_status tpf_ppe_batch_destroy(tpf_ppe_batch* b) {
  if (!b) return TPF_OK;
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  if (b->d_tab) (void)hipFree(b->d_tab);
  if (b->h_tab) (void)hipHostFree(b->h_tab);
  if (b->d_begin) (void)hipFree(b->d_begin);
  if (b->d_flags) (void)hipFree(b->d_flags);
  if (b->h_flags) (void)hipHostFree(b->h_flags);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return TPF_OK;
}

extern "C" tpf_status tpf_ppe_batch_create(tpf_ppe_batch** out, tpf_ppe* const* handles, int n,
                                           void* stream) {
  if (!out || !handles || n < 1) return fail(TPF_ERR_VALUE, "need at least one PPE handle");
  *out = nullptr;
  auto* b = new tpf_ppe_batch();
  b->n = n;
  std::vector<unsigned> begin(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    if (!handles[i]) {
      delete b;
      return fail(TPF_ERR_VALUE, "null PPE handle in batch");
    }
    for (int j = 0; j < i; ++j)
      if (handles[j] == handles[i]) {
        delete b;
        return fail(TPF_ERR_VALUE, "a PPE handle appears twice in one batch");
      }
    b->h.push_back(handles[i]);
    begin[i + 1] = begin[i] + (unsigned)handles[i]->grid;
  }
  b->blocks = begin[n];
  b->user_stream = (hipStream_t)stream;
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&b->d_tab, sizeof(tpf::PcgProb) * n);
  if (e == hipSuccess) e = hipHostMalloc(&b->h_tab, sizeof(tpf::PcgProb) * n);
  if (e == hipSuccess) e = hipMalloc(&b->d_begin, sizeof(unsigned) * (n + 1));
  if (e == hipSuccess) e = hipMalloc(&b->d_flags, sizeof(tpf::PcgFlags) * n);
  if (e == hipSuccess) e = hipHostMalloc(&b->h_flags, sizeof(tpf::PcgFlags) * n);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_begin, begin.data(), sizeof(unsigned) * (n + 1), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    tpf_ppe_batch_destroy(b);
    return fail(TPF_ERR_RUNTIME, std::string("PPE batch allocation: ") + hipGetErrorString(e));
  }
  *out = b;
  return TPF_OK;
}

extern "C" tpf_status tpf_ppe_batch_solve(tpf_ppe_batch* b, const double* const* rhs,
                                          double* const* x, const tpf_solve_opts* opts,
                                          tpf_solve_report* reports) {
  using namespace tpf;
  if (!b || !rhs || !x || !opts) return fail(TPF_ERR_VALUE, "null argument");
  if (opts->precond != 1) return fail(TPF_ERR_VALUE, "batched CG runs Jacobi-PCG (precond 1)");
  if (opts->mixed) return fail(TPF_ERR_VALUE, "batched CG runs in FP64");
  if (!(opts->tol > 0.0) || opts->maxiter < 1) return fail(TPF_ERR_VALUE, "bad tol/maxiter");
  const int n = b->n;
  int n_on = 0;
  for (int i = 0; i < n; ++i) {
    if (!rhs[i] != !x[i]) return fail(TPF_ERR_VALUE, "rhs and x must both be set or both null");
    if (!rhs[i]) continue;  // problem i sits this solve out
    ++n_on;
    if (!b->h[i]->rho) return fail(TPF_ERR_RUNTIME, "PPE density not set for a batch member");
  }
  if (n_on == 0) {
    if (reports)
      for (int i = 0; i < n; ++i) reports[i] = tpf_solve_report{};
    return TPF_OK;
  }
  {  // order the batch stream after the caller's
    hipEvent_t ev;
    TPF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    TPF_HIP(hipEventRecord(ev, b->user_stream));
    TPF_HIP(hipStreamWaitEvent(b->stream, ev, 0));
    TPF_HIP(hipEventDestroy(ev));
  }
  std::vector<int64_t> iters(n, 0);
  std::vector<int> active(n, 0), restart(n, 0);
  for (int i = 0; i < n; ++i) active[i] = rhs[i] ? 1 : 0;
  std::vector<double> rr(n, 0.0), bb(n, 0.0), tr(n, 0.0);
  const int chunk = opts->check_every > 0 ? opts->check_every : 32;
  for (int i = 0; i < n; ++i) {
    tpf_ppe* s = b->h[i];
    PcgProb& P = b->h_tab[i];
    P.g = s->g;
    P.rho = s->rho;
    P.dinv = s->dinv;
    P.rhs = rhs[i];
    P.x = x[i];
    P.b = s->b;
    P.r = s->r;
    P.p = s->p;
    P.q = s->q;
    P.partials = s->partials;
    P.st = s->st;
    P.grid = s->grid;
    P.active = active[i];
    P.zero_x = opts->use_x0 ? 0 : 1;
    if (!active[i]) {  // a skipped problem's kernels return at once (done = 1)
      PcgState idle{};
      idle.done = 1;
      *s->st_host = idle;
      TPF_HIP(hipMemcpyAsync(s->st, s->st_host, sizeof(idle), hipMemcpyHostToDevice, b->stream));
    }
  }
  TPF_TRY(batch_upload(b));
  if (opts->use_x0) {
    hipLaunchKernelGGL(kb_pin_zero, dim3((n + 63) / 64), dim3(64), 0, b->stream,
                       (const PcgProb*)b->d_tab, n);
    TPF_HIP(hipGetLastError());
  }
  TPF_TRY(batch_graph(b, chunk));
  for (int round = 0;; ++round) {
    bool any = false;
    for (int i = 0; i < n; ++i) {
      if (!active[i]) continue;
      any = true;
      PcgState init{};
      init.tol2 = opts->tol * opts->tol;
      init.maxiter = opts->maxiter - iters[i];
      *b->h[i]->st_host = init;
      TPF_HIP(hipMemcpyAsync(b->h[i]->st, b->h[i]->st_host, sizeof(init), hipMemcpyHostToDevice,
                             b->stream));
    }
    if (!any) break;
    hipLaunchKernelGGL(kb_init, dim3(b->blocks), dim3(kBlock), 0, b->stream,
                       (const PcgProb*)b->d_tab, (const unsigned*)b->d_begin, n);
    hipLaunchKernelGGL(kb_fin_init, dim3(n), dim3(kBlock), 0, b->stream, (const PcgProb*)b->d_tab);
    TPF_HIP(hipGetLastError());
    // inactive members keep done = 1 from their last round
    TPF_TRY(batch_gather(b));
    for (;;) {
      bool all_done = true;
      for (int i = 0; i < n; ++i) all_done = all_done && b->h_flags[i].done;
      if (all_done) break;
      TPF_HIP(hipGraphLaunch(b->graph, b->stream));
      TPF_TRY(batch_gather(b));
    }
    hipLaunchKernelGGL(kb_finish_x, dim3(b->blocks), dim3(kBlock), 0, b->stream,
                       (const PcgProb*)b->d_tab, (const unsigned*)b->d_begin, n);
    hipLaunchKernelGGL(kb_true_res, dim3(b->blocks), dim3(kBlock), 0, b->stream,
                       (const PcgProb*)b->d_tab, (const unsigned*)b->d_begin, n);
    hipLaunchKernelGGL(kb_fin_true, dim3(n), dim3(kBlock), 0, b->stream, (const PcgProb*)b->d_tab);
    TPF_HIP(hipGetLastError());
    TPF_TRY(batch_gather(b));
    bool again = false;
    for (int i = 0; i < n; ++i) {
      if (!active[i]) continue;
      const PcgFlags& f = b->h_flags[i];
      iters[i] += f.iter;
      rr[i] = f.rr;
      bb[i] = f.bb;
      tr[i] = f.true_bb > 0.0 ? std::sqrt(f.true_rr / f.true_bb) : std::sqrt(f.true_rr);
      const bool rec_ok = bb[i] > 0 ? std::sqrt(rr[i] / bb[i]) <= opts->tol : true;
      // residual replacement, as tpf_ppe_solve: restart from x (at most 4 times)
      const bool more = !(tr[i] <= opts->tol || !rec_ok || iters[i] >= opts->maxiter ||
                          restart[i] >= 4);
      active[i] = more ? 1 : 0;
      if (more) {
        restart[i] += 1;
        again = true;
      }
      b->h_tab[i].active = active[i];
      b->h_tab[i].zero_x = 0;
    }
    if (!again) break;
    TPF_TRY(batch_upload(b));
  }
  std::string failed;
  for (int i = 0; i < n; ++i) {
    tpf_solve_report r{};
    if (!rhs[i]) {
      if (reports) reports[i] = r;
      continue;
    }
    r.iterations = iters[i];
    r.relres = bb[i] > 0 ? std::sqrt(rr[i] / bb[i]) : 0.0;
    r.true_relres = tr[i];
    r.converged = tr[i] <= opts->tol ? 1 : 0;
    if (reports) reports[i] = r;
    if (r.relres > opts->tol) failed += (failed.empty() ? "" : ", ") + std::to_string(i);
  }
  {  // the caller's stream waits for the batch
    hipEvent_t ev;
    TPF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    TPF_HIP(hipEventRecord(ev, b->stream));
    TPF_HIP(hipStreamWaitEvent(b->user_stream, ev, 0));
    TPF_HIP(hipEventDestroy(ev));
  }
  if (!failed.empty())
    return fail(TPF_ERR_RUNTIME, "batched PPE CG did not converge cleanly for problem(s) " +
                                     failed + " (maxiter " + std::to_string(opts->maxiter) + ")");
  return TPF_OK;
}
