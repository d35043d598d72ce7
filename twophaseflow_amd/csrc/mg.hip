// Geometric multigrid V-cycle for the CG system of the FVM PPE (FP32).
//
// No counterpart in the reference: this is the backbone of the learned
// preconditioner.  The fine operator is the reference's CG system
// (-W L) of ppe/fvm_matrixfree.py:163-221 written as a symmetric 7-point
// conductance stencil: (A x)_i = sum_f T_f (x_i - x_j) with
// T_f = (transverse control-volume area) * 2/(rho_i + rho_j) / h_a, and the
// pin node held at zero (its row and column are identity in the CG system).
//
// Hierarchy: node-centred coarsening n -> (n - 1)/2 + 1 per axis while n is
// odd and >= 5 and the pin stays on an even node.  Coarse conductances are
// homogenised: the two fine faces spanned by a coarse face are combined in
// series (harmonic) and the transverse neighbours in parallel with weights
// 1/2, 1, 1/2 per transverse axis, which reproduces the 2h rediscretisation
// for constant density and keeps high-contrast (density ratio 1000) layers
// resistive.  Restriction is full weighting (the transpose of trilinear
// interpolation), smoothing symmetric red-black Gauss-Seidel, and a
// coarsest level of <= 1024 nodes is solved in one workgroup by red-black
// sweeps in LDS (a larger stop level only gets the symmetric sweeps).
// Everything is FP32; the outer Krylov method is FP64.
//
// Layout: every level stores T[a][i] for the face between node i and its
// +1 neighbour along slot a (zero on the last plane), 1/diag, and x, b, r.
// All entry points enqueue on the handle's user stream and return without
// synchronising.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ppe_impl.hpp"
#include "reduce.hpp"

namespace tpf {
namespace {

constexpr int kMaxLevels = 12;
constexpr int kCoarseMax = 1024;  // nodes of the level solved in one block
constexpr int kMaxDefl = 64;      // learned coarse vectors
constexpr int kDeflBlocks = 1024; // dot-product blocks per vector

struct LevelGeom {
  int n[3];
  int64_t total;
  int64_t pin;  // node held at zero (-1: none)
};

// Fine-level operator evaluated from an FP32 copy of rho (face conductances
// formed on the fly: 1 array read per node instead of 3 conductances + 1/diag).
struct FineGeom {
  const float* rho;
  float h[3];    // node spacing per slot
  float ih2[3];  // 2 / h
};

struct MgLevel {
  LevelGeom g{};
  float* T[3] = {nullptr, nullptr, nullptr};
  float* dinv = nullptr;
  float* x = nullptr;
  float* b = nullptr;
  float* r = nullptr;
  float* rho32 = nullptr;  // fine level only
  FineGeom fg{};
};

__host__ __device__ inline int64_t lstride(const LevelGeom& g, int a) {
  return a == 0 ? (int64_t)g.n[1] * g.n[2] : (a == 1 ? (int64_t)g.n[2] : 1);
}

__device__ inline void lcoords(const LevelGeom& g, int64_t idx, int (&c)[3]) {
  const uint32_t plane = (uint32_t)g.n[1] * (uint32_t)g.n[2];
  const uint32_t u = (uint32_t)idx;
  c[0] = (int)(u / plane);
  const uint32_t rem = u - (uint32_t)c[0] * plane;
  c[1] = (int)(rem / (uint32_t)g.n[2]);
  c[2] = (int)(rem - (uint32_t)c[1] * (uint32_t)g.n[2]);
}

#define GRID_LOOP(idx, total)                                                  \
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < (total); \
       idx += (int64_t)gridDim.x * kBlock)


// Fine conductances from rho (FP64 in, FP32 out).
__global__ __launch_bounds__(kBlock) void k_fine_T(PpeGeom pg, LevelGeom g,
                                                   const double* __restrict__ rho,
                                                   float* __restrict__ T0, float* __restrict__ T1,
                                                   float* __restrict__ T2) {
  float* T[3] = {T0, T1, T2};
  GRID_LOOP(idx, g.total) {
    int c[3];
    lcoords(g, idx, c);
    double dv[3];
#pragma unroll
    for (int b = 0; b < 3; ++b)
      dv[b] = g.n[b] == 1 ? 1.0 : ((c[b] == 0 || c[b] == g.n[b] - 1) ? 0.5 * pg.h[b] : pg.h[b]);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (g.n[a] == 1) continue;
      float t = 0.0f;
      if (c[a] < g.n[a] - 1) {
        double area = 1.0;
#pragma unroll
        for (int b = 0; b < 3; ++b)
          if (b != a && g.n[b] > 1) area *= dv[b];
        const int64_t j = idx + lstride(g, a);
        t = (float)(area * (2.0 / (rho[idx] + rho[j])) * pg.ih[a]);
      }
      T[a][idx] = t;
    }
  }
}

// Homogenised coarse conductances.
__global__ __launch_bounds__(kBlock) void k_coarse_T(LevelGeom f, LevelGeom c,
                                                     const float* __restrict__ F0,
                                                     const float* __restrict__ F1,
                                                     const float* __restrict__ F2,
                                                     float* __restrict__ C0, float* __restrict__ C1,
                                                     float* __restrict__ C2) {
  const float* F[3] = {F0, F1, F2};
  float* Cc[3] = {C0, C1, C2};
  GRID_LOOP(idx, c.total) {
    int I[3];
    lcoords(c, idx, I);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (c.n[a] == 1) continue;
      float acc = 0.0f;
      if (I[a] < c.n[a] - 1) {
        const int b1 = a == 0 ? 1 : 0, b2 = a == 2 ? 1 : 2;
        for (int d1 = -1; d1 <= 1; ++d1) {
          if (c.n[b1] == 1 && d1 != 0) continue;
          const int j1 = 2 * I[b1] + d1;
          if (j1 < 0 || j1 >= f.n[b1]) continue;
          const float w1 = d1 == 0 ? 1.0f : 0.5f;
          for (int d2 = -1; d2 <= 1; ++d2) {
            if (c.n[b2] == 1 && d2 != 0) continue;
            const int j2 = 2 * I[b2] + d2;
            if (j2 < 0 || j2 >= f.n[b2]) continue;
            const float w2 = d2 == 0 ? 1.0f : 0.5f;
            int fc[3];
            fc[a] = 2 * I[a];
            fc[b1] = j1;
            fc[b2] = j2;
            const int64_t fi = ((int64_t)fc[0] * f.n[1] + fc[1]) * f.n[2] + fc[2];
            const float t1 = F[a][fi], t2 = F[a][fi + lstride(f, a)];
            const float s = (t1 > 0.0f && t2 > 0.0f) ? t1 * t2 / (t1 + t2) : 0.0f;
            acc += w1 * w2 * s;
          }
        }
      }
      Cc[a][idx] = acc;
    }
  }
}

// 1/diag; the pin row gets 0 so the smoother leaves x[pin] = 0.
__global__ __launch_bounds__(kBlock) void k_dinv(LevelGeom g, const float* __restrict__ T0,
                                                 const float* __restrict__ T1,
                                                 const float* __restrict__ T2,
                                                 float* __restrict__ dinv) {
  const float* T[3] = {T0, T1, T2};
  GRID_LOOP(idx, g.total) {
    int c[3];
    lcoords(g, idx, c);
    float d = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (g.n[a] == 1) continue;
      d += T[a][idx];
      if (c[a] > 0) d += T[a][idx - lstride(g, a)];
    }
    dinv[idx] = (idx == g.pin || !(d > 0.0f)) ? 0.0f : 1.0f / d;
  }
}

// sum_f T_f x_nbr and diag at node idx
__device__ inline float nbr_sum(const LevelGeom& g, const float* const (&T)[3],
                                const float* __restrict__ x, int64_t idx, const int (&c)[3],
                                float& diag) {
  float s = 0.0f;
  diag = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (g.n[a] == 1) continue;
    const int64_t st = lstride(g, a);
    if (c[a] < g.n[a] - 1) {
      const float t = T[a][idx];
      s += t * x[idx + st];
      diag += t;
    }
    if (c[a] > 0) {
      const float t = T[a][idx - st];
      s += t * x[idx - st];
      diag += t;
    }
  }
  return s;
}

// One red-black Gauss-Seidel half sweep: nodes with (c0+c1+c2) % 2 == color.
// Threads cover only the nodes of `color` (as k_rbgs_rho).
__global__ __launch_bounds__(kBlock) void k_rbgs(LevelGeom g, int color, const float* __restrict__ T0,
                                                 const float* __restrict__ T1,
                                                 const float* __restrict__ T2,
                                                 const float* __restrict__ dinv,
                                                 const float* __restrict__ b, float* x) {
  const float* const T[3] = {T0, T1, T2};
  const int half = (g.n[2] + 1) >> 1;
  const int64_t slots = (int64_t)g.n[0] * g.n[1] * half;
  CHUNK_LOOP(u, slots) {
    const uint32_t row = (uint32_t)u / (uint32_t)half;
    const int k = (int)((uint32_t)u - row * (uint32_t)half);
    int c[3];
    c[0] = (int)(row / (uint32_t)g.n[1]);
    c[1] = (int)(row - (uint32_t)c[0] * (uint32_t)g.n[1]);
    c[2] = ((color + c[0] + c[1]) & 1) + 2 * k;
    if (c[2] >= g.n[2]) continue;
    const int64_t idx = (int64_t)row * g.n[2] + c[2];
    float d;
    const float s = nbr_sum(g, T, x, idx, c, d);
    x[idx] = (b[idx] + s) * dinv[idx];
  }
}

// First forward half-sweep from x = 0 on a coarse level (fused zeroing).
__global__ __launch_bounds__(kBlock) void k_rbgs_first(LevelGeom g, const float* __restrict__ dinv,
                                                       const float* __restrict__ b,
                                                       float* __restrict__ x) {
  CHUNK_LOOP(idx, g.total) {
    int c[3];
    lcoords(g, idx, c);
    x[idx] = ((c[0] + c[1] + c[2]) & 1) == 0 ? b[idx] * dinv[idx] : 0.0f;
  }
}

// r = b - A x (r[pin] = 0)
__global__ __launch_bounds__(kBlock) void k_resid(LevelGeom g, const float* __restrict__ T0,
                                                  const float* __restrict__ T1,
                                                  const float* __restrict__ T2,
                                                  const float* __restrict__ b,
                                                  const float* __restrict__ x,
                                                  float* __restrict__ r) {
  const float* const T[3] = {T0, T1, T2};
  CHUNK_LOOP(idx, g.total) {
    int c[3];
    lcoords(g, idx, c);
    float d;
    const float s = nbr_sum(g, T, x, idx, c, d);
    r[idx] = idx == g.pin ? 0.0f : b[idx] - (d * x[idx] - s);
  }
}

// sum_f t_f x_nbr and diag from rho (fine level)
__device__ inline float nbr_sum_rho(const LevelGeom& g, const FineGeom& fg,
                                    const float* __restrict__ x, int64_t idx, const int (&c)[3],
                                    float& diag) {
  float dv[3];
#pragma unroll
  for (int b = 0; b < 3; ++b)
    dv[b] = g.n[b] == 1 ? 1.0f
                        : ((c[b] == 0 || c[b] == g.n[b] - 1) ? 0.5f * fg.h[b] : fg.h[b]);
  const float ri = fg.rho[idx];
  float s = 0.0f, d = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (g.n[a] == 1) continue;
    float w = fg.ih2[a];
#pragma unroll
    for (int b = 0; b < 3; ++b)
      if (b != a && g.n[b] > 1) w *= dv[b];
    const int64_t st = lstride(g, a);
    if (c[a] < g.n[a] - 1) {
      const float t = w * __builtin_amdgcn_rcpf(ri + fg.rho[idx + st]);
      s += t * x[idx + st];
      d += t;
    }
    if (c[a] > 0) {
      const float t = w * __builtin_amdgcn_rcpf(ri + fg.rho[idx - st]);
      s += t * x[idx - st];
      d += t;
    }
  }
  diag = d;
  return s;
}

// Threads cover only the nodes of `color`: row (c0, c1) holds the nodes
// c2 = ((color + c0 + c1) & 1) + 2k, k < (n2 + 1) / 2.
__global__ __launch_bounds__(kBlock) void k_rbgs_rho(LevelGeom g, FineGeom fg, int color,
                                                     const float* __restrict__ b, float* x) {
  const int half = (g.n[2] + 1) >> 1;
  const int64_t slots = (int64_t)g.n[0] * g.n[1] * half;
  CHUNK_LOOP(u, slots) {  // 32-bit index arithmetic (levels hold < 2^32 nodes, like lcoords)
    const uint32_t row = (uint32_t)u / (uint32_t)half;
    const int k = (int)((uint32_t)u - row * (uint32_t)half);
    int c[3];
    c[0] = (int)(row / (uint32_t)g.n[1]);
    c[1] = (int)(row - (uint32_t)c[0] * (uint32_t)g.n[1]);
    c[2] = ((color + c[0] + c[1]) & 1) + 2 * k;
    if (c[2] >= g.n[2]) continue;
    const int64_t idx = (int64_t)row * g.n[2] + c[2];
    float d;
    const float s = nbr_sum_rho(g, fg, x, idx, c, d);
    x[idx] = idx == g.pin ? 0.0f : (b[idx] + s) * __builtin_amdgcn_rcpf(d);
  }
}

// First forward half-sweep from x = 0 (fused with the zeroing): colour-0
// nodes take b / diag (their neighbours are all 0), colour-1 nodes 0.  A
// thread takes the node pair (2k, 2k+1) of a row, so the diagonal is formed
// for one node per thread without a checkerboard branch in the wavefront.
__global__ __launch_bounds__(kBlock) void k_rbgs_rho_first(LevelGeom g, FineGeom fg,
                                                           const float* __restrict__ b,
                                                           float* __restrict__ x) {
  const int half = (g.n[2] + 1) >> 1;
  const int64_t slots = (int64_t)g.n[0] * g.n[1] * half;
  CHUNK_LOOP(u, slots) {
    const uint32_t row = (uint32_t)u / (uint32_t)half;
    const int k = (int)((uint32_t)u - row * (uint32_t)half);
    int c[3];
    c[0] = (int)(row / (uint32_t)g.n[1]);
    c[1] = (int)(row - (uint32_t)c[0] * (uint32_t)g.n[1]);
    const int odd = (c[0] + c[1]) & 1;  // colour-0 node of the pair: 2k + odd
    c[2] = 2 * k + odd;
    const int64_t r0 = (int64_t)row * g.n[2];
    if (2 * k + (1 - odd) < g.n[2]) x[r0 + 2 * k + (1 - odd)] = 0.0f;
    if (c[2] >= g.n[2]) continue;
    const int64_t idx = r0 + c[2];
    float v = 0.0f;
    if (idx != g.pin) {
      float dv[3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
        dv[a] = g.n[a] == 1 ? 1.0f
                            : ((c[a] == 0 || c[a] == g.n[a] - 1) ? 0.5f * fg.h[a] : fg.h[a]);
      const float ri = fg.rho[idx];
      float d = 0.0f;
#pragma unroll
      for (int a = 0; a < 3; ++a) {  // the diagonal exactly as nbr_sum_rho forms it
        if (g.n[a] == 1) continue;
        float w = fg.ih2[a];
#pragma unroll
        for (int e = 0; e < 3; ++e)
          if (e != a && g.n[e] > 1) w *= dv[e];
        const int64_t st = lstride(g, a);
        if (c[a] < g.n[a] - 1) d += w * __builtin_amdgcn_rcpf(ri + fg.rho[idx + st]);
        if (c[a] > 0) d += w * __builtin_amdgcn_rcpf(ri + fg.rho[idx - st]);
      }
      v = b[idx] * __builtin_amdgcn_rcpf(d);
    }
    x[idx] = v;
  }
}

__global__ __launch_bounds__(kBlock) void k_resid_rho(LevelGeom g, FineGeom fg,
                                                      const float* __restrict__ b,
                                                      const float* __restrict__ x,
                                                      float* __restrict__ r) {
  CHUNK_LOOP(idx, g.total) {
    int c[3];
    lcoords(g, idx, c);
    float d;
    const float s = nbr_sum_rho(g, fg, x, idx, c, d);
    r[idx] = idx == g.pin ? 0.0f : b[idx] - (d * x[idx] - s);
  }
}

__global__ __launch_bounds__(kBlock) void k_rho32(int64_t n, const double* __restrict__ rho,
                                                  float* __restrict__ r32) {
  GRID_LOOP(i, n) r32[i] = (float)rho[i];
}

// y = A x (y[pin] = 0)
__global__ __launch_bounds__(kBlock) void k_apply(LevelGeom g, const float* __restrict__ T0,
                                                  const float* __restrict__ T1,
                                                  const float* __restrict__ T2,
                                                  const float* __restrict__ x,
                                                  float* __restrict__ y) {
  const float* const T[3] = {T0, T1, T2};
  GRID_LOOP(idx, g.total) {
    int c[3];
    lcoords(g, idx, c);
    float d;
    const float s = nbr_sum(g, T, x, idx, c, d);
    y[idx] = idx == g.pin ? 0.0f : d * x[idx] - s;
  }
}

// b_c = R r_f, full weighting (1/2 per off-centre offset and axis); b_c[pin] = 0.
__device__ inline float restrict_node(const LevelGeom& f, const LevelGeom& c, const float* rf,
                                      int64_t idx) {
  int I[3];
  lcoords(c, idx, I);
  float acc = 0.0f;
  for (int d0 = -1; d0 <= 1; ++d0) {
    const int j0 = 2 * I[0] + d0;
    if (c.n[0] == 1 ? d0 != 0 : (j0 < 0 || j0 >= f.n[0])) continue;
    const float w0 = d0 == 0 ? 1.0f : 0.5f;
    for (int d1 = -1; d1 <= 1; ++d1) {
      const int j1 = 2 * I[1] + d1;
      if (c.n[1] == 1 ? d1 != 0 : (j1 < 0 || j1 >= f.n[1])) continue;
      const float w1 = w0 * (d1 == 0 ? 1.0f : 0.5f);
      const int64_t row = ((int64_t)(c.n[0] == 1 ? 0 : j0) * f.n[1] + (c.n[1] == 1 ? 0 : j1)) *
                          f.n[2];
      for (int d2 = -1; d2 <= 1; ++d2) {
        const int j2 = 2 * I[2] + d2;
        if (j2 < 0 || j2 >= f.n[2]) continue;
        acc += w1 * (d2 == 0 ? 1.0f : 0.5f) * rf[row + j2];
      }
    }
  }
  return idx == c.pin ? 0.0f : acc;
}

__global__ __launch_bounds__(kBlock) void k_restrict(LevelGeom f, LevelGeom c,
                                                     const float* __restrict__ rf,
                                                     float* __restrict__ bc) {
  CHUNK_LOOP(idx, c.total) bc[idx] = restrict_node(f, c, rf, idx);
}

// x_f += P x_c (trilinear interpolation); the fine pin stays 0.
__device__ inline float prolong_node(const LevelGeom& f, const LevelGeom& c, const float* xc,
                                     int64_t idx) {
  int j[3];
  lcoords(f, idx, j);
  int lo[3], hi[3];
  float wl[3], wh[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = j[a] >> 1;
    if (j[a] & 1) {
      hi[a] = lo[a] + 1;
      wl[a] = 0.5f;
      wh[a] = 0.5f;
    } else {
      hi[a] = lo[a];
      wl[a] = 1.0f;
      wh[a] = 0.0f;
    }
  }
  float acc = 0.0f;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int i0 = (s & 4) ? hi[0] : lo[0];
    const int i1 = (s & 2) ? hi[1] : lo[1];
    const int i2 = (s & 1) ? hi[2] : lo[2];
    const float w = ((s & 4) ? wh[0] : wl[0]) * ((s & 2) ? wh[1] : wl[1]) *
                    ((s & 1) ? wh[2] : wl[2]);
    if (w != 0.0f) acc += w * xc[((int64_t)i0 * c.n[1] + i1) * c.n[2] + i2];
  }
  return acc;
}

__global__ __launch_bounds__(kBlock) void k_prolong_add(LevelGeom f, LevelGeom c,
                                                        const float* __restrict__ xc,
                                                        float* __restrict__ xf) {
  CHUNK_LOOP(idx, f.total) {
    if (idx == f.pin) continue;
    xf[idx] += prolong_node(f, c, xc, idx);
  }
}

__global__ __launch_bounds__(kBlock) void k_zero(int64_t n, float* __restrict__ x) {
  GRID_LOOP(i, n) x[i] = 0.0f;
}
__global__ __launch_bounds__(kBlock) void k_copy(int64_t n, const float* __restrict__ a,
                                                 float* __restrict__ b) {
  GRID_LOOP(i, n) b[i] = a[i];
}

// Coarsest level: `sweeps` symmetric red-black sweeps from x = 0, in LDS,
// one node per thread (total <= 1024).
__global__ __launch_bounds__(1024) void k_coarse_solve(LevelGeom g, const float* __restrict__ T0,
                                                       const float* __restrict__ T1,
                                                       const float* __restrict__ T2,
                                                       const float* __restrict__ dinv,
                                                       const float* __restrict__ b,
                                                       float* __restrict__ x, int sweeps) {
  __shared__ float xs[kCoarseMax];
  const int64_t idx = threadIdx.x;
  const bool live = idx < g.total;
  int c[3] = {0, 0, 0};
  float t_hi[3] = {0, 0, 0}, t_lo[3] = {0, 0, 0};
  float bi = 0.0f, di = 0.0f;
  int color = 0;
  if (live) {
    lcoords(g, idx, c);
    color = (c[0] + c[1] + c[2]) & 1;
    const float* const T[3] = {T0, T1, T2};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (g.n[a] == 1) continue;
      if (c[a] < g.n[a] - 1) t_hi[a] = T[a][idx];
      if (c[a] > 0) t_lo[a] = T[a][idx - lstride(g, a)];
    }
    bi = b[idx];
    di = dinv[idx];
    xs[idx] = 0.0f;
  }
  __syncthreads();
  for (int s = 0; s < 4 * sweeps; ++s) {
    // `sweeps` forward sweeps (red, black), then `sweeps` reverse (black, red)
    const int col = (s < 2 * sweeps) ? (s & 1) : 1 - (s & 1);
    if (live && color == col) {
      float acc = bi;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (g.n[a] == 1) continue;
        const int64_t st = lstride(g, a);
        if (c[a] < g.n[a] - 1) acc += t_hi[a] * xs[idx + st];
        if (c[a] > 0) acc += t_lo[a] * xs[idx - st];
      }
      xs[idx] = acc * di;
    }
    __syncthreads();
  }
  if (live) x[idx] = xs[idx];
}

// A whole V-visit of the last two levels in one workgroup: level f
// (<= kTwoLevelMax nodes, kTwoLevelPer per thread, coefficients held in
// registers, iterate and residual in LDS) and the coarsest level c
// (<= kCoarseMax nodes, one per thread, in LDS).  Pre-smoothing from x_f = 0
// (or from the current x_f), residual, restriction, `sweeps` symmetric
// coarse sweeps, prolongation, reverse post-smoothing; the per-node
// arithmetic is that of k_rbgs, k_resid, k_restrict, k_coarse_solve and
// k_prolong_add.  It replaces ~12 latency-bound launches by one.
constexpr int kTwoLevelPer = 5;
constexpr int kTwoLevelMax = 1024 * kTwoLevelPer;

__global__ __launch_bounds__(1024) void k_two_level(
    LevelGeom gf, const float* __restrict__ Tf0, const float* __restrict__ Tf1,
    const float* __restrict__ Tf2, const float* __restrict__ dinvf, const float* __restrict__ bf,
    float* __restrict__ xf, float* __restrict__ rf, LevelGeom gc, const float* __restrict__ Tc0,
    const float* __restrict__ Tc1, const float* __restrict__ Tc2, const float* __restrict__ dinvc,
    float* __restrict__ bc, float* __restrict__ xc, int nu, int sweeps, int from_zero) {
  __shared__ float xs[kTwoLevelMax];
  __shared__ float rs[kTwoLevelMax];
  __shared__ float xcs[kCoarseMax];
  __shared__ float bcs[kCoarseMax];
  const int tid = threadIdx.x;
  const int bd = (int)blockDim.x;  // >= coarse nodes, >= fine nodes / kTwoLevelPer
  // fine-level coefficients of this thread's nodes
  float thi[kTwoLevelPer][3], tlo[kTwoLevelPer][3], di[kTwoLevelPer], bi[kTwoLevelPer];
  int cf[kTwoLevelPer][3], colf[kTwoLevelPer];
  const float* const Tf[3] = {Tf0, Tf1, Tf2};
#pragma unroll
  for (int k = 0; k < kTwoLevelPer; ++k) {
    const int64_t idx = tid + (int64_t)k * bd;
    colf[k] = -1;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      thi[k][a] = tlo[k][a] = 0.0f;
      cf[k][a] = 0;
    }
    di[k] = bi[k] = 0.0f;
    if (idx < gf.total) {
      lcoords(gf, idx, cf[k]);
      colf[k] = (cf[k][0] + cf[k][1] + cf[k][2]) & 1;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (gf.n[a] == 1) continue;
        if (cf[k][a] < gf.n[a] - 1) thi[k][a] = Tf[a][idx];
        if (cf[k][a] > 0) tlo[k][a] = Tf[a][idx - lstride(gf, a)];
      }
      di[k] = dinvf[idx];
      bi[k] = bf[idx];
      xs[idx] = from_zero ? 0.0f : xf[idx];
    }
  }
  __syncthreads();
  // neighbour sum in nbr_sum's order
  auto nsum = [&](int k, int64_t idx, float& diag) {
    float s = 0.0f;
    diag = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (gf.n[a] == 1) continue;
      const int64_t st = lstride(gf, a);
      if (cf[k][a] < gf.n[a] - 1) {
        s += thi[k][a] * xs[idx + st];
        diag += thi[k][a];
      }
      if (cf[k][a] > 0) {
        s += tlo[k][a] * xs[idx - st];
        diag += tlo[k][a];
      }
    }
    return s;
  };
  auto smooth_f = [&](bool reverse) {
    for (int sw = 0; sw < nu; ++sw)
      for (int h = 0; h < 2; ++h) {
        const int color = reverse ? 1 - h : h;
#pragma unroll
        for (int k = 0; k < kTwoLevelPer; ++k) {
          const int64_t idx = tid + (int64_t)k * bd;
          if (colf[k] == color) {
            float d;
            const float s = nsum(k, idx, d);
            xs[idx] = (bi[k] + s) * di[k];
          }
        }
        __syncthreads();
      }
  };
  smooth_f(false);
  // residual (pin row 0)
#pragma unroll
  for (int k = 0; k < kTwoLevelPer; ++k) {
    const int64_t idx = tid + (int64_t)k * bd;
    if (colf[k] >= 0) {
      float d;
      const float s = nsum(k, idx, d);
      const float r = idx == gf.pin ? 0.0f : bi[k] - (d * xs[idx] - s);
      rs[idx] = r;
      rf[idx] = r;
    }
  }
  __syncthreads();
  // restriction and the coarsest level's symmetric sweeps from x_c = 0
  const bool live = tid < gc.total;
  int c[3] = {0, 0, 0};
  float t_hi[3] = {0, 0, 0}, t_lo[3] = {0, 0, 0};
  float bci = 0.0f, dci = 0.0f;
  int color = 0;
  if (live) {
    bci = restrict_node(gf, gc, rs, tid);
    bcs[tid] = bci;
    bc[tid] = bci;
    lcoords(gc, tid, c);
    color = (c[0] + c[1] + c[2]) & 1;
    const float* const Tc[3] = {Tc0, Tc1, Tc2};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (gc.n[a] == 1) continue;
      if (c[a] < gc.n[a] - 1) t_hi[a] = Tc[a][tid];
      if (c[a] > 0) t_lo[a] = Tc[a][tid - lstride(gc, a)];
    }
    dci = dinvc[tid];
    xcs[tid] = 0.0f;
  }
  __syncthreads();
  for (int s = 0; s < 4 * sweeps; ++s) {
    const int col = (s < 2 * sweeps) ? (s & 1) : 1 - (s & 1);
    if (live && color == col) {
      float acc = bci;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (gc.n[a] == 1) continue;
        const int64_t st = lstride(gc, a);
        if (c[a] < gc.n[a] - 1) acc += t_hi[a] * xcs[tid + st];
        if (c[a] > 0) acc += t_lo[a] * xcs[tid - st];
      }
      xcs[tid] = acc * dci;
    }
    __syncthreads();
  }
  if (live) xc[tid] = xcs[tid];
  // prolongation (fine pin stays 0) and reverse post-smoothing
#pragma unroll
  for (int k = 0; k < kTwoLevelPer; ++k) {
    const int64_t idx = tid + (int64_t)k * bd;
    if (colf[k] >= 0 && idx != gf.pin) xs[idx] += prolong_node(gf, gc, xcs, idx);
  }
  __syncthreads();
  smooth_f(true);
#pragma unroll
  for (int k = 0; k < kTwoLevelPer; ++k) {
    const int64_t idx = tid + (int64_t)k * bd;
    if (colf[k] >= 0) xf[idx] = xs[idx];
  }
}

// Learned coarse correction on level L (A-DEF2): c = W t_L (k dot products,
// block partials per vector), y = E^-1 c, u = sum_j y_j W_j.
__global__ __launch_bounds__(kBlock) void k_defl_dot(int64_t n, const float* __restrict__ W,
                                                     const float* __restrict__ t,
                                                     double* __restrict__ part) {
  __shared__ double sh[kBlock / 64];
  const int j = blockIdx.y;
  const float* w = W + (int64_t)j * n;
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    v[0] += (double)w[i] * (double)t[i];
  block_sum<1>(v, sh);
  if (threadIdx.x == 0) part[(int64_t)j * gridDim.x + blockIdx.x] = v[0];
}

// One workgroup: c_j = sum_b part[j][b] (fixed order), y = Einv c (FP64).
__global__ __launch_bounds__(kBlock) void k_defl_solve(int k, int nb, const double* __restrict__ part,
                                                       const double* __restrict__ Einv,
                                                       float* __restrict__ y) {
  __shared__ double c[kMaxDefl];
  const int j = threadIdx.x;
  if (j < k) {
    double a = 0.0;
    for (int b = 0; b < nb; ++b) a += part[(int64_t)j * nb + b];
    c[j] = a;
  }
  __syncthreads();
  if (j < k) {
    double a = 0.0;
    for (int i = 0; i < k; ++i) a += Einv[j * k + i] * c[i];
    y[j] = (float)a;
  }
}

__global__ __launch_bounds__(kBlock) void k_defl_comb(int64_t n, int k, const float* __restrict__ W,
                                                      const float* __restrict__ y,
                                                      float* __restrict__ u) {
  GRID_LOOP(i, n) {
    float a = 0.0f;
    for (int j = 0; j < k; ++j) a += y[j] * W[(int64_t)j * n + i];
    u[i] = a;
  }
}

// ---- host ------------------------------------------------------------------
struct MgState {
  int nlev = 0;
  MgLevel lv[kMaxLevels];
  const double* rho_src = nullptr;
  int max_levels = 0;
  // captured cycle (tpf::mg_cycle)
  hipGraphExec_t cyc = nullptr;
  int cyc_key[4] = {0, 0, 0, 0};
  hipStream_t cyc_stream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;  // tpf_mg_vcycle's stream ordering
  double* defl_part = nullptr;  // [kMaxDefl][blocks] dot partials
  double* defl_einv = nullptr;  // [kMaxDefl^2]
  float* defl_y = nullptr;      // [kMaxDefl]
};

MgState* mg_of(tpf_ppe* s) { return static_cast<MgState*>(s->mg); }

unsigned grid_of(int64_t total) {
  const int64_t blocks = (total + kBlock - 1) / kBlock;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(blocks, 256 * 8));
}

void free_mg(MgState* m) {
  if (m->cyc) (void)hipGraphExecDestroy(m->cyc);
  if (m->ev_in) (void)hipEventDestroy(m->ev_in);
  if (m->ev_out) (void)hipEventDestroy(m->ev_out);
  if (m->defl_part) (void)hipFree(m->defl_part);
  if (m->defl_einv) (void)hipFree(m->defl_einv);
  if (m->defl_y) (void)hipFree(m->defl_y);
  for (int l = 0; l < m->nlev; ++l) {
    MgLevel& L = m->lv[l];
    float* bufs[] = {L.T[0], L.T[1], L.T[2], L.dinv, L.x, L.b, L.r, L.rho32};
    for (float* p : bufs)
      if (p) (void)hipFree(p);
  }
  delete m;
}

#define MG_LAUNCH(kern, total, ...)                                                       \
  do {                                                                                    \
    hipLaunchKernelGGL(kern, dim3(grid_of(total)), dim3(kBlock), 0, st, __VA_ARGS__);     \
    TPF_HIP(hipGetLastError());                                                           \
  } while (0)

// `from_zero`: x is not read and starts at 0 (the zeroing is fused into
// the first forward half-sweep).
tpf_status smooth(MgLevel& L, hipStream_t st, const float* b, float* x, int sweeps,
                  bool reverse, bool from_zero = false) {
  for (int k = 0; k < sweeps; ++k)
    for (int h = 0; h < 2; ++h) {
      const int color = reverse ? 1 - h : h;
      if (from_zero && k == 0 && h == 0) {
        if (reverse) return fail(TPF_ERR_VALUE, "zero start needs a forward sweep");
        if (L.rho32)
          MG_LAUNCH(k_rbgs_rho_first, (L.g.total + 1) / 2, L.g, L.fg, b, x);
        else
          MG_LAUNCH(k_rbgs_first, L.g.total, L.g, L.dinv, b, x);
        continue;
      }
      if (L.rho32)
        MG_LAUNCH(k_rbgs_rho, (L.g.total + 1) / 2, L.g, L.fg, color, b, x);
      else
        MG_LAUNCH(k_rbgs, (L.g.total + 1) / 2, L.g, color, L.T[0], L.T[1], L.T[2], L.dinv, b, x);
    }
  return TPF_OK;
}

tpf_status coarse_solve(MgLevel& L, hipStream_t st, int sweeps) {
  // one node per thread, whole wavefronts (fewer wavefronts per barrier)
  const int threads = (int)std::min<int64_t>(1024, (L.g.total + 63) / 64 * 64);
  hipLaunchKernelGGL(k_coarse_solve, dim3(1), dim3(threads), 0, st, L.g, L.T[0], L.T[1], L.T[2],
                     L.dinv, L.b, L.x, sweeps);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

// Descend from level `from` (whose b is set) to level `to`: pre-smooth,
// residual, restrict.  x of each level is its pre-smoothed iterate.
tpf_status descend(MgState* m, hipStream_t st, int from, int to, int nu) {
  for (int l = from; l < to; ++l) {
    MgLevel& L = m->lv[l];
    MgLevel& C = m->lv[l + 1];
    if (nu == 0) MG_LAUNCH(k_zero, L.g.total, L.g.total, L.x);
    TPF_TRY(smooth(L, st, L.b, L.x, nu, false, true));
    if (L.rho32)
      MG_LAUNCH(k_resid_rho, L.g.total, L.g, L.fg, L.b, L.x, L.r);
    else
      MG_LAUNCH(k_resid, L.g.total, L.g, L.T[0], L.T[1], L.T[2], L.b, L.x, L.r);
    MG_LAUNCH(k_restrict, C.g.total, L.g, C.g, L.r, C.b);
  }
  return TPF_OK;
}

// Ascend from level `from` (whose x holds the correction) to level `to`.
tpf_status ascend(MgState* m, hipStream_t st, int from, int to, int nu) {
  for (int l = from - 1; l >= to; --l) {
    MgLevel& L = m->lv[l];
    MgLevel& C = m->lv[l + 1];
    MG_LAUNCH(k_prolong_add, L.g.total, L.g, C.g, C.x, L.x);
    TPF_TRY(smooth(L, st, L.b, L.x, nu, true));
  }
  return TPF_OK;
}

// One gamma-cycle on level l (b_l set).  A first visit starts from x_l = 0,
// a repeated visit (W-cycle) continues from the current x_l.
tpf_status cycle(MgState* m, hipStream_t st, int l, int last, int gamma, bool first, int nu_pre,
                 int nu_post, int coarse_sweeps, int w_levels = 0) {
  MgLevel& L = m->lv[l];
  if (l == last) {
    if (!first) return TPF_OK;  // same right-hand side, same coarse result
    if (L.g.total <= kCoarseMax) return coarse_solve(L, st, coarse_sweeps);
    // truncated hierarchy: symmetric smoothing only on the stop level
    if (coarse_sweeps == 0) MG_LAUNCH(k_zero, L.g.total, L.g.total, L.x);
    TPF_TRY(smooth(L, st, L.b, L.x, coarse_sweeps, false, true));
    return smooth(L, st, L.b, L.x, coarse_sweeps, true);
  }
  MgLevel& C = m->lv[l + 1];
  const int g_here = (w_levels <= 0 || l < w_levels) ? gamma : 1;
  if (l + 1 == last && g_here == 1 && !L.rho32 && nu_pre == nu_post &&
      L.g.total <= kTwoLevelMax && C.g.total <= kCoarseMax) {
    // the fewest whole wavefronts that hold the coarse level at one node and
    // the fine level at kTwoLevelPer nodes per thread: fewer wavefronts per
    // barrier on 2-D hierarchies (coarsest 17^2)
    const int64_t need = std::max<int64_t>(C.g.total, (L.g.total + kTwoLevelPer - 1) / kTwoLevelPer);
    const int threads = (int)std::min<int64_t>(1024, (need + 63) / 64 * 64);
    hipLaunchKernelGGL(k_two_level, dim3(1), dim3(threads), 0, st, L.g, L.T[0], L.T[1], L.T[2],
                       L.dinv, L.b, L.x, L.r, C.g, C.T[0], C.T[1], C.T[2], C.dinv, C.b, C.x,
                       nu_pre, coarse_sweeps, first ? 1 : 0);
    TPF_HIP(hipGetLastError());
    return TPF_OK;
  }
  if (first && nu_pre == 0) MG_LAUNCH(k_zero, L.g.total, L.g.total, L.x);
  TPF_TRY(smooth(L, st, L.b, L.x, nu_pre, false, first));
  if (L.rho32)
    MG_LAUNCH(k_resid_rho, L.g.total, L.g, L.fg, L.b, L.x, L.r);
  else
    MG_LAUNCH(k_resid, L.g.total, L.g, L.T[0], L.T[1], L.T[2], L.b, L.x, L.r);
  MG_LAUNCH(k_restrict, C.g.total, L.g, C.g, L.r, C.b);
  // gamma repeats apply on the top w_levels levels (all when w_levels <= 0)
  for (int k = 0; k < g_here; ++k)
    TPF_TRY(cycle(m, st, l + 1, last, gamma, k == 0, nu_pre, nu_post, coarse_sweeps, w_levels));
  MG_LAUNCH(k_prolong_add, L.g.total, L.g, C.g, C.x, L.x);
  return smooth(L, st, L.b, L.x, nu_post, true);
}

tpf_status check_level(tpf_ppe* s, int level) {
  MgState* m = mg_of(s);
  if (!m || m->nlev == 0) return fail(TPF_ERR_RUNTIME, "multigrid hierarchy not built");
  if (level < 0 || level >= m->nlev) return fail(TPF_ERR_VALUE, "level out of range");
  if (m->rho_src != s->rho)
    return fail(TPF_ERR_RUNTIME, "density changed since tpf_mg_setup (call it again)");
  return TPF_OK;
}

}  // namespace
}  // namespace tpf

using namespace tpf;

void tpf_mg_release(tpf_ppe* s) {
  if (s->mg) free_mg(static_cast<MgState*>(s->mg));
  s->mg = nullptr;
}

extern "C" tpf_status tpf_mg_setup(tpf_ppe* s, int max_levels, int* levels) {
  if (!s || !levels) return fail(TPF_ERR_VALUE, "null argument");
  if (!s->rho) return fail(TPF_ERR_RUNTIME, "PPE density not set (call tpf_ppe_set_density)");
  if (max_levels < 1) return fail(TPF_ERR_VALUE, "max_levels must be >= 1");
  if (s->g.total >= ((int64_t)1 << 32))
    return fail(TPF_ERR_VALUE, "multigrid levels index nodes in 32 bits (< 2^32 nodes)");
  max_levels = std::min(max_levels, kMaxLevels);
  // level sizes
  std::vector<LevelGeom> geo;
  LevelGeom g0{{s->g.n0, s->g.n1, s->g.n2}, s->g.total, s->g.pin};
  geo.push_back(g0);
  while ((int)geo.size() < max_levels) {
    const LevelGeom& f = geo.back();
    if (f.total <= kCoarseMax) break;
    bool ok = true;
    int pc[3];
    {
      int64_t p = f.pin, pl = (int64_t)f.n[1] * f.n[2];
      pc[0] = (int)(p / pl);
      pc[1] = (int)((p % pl) / f.n[2]);
      pc[2] = (int)(p % f.n[2]);
    }
    LevelGeom c{};
    for (int a = 0; a < 3; ++a) {
      if (f.n[a] == 1) {
        c.n[a] = 1;
        continue;
      }
      if (f.n[a] < 5 || (f.n[a] & 1) == 0 || (pc[a] & 1)) ok = false;
      c.n[a] = (f.n[a] - 1) / 2 + 1;
    }
    if (!ok) break;
    c.total = (int64_t)c.n[0] * c.n[1] * c.n[2];
    c.pin = ((int64_t)(pc[0] / 2) * c.n[1] + pc[1] / 2) * c.n[2] + pc[2] / 2;
    geo.push_back(c);
  }
  MgState* m = mg_of(s);
  bool rebuild = !m || m->nlev != (int)geo.size();
  if (m && rebuild) {
    free_mg(m);
    s->mg = nullptr;
    m = nullptr;
  }
  if (!m) {
    m = new MgState();
    s->mg = m;
    m->nlev = (int)geo.size();
    for (int l = 0; l < m->nlev; ++l) {
      MgLevel& L = m->lv[l];
      L.g = geo[l];
      const size_t bytes = sizeof(float) * (size_t)L.g.total;
      for (int a = 0; a < 3; ++a) TPF_HIP(hipMalloc(&L.T[a], bytes));
      TPF_HIP(hipMalloc(&L.dinv, bytes));
      TPF_HIP(hipMalloc(&L.x, bytes));
      TPF_HIP(hipMalloc(&L.b, bytes));
      TPF_HIP(hipMalloc(&L.r, bytes));
      for (int a = 0; a < 3; ++a) TPF_HIP(hipMemset(L.T[a], 0, bytes));
    }
  }
  hipStream_t st = s->user_stream;
  MgLevel& F = m->lv[0];
  MG_LAUNCH(k_fine_T, F.g.total, s->g, F.g, s->rho, F.T[0], F.T[1], F.T[2]);
  if (!F.rho32) TPF_HIP(hipMalloc(&F.rho32, sizeof(float) * (size_t)F.g.total));
  MG_LAUNCH(k_rho32, F.g.total, F.g.total, s->rho, F.rho32);
  F.fg.rho = F.rho32;
  for (int a = 0; a < 3; ++a) {
    F.fg.h[a] = (float)s->g.h[a];
    F.fg.ih2[a] = (float)(2.0 * s->g.ih[a]);
  }
  for (int l = 0; l < m->nlev; ++l) {
    MgLevel& L = m->lv[l];
    if (l > 0) {
      MgLevel& P = m->lv[l - 1];
      MG_LAUNCH(k_coarse_T, L.g.total, P.g, L.g, P.T[0], P.T[1], P.T[2], L.T[0], L.T[1], L.T[2]);
    }
    MG_LAUNCH(k_dinv, L.g.total, L.g, L.T[0], L.T[1], L.T[2], L.dinv);
  }
  TPF_HIP(hipStreamSynchronize(st));
  m->rho_src = s->rho;
  m->max_levels = max_levels;
  *levels = m->nlev;
  return TPF_OK;
}

namespace tpf {
tpf_status mg_ensure(tpf_ppe* s, int max_levels) {
  MgState* m = mg_of(s);
  if (m && m->nlev > 0 && m->rho_src == s->rho && m->max_levels == max_levels) return TPF_OK;
  int lv = 0;
  return tpf_mg_setup(s, max_levels, &lv);
}

void mg_invalidate(tpf_ppe* s) {
  if (MgState* m = mg_of(s)) m->rho_src = nullptr;
}

float* mg_fine_b(tpf_ppe* s) { return mg_of(s)->lv[0].b; }
float* mg_fine_x(tpf_ppe* s) { return mg_of(s)->lv[0].x; }

tpf_status mg_cycle(tpf_ppe* s, hipStream_t st, int nu, int gamma, int coarse_sweeps,
                    int w_levels) {
  MgState* m = mg_of(s);
  if (!m || m->nlev == 0) return fail(TPF_ERR_RUNTIME, "multigrid hierarchy not built");
  const int key[4] = {nu, gamma, coarse_sweeps, w_levels};
  if (!(m->cyc && m->cyc_stream == st && key[0] == m->cyc_key[0] && key[1] == m->cyc_key[1] &&
        key[2] == m->cyc_key[2] && key[3] == m->cyc_key[3])) {
    if (m->cyc) {
      TPF_HIP(hipGraphExecDestroy(m->cyc));
      m->cyc = nullptr;
    }
    hipGraph_t graph;
    TPF_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const tpf_status cs = cycle(m, st, 0, m->nlev - 1, gamma, true, nu, nu, coarse_sweeps,
                                w_levels);
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    if (cs != TPF_OK) return cs;
    if (ec != hipSuccess) return fail(TPF_ERR_RUNTIME, std::string("capture: ") + hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&m->cyc, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess)
      return fail(TPF_ERR_RUNTIME, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    for (int i = 0; i < 4; ++i) m->cyc_key[i] = key[i];
    m->cyc_stream = st;
  }
  TPF_HIP(hipGraphLaunch(m->cyc, st));
  return TPF_OK;
}

tpf_status mg_deflate_setup(tpf_ppe* s, int k, int level, const double* einv_host) {
  MgState* m = mg_of(s);
  if (!m || m->nlev == 0) return fail(TPF_ERR_RUNTIME, "multigrid hierarchy not built");
  if (k < 1 || k > kMaxDefl) return fail(TPF_ERR_VALUE, "deflation: 1 <= k <= 64 vectors");
  if (level < 1 || level >= m->nlev) return fail(TPF_ERR_VALUE, "deflation level out of range");
  if (!m->defl_part) {
    TPF_HIP(hipMalloc(&m->defl_part, sizeof(double) * kMaxDefl * kDeflBlocks));
    TPF_HIP(hipMalloc(&m->defl_einv, sizeof(double) * kMaxDefl * kMaxDefl));
    TPF_HIP(hipMalloc(&m->defl_y, sizeof(float) * kMaxDefl));
  }
  TPF_HIP(hipMemcpy(m->defl_einv, einv_host, sizeof(double) * k * k, hipMemcpyHostToDevice));
  return TPF_OK;
}

tpf_status mg_deflate(tpf_ppe* s, hipStream_t st, const float* W, int k, int level) {
  MgState* m = mg_of(s);
  MgLevel& F = m->lv[0];
  // t = r - A z on the fine level (level-0 b holds r, x holds z = V(r))
  if (F.rho32)
    MG_LAUNCH(k_resid_rho, F.g.total, F.g, F.fg, (const float*)F.b, (const float*)F.x, F.r);
  else
    MG_LAUNCH(k_resid, F.g.total, F.g, F.T[0], F.T[1], F.T[2], (const float*)F.b,
              (const float*)F.x, F.r);
  for (int l = 0; l < level; ++l)
    MG_LAUNCH(k_restrict, m->lv[l + 1].g.total, m->lv[l].g, m->lv[l + 1].g,
              (const float*)m->lv[l].r, m->lv[l + 1].r);
  MgLevel& C = m->lv[level];
  const int64_t nc = C.g.total;
  const int nb = (int)std::min<int64_t>((nc + kBlock - 1) / kBlock, kDeflBlocks);
  hipLaunchKernelGGL(k_defl_dot, dim3(nb, k), dim3(kBlock), 0, st, nc, W, (const float*)C.r,
                     m->defl_part);
  hipLaunchKernelGGL(k_defl_solve, dim3(1), dim3(kBlock), 0, st, k, nb,
                     (const double*)m->defl_part, (const double*)m->defl_einv, m->defl_y);
  MG_LAUNCH(k_defl_comb, nc, nc, k, W, (const float*)m->defl_y, C.x);
  // z += P ... P u
  for (int l = level - 1; l >= 0; --l) {
    MgLevel& L = m->lv[l];
    if (l > 0) MG_LAUNCH(k_zero, L.g.total, L.g.total, L.x);
    MG_LAUNCH(k_prolong_add, L.g.total, L.g, m->lv[l + 1].g, (const float*)m->lv[l + 1].x, L.x);
  }
  return TPF_OK;
}

tpf_status mg_level_total(tpf_ppe* s, int level, int64_t* total) {
  MgState* m = mg_of(s);
  if (!m || level < 0 || level >= m->nlev) return fail(TPF_ERR_VALUE, "level out of range");
  *total = m->lv[level].g.total;
  return TPF_OK;
}
}  // namespace tpf

extern "C" tpf_status tpf_mg_level_info(tpf_ppe* s, int level, int64_t* n3, int64_t* pin) {
  MgState* m = mg_of(s);
  if (!m || level < 0 || level >= m->nlev) return fail(TPF_ERR_VALUE, "level out of range");
  for (int a = 0; a < 3; ++a) n3[a] = m->lv[level].g.n[a];
  if (pin) *pin = m->lv[level].g.pin;
  return TPF_OK;
}

extern "C" tpf_status tpf_mg_vcycle(tpf_ppe* s, const float* r, float* z, int nu_pre, int nu_post,
                                    int coarse_sweeps, int stop_level, int gamma) {
  if (!s || !r || !z) return fail(TPF_ERR_VALUE, "null argument");
  if (gamma < 1 || gamma > 2) return fail(TPF_ERR_VALUE, "gamma must be 1 (V) or 2 (W)");
  TPF_TRY(check_level(s, 0));
  MgState* m = mg_of(s);
  const int last = (stop_level < 0 || stop_level >= m->nlev) ? m->nlev - 1 : stop_level;
  MgLevel& F = m->lv[0];
  if (last == m->nlev - 1 && nu_pre == nu_post && s->stream) {
    // full cycle: replay the captured graph on the handle's stream, ordered
    // after and before the caller's stream (~150 launches become one)
    if (!m->ev_in) {
      TPF_HIP(hipEventCreateWithFlags(&m->ev_in, hipEventDisableTiming));
      TPF_HIP(hipEventCreateWithFlags(&m->ev_out, hipEventDisableTiming));
    }
    hipStream_t st = s->stream;
    TPF_HIP(hipEventRecord(m->ev_in, s->user_stream));
    TPF_HIP(hipStreamWaitEvent(st, m->ev_in, 0));
    MG_LAUNCH(k_copy, F.g.total, F.g.total, r, F.b);
    TPF_TRY(mg_cycle(s, st, nu_pre, gamma, coarse_sweeps, 0));
    MG_LAUNCH(k_copy, F.g.total, F.g.total, F.x, z);
    TPF_HIP(hipEventRecord(m->ev_out, st));
    TPF_HIP(hipStreamWaitEvent(s->user_stream, m->ev_out, 0));
    return TPF_OK;
  }
  hipStream_t st = s->user_stream;
  MG_LAUNCH(k_copy, F.g.total, F.g.total, r, F.b);
  TPF_TRY(cycle(m, st, 0, last, gamma, true, nu_pre, nu_post, coarse_sweeps));
  MG_LAUNCH(k_copy, F.g.total, F.g.total, F.x, z);
  return TPF_OK;
}

extern "C" tpf_status tpf_mg_down(tpf_ppe* s, const float* r, int level, int nu_pre, float* bc) {
  if (!s || !r || !bc) return fail(TPF_ERR_VALUE, "null argument");
  TPF_TRY(check_level(s, level));
  MgState* m = mg_of(s);
  hipStream_t st = s->user_stream;
  MgLevel& F = m->lv[0];
  MG_LAUNCH(k_copy, F.g.total, F.g.total, r, F.b);
  TPF_TRY(descend(m, st, 0, level, nu_pre));
  MgLevel& C = m->lv[level];
  MG_LAUNCH(k_copy, C.g.total, C.g.total, (const float*)C.b, bc);
  return TPF_OK;
}

extern "C" tpf_status tpf_mg_up(tpf_ppe* s, int level, const float* xc, int nu_post, float* z) {
  if (!s || !xc || !z) return fail(TPF_ERR_VALUE, "null argument");
  TPF_TRY(check_level(s, level));
  MgState* m = mg_of(s);
  hipStream_t st = s->user_stream;
  MgLevel& C = m->lv[level];
  MG_LAUNCH(k_copy, C.g.total, C.g.total, xc, C.x);
  TPF_TRY(ascend(m, st, level, 0, nu_post));
  MgLevel& F = m->lv[0];
  MG_LAUNCH(k_copy, F.g.total, F.g.total, (const float*)F.x, z);
  return TPF_OK;
}

extern "C" tpf_status tpf_mg_prolong(tpf_ppe* s, int level, const float* xc, float* z) {
  if (!s || !xc || !z) return fail(TPF_ERR_VALUE, "null argument");
  TPF_TRY(check_level(s, level));
  MgState* m = mg_of(s);
  hipStream_t st = s->user_stream;
  MgLevel& C = m->lv[level];
  MG_LAUNCH(k_copy, C.g.total, C.g.total, xc, C.x);
  for (int l = level - 1; l >= 0; --l) {
    MgLevel& L = m->lv[l];
    MG_LAUNCH(k_zero, L.g.total, L.g.total, L.x);
    MG_LAUNCH(k_prolong_add, L.g.total, L.g, m->lv[l + 1].g, (const float*)m->lv[l + 1].x, L.x);
  }
  MG_LAUNCH(k_copy, m->lv[0].g.total, m->lv[0].g.total, (const float*)m->lv[0].x, z);
  return TPF_OK;
}

extern "C" tpf_status tpf_mg_restrict(tpf_ppe* s, const float* r, int level, float* bc) {
  if (!s || !r || !bc) return fail(TPF_ERR_VALUE, "null argument");
  TPF_TRY(check_level(s, level));
  MgState* m = mg_of(s);
  hipStream_t st = s->user_stream;
  MG_LAUNCH(k_copy, m->lv[0].g.total, m->lv[0].g.total, r, m->lv[0].r);
  for (int l = 0; l < level; ++l) {
    MgLevel& L = m->lv[l];
    MG_LAUNCH(k_restrict, m->lv[l + 1].g.total, L.g, m->lv[l + 1].g, (const float*)L.r,
              m->lv[l + 1].r);
  }
  MgLevel& C = m->lv[level];
  MG_LAUNCH(k_copy, C.g.total, C.g.total, (const float*)C.r, bc);
  return TPF_OK;
}

extern "C" tpf_status tpf_mg_apply(tpf_ppe* s, int level, const float* x, float* y) {
  if (!s || !x || !y) return fail(TPF_ERR_VALUE, "null argument");
  TPF_TRY(check_level(s, level));
  MgLevel& L = mg_of(s)->lv[level];
  hipStream_t st = s->user_stream;
  MG_LAUNCH(k_apply, L.g.total, L.g, L.T[0], L.T[1], L.T[2], x, y);
  return TPF_OK;
}

extern "C" tpf_status tpf_mg_smooth(tpf_ppe* s, int level, const float* b, float* x, int sweeps) {
  if (!s || !b || !x) return fail(TPF_ERR_VALUE, "null argument");
  TPF_TRY(check_level(s, level));
  MgLevel& L = mg_of(s)->lv[level];
  hipStream_t st = s->user_stream;
  TPF_TRY(smooth(L, st, b, x, sweeps, false));
  TPF_TRY(smooth(L, st, b, x, sweeps, true));
  return TPF_OK;
}

extern "C" tpf_status tpf_mg_coeffs(tpf_ppe* s, int level, float* out) {
  if (!s || !out) return fail(TPF_ERR_VALUE, "null argument");
  TPF_TRY(check_level(s, level));
  MgLevel& L = mg_of(s)->lv[level];
  const size_t n = (size_t)L.g.total;
  hipStream_t st = s->user_stream;
  for (int a = 0; a < 3; ++a)
    TPF_HIP(hipMemcpyAsync(out + a * n, L.T[a], sizeof(float) * n, hipMemcpyDeviceToDevice, st));
  TPF_HIP(hipMemcpyAsync(out + 3 * n, L.dinv, sizeof(float) * n, hipMemcpyDeviceToDevice, st));
  return TPF_OK;
}
