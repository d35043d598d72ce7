// CCD line-solve kernels for gfx950 and the C-ABI entry points that launch
// them (replace ccd/ccd_solver_helpers.py:38-77 and :426-465).
//
// Kernels (one line = the 2x2-block tridiagonal system of one grid line, cut
// into P partitions of <= M rows and coupled through precomputed reduced
// (SPIKE) rows):
//  * ccd_line_kernel<M>: 64 lines x P partitions per workgroup, lane = line,
//    wavefront = partition (tables wavefront-uniform, staged in LDS).  A lane
//    keeps its partition (<= M rows, 2 FP64 unknowns per row) in registers:
//    RHS -> local block LU forward/back sweep -> partition ends to LDS ->
//    coupling -> spike correction -> store.  The contiguous (last) axis is
//    staged through LDS as one coalesced [64][nax] tile.
//  * ccd_explicit_kernel<M>: strided axes, same mapping, the local solve as
//    the precomputed partition inverse applied to first differences (no
//    recurrence state: 64 VGPRs, two workgroups per CU).
//  * ccd_wave_kernel<M>: few long lines; lane = partition, wavefront = line.
//  * ccd_line_kernel<M, true>: FP32 solves with one FP64 refinement.
//  * batch kernels: the same per-line code over a device table of slots.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "ccd_plan.hpp"
#include "../../include/tpf_testing.h"

namespace tpf {

namespace {

// Wavefront-uniform table reads through the scalar cache.
typedef const __attribute__((address_space(4))) double* kdptr;
typedef const __attribute__((address_space(4))) float* kfptr;

constexpr int kStageRows = 16;              // rows per LDS re-staging chunk
constexpr int kStagePitch = kStageRows + 1; // padded line pitch (doubles)

// One axis slot of a launch: the axis geometry, its line addressing and the
// arrays it reads and writes.  A launch covers up to 3 slots (all axes of
// tpf_ccd_apply in one grid); workgroups [block_begin[s], block_begin[s+1])
// belong to slot s.
struct CcdSlot {
  CcdAxisGeom g;
  CcdLines ln;
  const double* f;
  double* d1;
  double* d2;
  int contig;  // line nodes are contiguous (last axis)
  int tiled;   // contiguous lines staged through LDS as one coalesced tile
};
struct CcdLaunch {
  int nslots;
  unsigned block_begin[4];
  CcdSlot slot[3];
};

constexpr int kTilePer = 20;  // tile doubles per thread when staging a contiguous-axis tile

// Copies n table values to LDS, dst[map(i)] = src[i]: each thread issues 8
// global loads before their LDS stores, so a staging loop waits on one
// round trip per 8 values instead of one per value.
template <typename T, typename Map>
__device__ __forceinline__ void stage_table(T* dst, const T* __restrict__ src, int n, Map map) {
  const int bd = (int)blockDim.x;
  for (int i0 = (int)threadIdx.x; i0 < n; i0 += 8 * bd) {
    T v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = i0 + k * bd < n ? src[i0 + k * bd] : T(0);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (i0 + k * bd < n) dst[map(i0 + k * bd)] = v[k];
  }
}
__device__ __forceinline__ int same_index(int i) { return i; }

template <int M>
__device__ __forceinline__ void ccd_lines(const CcdAxisGeom& g, const CcdLines& ln,
                                          const bool CONTIG, const bool TILED,
                                          const double* __restrict__ f,
                                          double* __restrict__ d1, double* __restrict__ d2,
                                          const unsigned block) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x & 63;
  const int p = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t b = (int64_t)block * 64 + lane;
  const bool active = b < ln.n_lines;
  int64_t base = 0;
  if (active) {
    if (CONTIG) {
      base = b * ln.inner_stride;
    } else {
      const uint32_t bi = (uint32_t)b, in = (uint32_t)ln.inner;
      base = (int64_t)(bi / in) * ln.outer + (int64_t)(bi % in) * ln.inner_stride;
    }
  }
  const int64_t st = ln.stride;
  const int s0 = g.s[p];
  const int m = g.s[p + 1] - s0;
  const bool periodic = g.periodic != 0;
  const bool wall_first = !periodic && p == 0;
  const bool wall_last = !periodic && p == g.P - 1;

  // ---- load f: fv[k] = node (s0 - 1 + k) periodic (wrapped), (s0 + k) wall
  double fv[M + 2];
  const int j0 = periodic ? s0 - 1 : s0;
  if (M == 16 && TILED) {
    // Contiguous axis: the workgroup's 64 lines are one contiguous block of
    // memory; stage it through LDS with coalesced loads, then each lane reads
    // its partition of its own line.  (Host guarantees 64*nax <= kTilePer*blockDim.)
    const int nax = (int)ln.inner_stride;
    const int64_t rem = ln.n_lines - (int64_t)block * 64;
    const int cnt = (int)(rem < 64 ? rem : 64) * nax;
    const double* src = f + (int64_t)block * 64 * nax;
    double tmp[kTilePer];
#pragma unroll
    for (int i = 0; i < kTilePer; ++i) {
      const int e = threadIdx.x + i * (int)blockDim.x;
      tmp[i] = e < cnt ? src[e] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < kTilePer; ++i) {
      const int e = threadIdx.x + i * (int)blockDim.x;
      if (e < cnt) lds[e] = tmp[i];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < M + 2; ++k) {
      fv[k] = 0.0;
      if (k < m + 2) {
        int j = j0 + k;
        if (periodic) j = j < 0 ? j + g.N : (j >= g.N ? j - g.N : j);
        if (active) fv[k] = lds[lane * nax + j];
      }
    }
    __syncthreads();  // the tile's LDS is reused below
  } else {
#pragma unroll
    for (int k = 0; k < M + 2; ++k) {
      fv[k] = 0.0;
      if (k < m + 2) {
        int j = j0 + k;
        if (periodic) j = j < 0 ? j + g.N : (j >= g.N ? j - g.N : j);
        if (active) fv[k] = f[base + (int64_t)j * st];
      }
    }
  }

  // ---- per-row LU/spike tables and coupling rows -> LDS (all WGs share them;
  // reading them from LDS keeps scalar-cache misses off the dependent sweeps).
  // A staged contiguous-axis tile has been consumed by now; its space is reused.
  // Uniform partitions (periodic, equal sizes) share partition 0's rows:
  // only those are staged.
  const bool kLdsTab = M <= 32;
  const int NQ = g.NQ;
  const int np = g.uniform ? 1 : g.P;  // partitions with their own tables
  const int pt = g.uniform ? 0 : p;    // this wavefront's table partition
  double* ltab = lds;
  double* lq = lds + (kLdsTab ? np * M * kRowTab : 0);
  double* G = lq + (kLdsTab ? np * 16 * NQ : 0);  // [4P][64]
  if (kLdsTab) {
    const int nt = np * M * kRowTab, nq = np * 16 * NQ;
    stage_table(ltab, g.rowtab, nt, same_index);
    stage_table(lq, g.qtab, nq, same_index);
    __syncthreads();
  }

  // ---- wall closure right-hand sides (ccd_boundary_helpers.py:64-116)
  double cl0 = 0.0, cl1 = 0.0;  // L . bc_lo  for the first interior row
  double bl0 = 0.0, bl1 = 0.0;  // bc_lo itself (node-0 output)
  if (wall_first) {
#pragma unroll
    for (int k = 0; k < 4; ++k) bl0 += g.cIl[k] * fv[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) bl1 += g.cIIl[k] * fv[k];
    cl0 = kAlpha1 * bl0 + kB1 * bl1;
    cl1 = kB2 * bl0 + kBeta2 * bl1;
  }
  double cr0 = 0.0, cr1 = 0.0, br0 = 0.0, br1 = 0.0;
  if (wall_last) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double v = active ? f[base + (int64_t)(g.N - k) * st] : 0.0;
      if (k < 4) br0 += g.cIr[k] * v;
      br1 += g.cIIr[k] * v;
    }
    cr0 = kAlpha1 * br0 - kB1 * br1;
    cr1 = -kB2 * br0 + kBeta2 * br1;
  }

  // ---- RHS + forward sweep: y_r = R_r - Lf_r y_{r-1}
  const double* tab = (kLdsTab ? ltab : g.rowtab) + (size_t)pt * M * kRowTab;
  double x[M][2];
#pragma unroll
  for (int r = 0; r < M; ++r) {
    if (r < m) {
      double r0 = kA1 * (fv[r + 2] - fv[r]);
      double r1 = kA2 * ((fv[r] - 2.0 * fv[r + 1]) + fv[r + 2]);
      if (r == 0) {
        r0 -= cl0;
        r1 -= cl1;
      }
      if (r == m - 1) {
        r0 -= cr0;
        r1 -= cr1;
      }
      if (r > 0) {
        const double* t = tab + r * kRowTab;
        r0 -= t[0] * x[r - 1][0] + t[1] * x[r - 1][1];
        r1 -= t[2] * x[r - 1][0] + t[3] * x[r - 1][1];
      }
      x[r][0] = r0;
      x[r][1] = r1;
    }
  }
  // ---- backward sweep: x_r = iD_r (y_r - U x_{r+1}); partition ends -> LDS
#pragma unroll
  for (int r = M - 1; r >= 0; --r) {
    if (r < m) {
      double a = x[r][0], c = x[r][1];
      if (r + 1 < M && r < m - 1) {
        a -= kAlpha1 * x[r + 1][0] - kB1 * x[r + 1][1];
        c -= -kB2 * x[r + 1][0] + kBeta2 * x[r + 1][1];
      }
      const double* t = tab + r * kRowTab + 4;
      x[r][0] = t[0] * a + t[1] * c;
      x[r][1] = t[2] * a + t[3] * c;
      if (r == m - 1) {
        G[(p * 4 + 2) * 64 + lane] = x[r][0];
        G[(p * 4 + 3) * 64 + lane] = x[r][1];
      }
    }
  }
  G[(p * 4 + 0) * 64 + lane] = x[0][0];
  G[(p * 4 + 1) * 64 + lane] = x[0][1];
  __syncthreads();

  // ---- coupling: (left, right) end values of the neighbours = Q_p . G
  double lr0 = 0.0, lr1 = 0.0, lr2 = 0.0, lr3 = 0.0;
  {
    const double* q = (kLdsTab ? lq : g.qtab) + (size_t)pt * 4 * NQ * 4;
    for (int t = 0; t < NQ; ++t) {
      int qq = g.rel ? p + g.d0 + t : t;
      if (periodic) qq = qq < 0 ? qq + g.P : (qq >= g.P ? qq - g.P : qq);
      else qq = qq < 0 ? 0 : (qq >= g.P ? g.P - 1 : qq);
      const double G0 = G[(qq * 4 + 0) * 64 + lane];
      const double G1 = G[(qq * 4 + 1) * 64 + lane];
      const double G2 = G[(qq * 4 + 2) * 64 + lane];
      const double G3 = G[(qq * 4 + 3) * 64 + lane];
      const double* q0 = q + (0 * NQ + t) * 4;
      const double* q1 = q + (1 * NQ + t) * 4;
      const double* q2 = q + (2 * NQ + t) * 4;
      const double* q3 = q + (3 * NQ + t) * 4;
      lr0 += q0[0] * G0 + q0[1] * G1 + q0[2] * G2 + q0[3] * G3;
      lr1 += q1[0] * G0 + q1[1] * G1 + q1[2] * G2 + q1[3] * G3;
      lr2 += q2[0] * G0 + q2[1] * G1 + q2[2] * G2 + q2[3] * G3;
      lr3 += q3[0] * G0 + q3[1] * G1 + q3[2] * G2 + q3[3] * G3;
    }
  }

  // ---- spike correction and scaling back to f', f''
  double xl0 = 0.0, xl1 = 0.0;  // last row after correction (wall node N)
#pragma unroll
  for (int r = 0; r < M; ++r) {
    if (r < m) {
      const double* t = tab + r * kRowTab + 8;
      const double u = x[r][0] - (t[0] * lr0 + t[1] * lr1 + t[4] * lr2 + t[5] * lr3);
      const double v = x[r][1] - (t[2] * lr0 + t[3] * lr1 + t[6] * lr2 + t[7] * lr3);
      x[r][0] = u;
      x[r][1] = v;
      if (r == m - 1) {
        xl0 = u;
        xl1 = v;
      }
    }
  }
  const int off = periodic ? 0 : 1;  // unknown r -> node s0 + r + off

  // boundary / image nodes (one per line); a staged tile takes them in LDS below
  if (active && !(M == 16 && TILED)) {
    if (periodic && p == 0) {
      if (d1) d1[base + (int64_t)g.N * st] = x[0][0] * g.inv_h;
      if (d2) d2[base + (int64_t)g.N * st] = x[0][1] * g.inv_h2;
    }
    if (wall_first) {
      const double u0 = g.Ml[0] * x[0][0] + g.Ml[1] * x[0][1] + bl0;
      const double v0 = g.Ml[2] * x[0][0] + g.Ml[3] * x[0][1] + bl1;
      if (d1) d1[base] = u0 * g.inv_h;
      if (d2) d2[base] = v0 * g.inv_h2;
    }
    if (wall_last) {
      const double uN = g.Mr[0] * xl0 + g.Mr[1] * xl1 + br0;
      const double vN = g.Mr[2] * xl0 + g.Mr[3] * xl1 + br1;
      if (d1) d1[base + (int64_t)g.N * st] = uN * g.inv_h;
      if (d2) d2[base + (int64_t)g.N * st] = vN * g.inv_h2;
    }
  }

  if (!CONTIG) {
    if (active) {
#pragma unroll
      for (int r = 0; r < M; ++r) {
        if (r < m) {
          const int64_t a = base + (int64_t)(s0 + r + off) * st;
          // streaming outputs: non-temporal stores keep f resident in the caches
          if (d1) __builtin_nontemporal_store(x[r][0] * g.inv_h, d1 + a);
          if (d2) __builtin_nontemporal_store(x[r][1] * g.inv_h2, d2 + a);
        }
      }
    }
  } else if (M == 16 && TILED) {
    // The workgroup's 64 lines are one contiguous, 128 B-aligned block of
    // 64 x nax doubles: assemble each output component as that tile in LDS
    // (interior rows, closure / image nodes) and store it linearly, so every
    // store instruction writes whole aligned cache lines.
    const int nax = (int)ln.inner_stride;
    const int64_t rem = ln.n_lines - (int64_t)block * 64;
    const int cnt = (int)(rem < 64 ? rem : 64) * nax;
    __syncthreads();  // G no longer read; the tile space is reused
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
      double* dst = comp == 0 ? d1 : d2;
      const double scale = comp == 0 ? g.inv_h : g.inv_h2;
      if (dst == nullptr) continue;  // uniform
      double* row = lds + lane * nax;
#pragma unroll
      for (int r = 0; r < M; ++r)
        if (r < m) row[s0 + r + off] = x[r][comp] * scale;
      if (periodic && p == 0) row[g.N] = x[0][comp] * scale;
      if (wall_first)
        row[0] = (comp == 0 ? g.Ml[0] * x[0][0] + g.Ml[1] * x[0][1] + bl0
                            : g.Ml[2] * x[0][0] + g.Ml[3] * x[0][1] + bl1) * scale;
      if (wall_last)
        row[g.N] = (comp == 0 ? g.Mr[0] * xl0 + g.Mr[1] * xl1 + br0
                              : g.Mr[2] * xl0 + g.Mr[3] * xl1 + br1) * scale;
      __syncthreads();
      double* out = dst + (int64_t)block * 64 * nax;
#pragma unroll
      for (int i = 0; i < kTilePer; ++i) {
        const int e = threadIdx.x + i * (int)blockDim.x;
        if (e < cnt) __builtin_nontemporal_store(lds[e], out + e);
      }
      __syncthreads();
    }
  } else {
    // Re-stage [64 lines][16 rows] per wavefront so a store instruction covers
    // 4 lines x 16 consecutive nodes.
    __syncthreads();  // G no longer read; its space is reused
    double* S = lds + (size_t)p * 64 * kStagePitch;
    const int64_t line0 = (int64_t)block * 64;
#pragma unroll
    for (int c0 = 0; c0 < M; c0 += kStageRows) {
#pragma unroll
      for (int comp = 0; comp < 2; ++comp) {
        double* dst = comp == 0 ? d1 : d2;
        const double scale = comp == 0 ? g.inv_h : g.inv_h2;
        if (dst == nullptr) continue;  // uniform
#pragma unroll
        for (int r = 0; r < kStageRows; ++r)
          if (c0 + r < M) S[lane * kStagePitch + r] = x[c0 + r][comp] * scale;
        __syncthreads();
#pragma unroll
        for (int it = 0; it < kStageRows; ++it) {
          const int e = lane + 64 * it;
          const int l = e >> 4, r = e & 15;
          const int64_t bl = line0 + l;
          if (c0 + r < m && bl < ln.n_lines)
            __builtin_nontemporal_store(S[l * kStagePitch + r],
                                        dst + bl * ln.inner_stride + (int64_t)(s0 + c0 + r + off));
        }
        __syncthreads();
      }
    }
  }
}

// ---- FP32 solve + FP64 refinement (lane-per-line mapping, M <= 16) --------
// One partition solve in FP32 on the RHS held in v: block-LU sweeps with the
// FP32 tables, partition ends -> G ([4P][64]), barrier, coupling through
// the reduced-system rows, spike correction.  Every thread of the block
// calls it (it contains a barrier).
template <int M>
__device__ __forceinline__ void solve32(const CcdAxisGeom& g, const float* __restrict__ tab,
                                        const float* __restrict__ qt, float (&v)[M][2], int m,
                                        int p, int lane, bool periodic, float* G) {
#pragma unroll
  for (int r = 1; r < M; ++r) {
    if (r < m) {
      const float* t = tab + r * kRowTab;
      v[r][0] -= t[0] * v[r - 1][0] + t[1] * v[r - 1][1];
      v[r][1] -= t[2] * v[r - 1][0] + t[3] * v[r - 1][1];
    }
  }
#pragma unroll
  for (int r = M - 1; r >= 0; --r) {
    if (r < m) {
      float a = v[r][0], c = v[r][1];
      if (r + 1 < M && r < m - 1) {
        a -= (float)kAlpha1 * v[r + 1][0] - (float)kB1 * v[r + 1][1];
        c -= -(float)kB2 * v[r + 1][0] + (float)kBeta2 * v[r + 1][1];
      }
      const float* t = tab + r * kRowTab + 4;
      v[r][0] = t[0] * a + t[1] * c;
      v[r][1] = t[2] * a + t[3] * c;
      if (r == m - 1) {
        G[(p * 4 + 2) * 64 + lane] = v[r][0];
        G[(p * 4 + 3) * 64 + lane] = v[r][1];
      }
    }
  }
  G[(p * 4 + 0) * 64 + lane] = v[0][0];
  G[(p * 4 + 1) * 64 + lane] = v[0][1];
  __syncthreads();
  float lr0 = 0.f, lr1 = 0.f, lr2 = 0.f, lr3 = 0.f;
  const int NQ = g.NQ;
  const float* q = qt + (size_t)p * 4 * NQ * 4;
  for (int t = 0; t < NQ; ++t) {
    int qq = g.rel ? p + g.d0 + t : t;
    if (periodic) qq = qq < 0 ? qq + g.P : (qq >= g.P ? qq - g.P : qq);
    else qq = qq < 0 ? 0 : (qq >= g.P ? g.P - 1 : qq);
    const float G0 = G[(qq * 4 + 0) * 64 + lane], G1 = G[(qq * 4 + 1) * 64 + lane];
    const float G2 = G[(qq * 4 + 2) * 64 + lane], G3 = G[(qq * 4 + 3) * 64 + lane];
    const float* q0 = q + (0 * NQ + t) * 4;
    const float* q1 = q + (1 * NQ + t) * 4;
    const float* q2 = q + (2 * NQ + t) * 4;
    const float* q3 = q + (3 * NQ + t) * 4;
    lr0 += q0[0] * G0 + q0[1] * G1 + q0[2] * G2 + q0[3] * G3;
    lr1 += q1[0] * G0 + q1[1] * G1 + q1[2] * G2 + q1[3] * G3;
    lr2 += q2[0] * G0 + q2[1] * G1 + q2[2] * G2 + q2[3] * G3;
    lr3 += q3[0] * G0 + q3[1] * G1 + q3[2] * G2 + q3[3] * G3;
  }
#pragma unroll
  for (int r = 0; r < M; ++r) {
    if (r < m) {
      const float* t = tab + r * kRowTab + 8;
      v[r][0] -= t[0] * lr0 + t[1] * lr1 + t[4] * lr2 + t[5] * lr3;
      v[r][1] -= t[2] * lr0 + t[3] * lr1 + t[6] * lr2 + t[7] * lr3;
    }
  }
}

// FP32-refined variant of ccd_lines: the FP64 right-hand side is solved in
// FP32 (x1), the residual of the exact block system (interior rows
// Lb x_{i-1} + x_i + Ub x_{i+1}, wall closure rows with diagonal I + Lb Ml /
// I + Ub Mr) is formed in FP64 from x1 and the neighbours' end rows, a second
// FP32 solve gives the correction d, and x = x1 + d is scaled and stored in
// FP64.  HBM traffic is that of the FP64 kernel.
template <int M>
__device__ __forceinline__ void ccd_lines_ref(const CcdAxisGeom& g, const CcdLines& ln,
                                              const bool CONTIG, const bool TILED,
                                              const double* __restrict__ f,
                                              double* __restrict__ d1, double* __restrict__ d2,
                                              const unsigned block) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x & 63;
  const int p = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t b = (int64_t)block * 64 + lane;
  const bool active = b < ln.n_lines;
  int64_t base = 0;
  if (active) {
    if (CONTIG) {
      base = b * ln.inner_stride;
    } else {
      const uint32_t bi = (uint32_t)b, in = (uint32_t)ln.inner;
      base = (int64_t)(bi / in) * ln.outer + (int64_t)(bi % in) * ln.inner_stride;
    }
  }
  const int64_t st = ln.stride;
  const int s0 = g.s[p];
  const int m = g.s[p + 1] - s0;
  const bool periodic = g.periodic != 0;
  const bool wall_first = !periodic && p == 0;
  const bool wall_last = !periodic && p == g.P - 1;

  double fv[M + 2];
  const int j0 = periodic ? s0 - 1 : s0;
  if (M == 16 && TILED) {
    const int nax = (int)ln.inner_stride;
    const int64_t rem = ln.n_lines - (int64_t)block * 64;
    const int cnt = (int)(rem < 64 ? rem : 64) * nax;
    const double* src = f + (int64_t)block * 64 * nax;
    double tmp[kTilePer];
#pragma unroll
    for (int i = 0; i < kTilePer; ++i) {
      const int e = threadIdx.x + i * (int)blockDim.x;
      tmp[i] = e < cnt ? src[e] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < kTilePer; ++i) {
      const int e = threadIdx.x + i * (int)blockDim.x;
      if (e < cnt) lds[e] = tmp[i];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < M + 2; ++k) {
      fv[k] = 0.0;
      if (k < m + 2) {
        int j = j0 + k;
        if (periodic) j = j < 0 ? j + g.N : (j >= g.N ? j - g.N : j);
        if (active) fv[k] = lds[lane * nax + j];
      }
    }
    __syncthreads();
  } else {
#pragma unroll
    for (int k = 0; k < M + 2; ++k) {
      fv[k] = 0.0;
      if (k < m + 2) {
        int j = j0 + k;
        if (periodic) j = j < 0 ? j + g.N : (j >= g.N ? j - g.N : j);
        if (active) fv[k] = f[base + (int64_t)j * st];
      }
    }
  }
  // FP32 tables, the partition-end exchange G and the x1 end rows E in LDS
  float* ltab = reinterpret_cast<float*>(lds);
  float* lq = ltab + g.P * M * kRowTab;
  float* G = lq + g.P * 16 * g.NQ;
  float* E = G + 4 * g.P * 64;
  {
    const int nt = g.P * M * kRowTab, nq = g.P * 16 * g.NQ;
    stage_table(ltab, g.rowtab32, nt, same_index);
    stage_table(lq, g.qtab32, nq, same_index);
    __syncthreads();
  }
  double cl0 = 0.0, cl1 = 0.0, bl0 = 0.0, bl1 = 0.0;
  if (wall_first) {
#pragma unroll
    for (int k = 0; k < 4; ++k) bl0 += g.cIl[k] * fv[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) bl1 += g.cIIl[k] * fv[k];
    cl0 = kAlpha1 * bl0 + kB1 * bl1;
    cl1 = kB2 * bl0 + kBeta2 * bl1;
  }
  double cr0 = 0.0, cr1 = 0.0, br0 = 0.0, br1 = 0.0;
  if (wall_last) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double v = active ? f[base + (int64_t)(g.N - k) * st] : 0.0;
      if (k < 4) br0 += g.cIr[k] * v;
      br1 += g.cIIr[k] * v;
    }
    cr0 = kAlpha1 * br0 - kB1 * br1;
    cr1 = -kB2 * br0 + kBeta2 * br1;
  }
  // FP64 right-hand side of row r (closure terms on the end rows)
  auto rhs = [&](int r, double& r0, double& r1) {
    r0 = kA1 * (fv[r + 2] - fv[r]);
    r1 = kA2 * ((fv[r] - 2.0 * fv[r + 1]) + fv[r + 2]);
    if (r == 0) {
      r0 -= cl0;
      r1 -= cl1;
    }
    if (r == m - 1) {
      r0 -= cr0;
      r1 -= cr1;
    }
  };
  const float* tab = ltab + (size_t)p * M * kRowTab;
  float x1[M][2], v[M][2];
#pragma unroll
  for (int r = 0; r < M; ++r) {
    x1[r][0] = x1[r][1] = 0.f;
    if (r < m) {
      double r0, r1;
      rhs(r, r0, r1);
      x1[r][0] = (float)r0;
      x1[r][1] = (float)r1;
    }
  }
  solve32<M>(g, tab, lq, x1, m, p, lane, periodic, G);
  // end rows of x1 for the neighbours' residual rows (static indices only)
  E[(p * 4 + 0) * 64 + lane] = x1[0][0];
  E[(p * 4 + 1) * 64 + lane] = x1[0][1];
#pragma unroll
  for (int r = 0; r < M; ++r)
    if (r == m - 1) {
      E[(p * 4 + 2) * 64 + lane] = x1[r][0];
      E[(p * 4 + 3) * 64 + lane] = x1[r][1];
    }
  __syncthreads();
  const bool has_left = periodic || p > 0, has_right = periodic || p < g.P - 1;
  const int pl = p > 0 ? p - 1 : g.P - 1, pr = p < g.P - 1 ? p + 1 : 0;
  double lx0 = 0.0, lx1 = 0.0, rx0 = 0.0, rx1 = 0.0;  // neighbour rows
  if (has_left) {
    lx0 = E[(pl * 4 + 2) * 64 + lane];
    lx1 = E[(pl * 4 + 3) * 64 + lane];
  }
  if (has_right) {
    rx0 = E[(pr * 4 + 0) * 64 + lane];
    rx1 = E[(pr * 4 + 1) * 64 + lane];
  }
#pragma unroll
  for (int r = 0; r < M; ++r) {
    v[r][0] = v[r][1] = 0.f;
    if (r < m) {
      double r0, r1;
      rhs(r, r0, r1);
      const double u = x1[r][0], w = x1[r][1];
      // diagonal block: identity, or the closure-modified end rows
      double a0 = u, a1 = w;
      if (wall_first && r == 0) {
        const double mu = g.Ml[0] * u + g.Ml[1] * w, mv = g.Ml[2] * u + g.Ml[3] * w;
        a0 += kAlpha1 * mu + kB1 * mv;
        a1 += kB2 * mu + kBeta2 * mv;
      }
      if (wall_last && r == m - 1) {
        const double mu = g.Mr[0] * u + g.Mr[1] * w, mv = g.Mr[2] * u + g.Mr[3] * w;
        a0 += kAlpha1 * mu - kB1 * mv;
        a1 += -kB2 * mu + kBeta2 * mv;
      }
      // lower neighbour (Lb) and upper neighbour (Ub)
      const bool lo = r > 0 || has_left, hi = r < m - 1 || has_right;
      const double l0 = r > 0 ? (double)x1[r - 1][0] : lx0, l1 = r > 0 ? (double)x1[r - 1][1] : lx1;
      const double h0 = r < m - 1 ? (double)x1[r + 1][0] : rx0;
      const double h1 = r < m - 1 ? (double)x1[r + 1][1] : rx1;
      if (lo) {
        a0 += kAlpha1 * l0 + kB1 * l1;
        a1 += kB2 * l0 + kBeta2 * l1;
      }
      if (hi) {
        a0 += kAlpha1 * h0 - kB1 * h1;
        a1 += -kB2 * h0 + kBeta2 * h1;
      }
      v[r][0] = (float)(r0 - a0);
      v[r][1] = (float)(r1 - a1);
    }
  }
  solve32<M>(g, tab, lq, v, m, p, lane, periodic, G);
  auto xo = [&](int r, int c) { return (double)x1[r][c] + (double)v[r][c]; };
  const double x00 = xo(0, 0), x01 = xo(0, 1);
  double xl0 = 0.0, xl1 = 0.0;  // last row (wall node N closure)
#pragma unroll
  for (int r = 0; r < M; ++r)
    if (r == m - 1) {
      xl0 = xo(r, 0);
      xl1 = xo(r, 1);
    }
  const int off = periodic ? 0 : 1;
  if (active && !(M == 16 && TILED)) {
    if (periodic && p == 0) {
      if (d1) d1[base + (int64_t)g.N * st] = x00 * g.inv_h;
      if (d2) d2[base + (int64_t)g.N * st] = x01 * g.inv_h2;
    }
    if (wall_first) {
      if (d1) d1[base] = (g.Ml[0] * x00 + g.Ml[1] * x01 + bl0) * g.inv_h;
      if (d2) d2[base] = (g.Ml[2] * x00 + g.Ml[3] * x01 + bl1) * g.inv_h2;
    }
    if (wall_last) {
      if (d1) d1[base + (int64_t)g.N * st] = (g.Mr[0] * xl0 + g.Mr[1] * xl1 + br0) * g.inv_h;
      if (d2) d2[base + (int64_t)g.N * st] = (g.Mr[2] * xl0 + g.Mr[3] * xl1 + br1) * g.inv_h2;
    }
  }
  if (!CONTIG) {
    if (active) {
#pragma unroll
      for (int r = 0; r < M; ++r) {
        if (r < m) {
          const int64_t a = base + (int64_t)(s0 + r + off) * st;
          if (d1) __builtin_nontemporal_store(xo(r, 0) * g.inv_h, d1 + a);
          if (d2) __builtin_nontemporal_store(xo(r, 1) * g.inv_h2, d2 + a);
        }
      }
    }
  } else if (M == 16 && TILED) {
    const int nax = (int)ln.inner_stride;
    const int64_t rem = ln.n_lines - (int64_t)block * 64;
    const int cnt = (int)(rem < 64 ? rem : 64) * nax;
    __syncthreads();  // tables and exchange buffers dead; the tile space is reused
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
      double* dst = comp == 0 ? d1 : d2;
      const double scale = comp == 0 ? g.inv_h : g.inv_h2;
      if (dst == nullptr) continue;
      double* row = lds + lane * nax;
#pragma unroll
      for (int r = 0; r < M; ++r)
        if (r < m) row[s0 + r + off] = xo(r, comp) * scale;
      if (periodic && p == 0) row[g.N] = (comp == 0 ? x00 : x01) * scale;
      if (wall_first)
        row[0] = (comp == 0 ? g.Ml[0] * x00 + g.Ml[1] * x01 + bl0
                            : g.Ml[2] * x00 + g.Ml[3] * x01 + bl1) * scale;
      if (wall_last)
        row[g.N] = (comp == 0 ? g.Mr[0] * xl0 + g.Mr[1] * xl1 + br0
                              : g.Mr[2] * xl0 + g.Mr[3] * xl1 + br1) * scale;
      __syncthreads();
      double* out = dst + (int64_t)block * 64 * nax;
#pragma unroll
      for (int i = 0; i < kTilePer; ++i) {
        const int e = threadIdx.x + i * (int)blockDim.x;
        if (e < cnt) __builtin_nontemporal_store(lds[e], out + e);
      }
      __syncthreads();
    }
  } else {
    __syncthreads();  // exchange buffers dead; the staging space is reused
    double* S = lds + (size_t)p * 64 * kStagePitch;
    const int64_t line0 = (int64_t)block * 64;
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
      double* dst = comp == 0 ? d1 : d2;
      const double scale = comp == 0 ? g.inv_h : g.inv_h2;
      if (dst == nullptr) continue;
#pragma unroll
      for (int r = 0; r < kStageRows; ++r)
        if (r < M) S[lane * kStagePitch + r] = xo(r, comp) * scale;
      __syncthreads();
#pragma unroll
      for (int it = 0; it < kStageRows; ++it) {
        const int e = lane + 64 * it;
        const int l = e >> 4, r = e & 15;
        const int64_t bl = line0 + l;
        if (r < m && bl < ln.n_lines)
          __builtin_nontemporal_store(S[l * kStagePitch + r],
                                      dst + bl * ln.inner_stride + (int64_t)(s0 + r + off));
      }
      __syncthreads();
    }
  }
}

// Wavefront-per-line mapping: lane = partition (P <= 64), LPW = 64 / P'
// lines per wavefront (P' = P rounded up to a power of two: 4 lines of 16
// partitions, 2 of 32, 1 of 64), and a host-chosen number of wavefronts per
// workgroup, so a workgroup covers LPW x waves lines.  The workgroup's lines
// are staged through LDS as one [lines][nax] tile, loaded and stored with
// coalesced accesses on either axis (consecutive lines share every node's
// cache line on a strided axis; they are one contiguous block on the
// contiguous axis).  Tile rows are padded by one double per 16 nodes so the
// per-lane partition reads (16 nodes apart) spread over the LDS banks.  The
// per-lane work is the lane-per-line solve above; tables come from L2; the
// partition end values are exchanged through the (then dead) tile space.
// Small workgroups (4 waves) with several lines per wavefront let four
// workgroups share a CU, so one loads while another solves or stores.
constexpr int kWaveLines = 16;  // lines per workgroup of the batch kernels

__host__ __device__ inline int wave_pitch(int nax) { return nax + (nax >> 4) + 1; }

// Non-uniform axes (wall lines) stage their distinct partition row tables in
// LDS after the tile region when they take at most this many doubles, so the
// dependent sweeps read them from LDS instead of L2.
// Each staged table is padded by 2 doubles so lanes of different tables read
// different LDS banks.
constexpr int kWaveTabMax = 2048;
__host__ __device__ inline int wave_tab_stride(int M) { return M * kRowTab + 2; }
__host__ __device__ inline bool wave_tab_lds(const CcdAxisGeom& g) {
  return !g.uniform && g.ntypes * wave_tab_stride(g.M) <= kWaveTabMax;
}
// Doubles of the tile region: the [lines][pitch] tile, which later holds the
// 4 x 64 exchange buffer of every wavefront.
__host__ __device__ inline int wave_region(int lines, int pitch, int waves) {
  return lines * pitch > waves * 4 * 64 ? lines * pitch : waves * 4 * 64;
}
__device__ inline int tix(int pitch, int l, int j) { return l * pitch + j + (j >> 4); }

// Cooperative tile transfer between global lines [line0, line0 + nl) and LDS.
template <bool LOAD>
__device__ __forceinline__ void wave_tile_io(const CcdLines& ln, int nax, int pitch,
                                             int64_t line0, int nl, int nlw, const double* src,
                                             double* dst, double* tile) {
  const int cnt = nl * nax;
  if (ln.stride == 1) {  // contiguous axis: one block of nl*nax doubles
    // Batches of kB elements per thread: all global loads of a batch are
    // issued before any LDS write; (line, node) of element e advance by
    // blockDim.x = q*nax + rr without a division per element.
    constexpr int kB = 8;
    const int64_t g0 = line0 * ln.inner_stride;
    const int bd = (int)blockDim.x;
    const int q = bd / nax, rr = bd - q * nax;
    int l = (int)threadIdx.x / nax, j = (int)threadIdx.x - l * nax;
    for (int e0 = (int)threadIdx.x; e0 < cnt; e0 += kB * bd) {
      double v[kB];
      int ti[kB];
#pragma unroll
      for (int i = 0; i < kB; ++i) {
        const int e = e0 + i * bd;
        ti[i] = tix(pitch, l, j);
        if (LOAD) v[i] = e < cnt ? src[g0 + e] : 0.0;
        else v[i] = e < cnt ? tile[ti[i]] : 0.0;
        l += q;
        j += rr;
        if (j >= nax) {
          j -= nax;
          ++l;
        }
      }
#pragma unroll
      for (int i = 0; i < kB; ++i) {
        const int e = e0 + i * bd;
        if (e < cnt) {
          if (LOAD) tile[ti[i]] = v[i];
          else __builtin_nontemporal_store(v[i], dst + g0 + e);
        }
      }
    }
  } else {  // strided axis: consecutive threads take consecutive lines of one node
    // The workgroup size is a multiple of nlw (64 / LPW threads per line
    // slot), so a thread keeps its line l and advances its node j by
    // blockDim.x / nlw: the line's base address is formed once, and batches
    // of kB elements per thread issue all global accesses before any LDS
    // access (one load in flight per thread serialised the tile on HBM
    // latency: 1025^2 strided axis 0.039 ms).
    constexpr int kB = 8;
    const int bd = (int)blockDim.x;
    const int l = (int)threadIdx.x % nlw;
    const int jstep = bd / nlw;
    const bool lok = l < nl;
    int64_t base = 0;
    if (lok) {
      const uint32_t bi = (uint32_t)(line0 + l), in = (uint32_t)ln.inner;
      base = (int64_t)(bi / in) * ln.outer + (int64_t)(bi % in) * ln.inner_stride;
    }
    for (int j = (int)threadIdx.x / nlw; j < nax; j += kB * jstep) {
      double v[kB];
#pragma unroll
      for (int i = 0; i < kB; ++i) {
        const int jj = j + i * jstep;
        const bool ok = lok && jj < nax;
        if (LOAD) v[i] = ok ? src[base + (int64_t)jj * ln.stride] : 0.0;
        else v[i] = ok ? tile[tix(pitch, l, jj)] : 0.0;
      }
#pragma unroll
      for (int i = 0; i < kB; ++i) {
        const int jj = j + i * jstep;
        if (lok && jj < nax) {
          if (LOAD) tile[tix(pitch, l, jj)] = v[i];
          else __builtin_nontemporal_store(v[i], dst + base + (int64_t)jj * ln.stride);
        }
      }
    }
  }
}

template <int M, bool UNI, int LPW>
__device__ __forceinline__ void ccd_wave_tile(const CcdAxisGeom& g, const CcdLines& ln,
                                              const double* __restrict__ f,
                                              double* __restrict__ d1, double* __restrict__ d2,
                                              const unsigned block) {
  extern __shared__ double lds[];
  constexpr int PW = 64 / LPW;  // lanes per line
  const int lane = threadIdx.x & 63;
  const int p = lane & (PW - 1);           // partition
  const int lbase = lane & ~(PW - 1);      // first lane of this lane's line
  const int nwaves = (int)(blockDim.x >> 6);
  const int nlw = LPW * nwaves;            // lines per workgroup
  const int wv = threadIdx.x >> 6;
  const int w = wv * LPW + lane / PW;      // line within the workgroup
  const int64_t line0 = (int64_t)block * nlw;
  const int64_t remain = ln.n_lines - line0;
  const int nl = (int)(remain < nlw ? remain : nlw);
  const int nax = g.N + 1;
  const int pitch = wave_pitch(nax);
  const bool active = w < nl && p < g.P;
  const int pp = p < g.P ? p : g.P - 1;  // idle lanes mirror the last partition (no stores)
  // partition pp holds unknowns [n pp / P, n (pp+1) / P); with UNI the
  // partitions are equal, so m is wavefront-uniform and the row loops below
  // take uniform branches instead of per-lane masks
  const int mu = (int)((unsigned)g.n / (unsigned)g.P);
  const int s0 = UNI ? pp * mu : (int)((unsigned)(g.n * pp) / (unsigned)g.P);
  const int m = UNI ? mu : (int)((unsigned)(g.n * (pp + 1)) / (unsigned)g.P) - s0;
  const bool periodic = g.periodic != 0;
  const bool wall_first = active && !periodic && p == 0;
  const bool wall_last = active && !periodic && p == g.P - 1;

  wave_tile_io<true>(ln, nax, pitch, line0, nl, nlw, f, nullptr, lds);
  const bool ltab = !UNI && wave_tab_lds(g);
  double* tabl = lds + wave_region(nlw, pitch, nwaves);
  if (ltab)
    stage_table(tabl, g.rowtab_u, g.ntypes * M * kRowTab, [](int i) {
      return (i / (M * kRowTab)) * wave_tab_stride(M) + i % (M * kRowTab);
    });
  __syncthreads();
  const int lw = w < nl ? w : 0;  // rows of inactive waves read line 0 (never stored)
  double fv[M + 2];
  const int j0 = periodic ? s0 - 1 : s0;
#pragma unroll
  for (int k = 0; k < M + 2; ++k) {
    fv[k] = 0.0;
    if (k < m + 2) {
      int j = j0 + k;
      if (periodic) j = j < 0 ? j + g.N : (j >= g.N ? j - g.N : j);
      fv[k] = lds[tix(pitch, lw, j)];
    }
  }
  double cl0 = 0.0, cl1 = 0.0, bl0 = 0.0, bl1 = 0.0;
  if (wall_first) {
#pragma unroll
    for (int k = 0; k < 4; ++k) bl0 += g.cIl[k] * fv[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) bl1 += g.cIIl[k] * fv[k];
    cl0 = kAlpha1 * bl0 + kB1 * bl1;
    cl1 = kB2 * bl0 + kBeta2 * bl1;
  }
  double cr0 = 0.0, cr1 = 0.0, br0 = 0.0, br1 = 0.0;
  if (wall_last) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double v = lds[tix(pitch, lw, g.N - k)];
      if (k < 4) br0 += g.cIr[k] * v;
      br1 += g.cIIr[k] * v;
    }
    cr0 = kAlpha1 * br0 - kB1 * br1;
    cr1 = -kB2 * br0 + kBeta2 * br1;
  }
  __syncthreads();  // tile consumed; its space holds the exchange buffer below
  double* Gw = lds + wv * 4 * 64;
  // reduced-system rows of all partitions (partition fastest, so a lane per
  // partition reads consecutive doubles) -> LDS after the exchange buffer,
  // when they fit the tile space (the host sized the tile for it)
  const int nq = g.P * 16 * g.NQ;
  const bool qlds = !UNI && nwaves * 4 * 64 + nq <= nlw * wave_pitch(nax);
  const double* qsrc = g.qtab_t;
  if (qlds) {
    double* qd = lds + nwaves * 4 * 64;
    stage_table(qd, g.qtab_t, nq, same_index);
    qsrc = qd;  // visible after the barrier that precedes the coupling
  }
  // UNI: every partition's rows are partition 0's, read wavefront-uniform
  // through the scalar cache instead of per lane on the dependent sweeps
  using TabPtr = typename std::conditional<UNI, kdptr, const double*>::type;
  const TabPtr tab = UNI    ? (TabPtr)g.rowtab
                     : ltab ? (TabPtr)(tabl + (int)g.rtype[pp] * wave_tab_stride(M))
                            : (TabPtr)(g.rowtab + (size_t)pp * M * kRowTab);
  double x[M][2];
#pragma unroll
  for (int r = 0; r < M; ++r) {
    if (r < m) {
      double r0 = kA1 * (fv[r + 2] - fv[r]);
      double r1 = kA2 * ((fv[r] - 2.0 * fv[r + 1]) + fv[r + 2]);
      if (r == 0) {
        r0 -= cl0;
        r1 -= cl1;
      }
      if (r == m - 1) {
        r0 -= cr0;
        r1 -= cr1;
      }
      if (r > 0) {
        const TabPtr t = tab + r * kRowTab;
        r0 -= t[0] * x[r - 1][0] + t[1] * x[r - 1][1];
        r1 -= t[2] * x[r - 1][0] + t[3] * x[r - 1][1];
      }
      x[r][0] = r0;
      x[r][1] = r1;
    }
  }
#pragma unroll
  for (int r = M - 1; r >= 0; --r) {
    if (r < m) {
      double a = x[r][0], c = x[r][1];
      if (r + 1 < M && r < m - 1) {
        a -= kAlpha1 * x[r + 1][0] - kB1 * x[r + 1][1];
        c -= -kB2 * x[r + 1][0] + kBeta2 * x[r + 1][1];
      }
      const TabPtr t = tab + r * kRowTab + 4;
      x[r][0] = t[0] * a + t[1] * c;
      x[r][1] = t[2] * a + t[3] * c;
      if (r == m - 1) {
        Gw[2 * 64 + lane] = x[r][0];
        Gw[3 * 64 + lane] = x[r][1];
      }
    }
  }
  Gw[0 * 64 + lane] = x[0][0];
  Gw[1 * 64 + lane] = x[0][1];
  __syncthreads();
  double lr0 = 0.0, lr1 = 0.0, lr2 = 0.0, lr3 = 0.0;
  {
    const int NQ = g.NQ;
    // entry (c, t, k) of this partition's rows: partition 0's [c][NQ*4]
    // block when UNI, else the transposed table [c][NQ*4][P]
    const int qs = UNI ? 1 : g.P;
    const TabPtr q = UNI ? (TabPtr)g.qtab : (TabPtr)(qsrc + pp);
    for (int t = 0; t < NQ; ++t) {
      int qq = g.rel ? pp + g.d0 + t : t;
      if (periodic) qq = qq < 0 ? qq + g.P : (qq >= g.P ? qq - g.P : qq);
      else qq = qq < 0 ? 0 : (qq >= g.P ? g.P - 1 : qq);
      const double G0 = Gw[0 * 64 + lbase + qq], G1 = Gw[1 * 64 + lbase + qq];
      const double G2 = Gw[2 * 64 + lbase + qq], G3 = Gw[3 * 64 + lbase + qq];
      const TabPtr q0 = q + (0 * NQ + t) * 4 * qs;
      const TabPtr q1 = q + (1 * NQ + t) * 4 * qs;
      const TabPtr q2 = q + (2 * NQ + t) * 4 * qs;
      const TabPtr q3 = q + (3 * NQ + t) * 4 * qs;
      lr0 += q0[0] * G0 + q0[qs] * G1 + q0[2 * qs] * G2 + q0[3 * qs] * G3;
      lr1 += q1[0] * G0 + q1[qs] * G1 + q1[2 * qs] * G2 + q1[3 * qs] * G3;
      lr2 += q2[0] * G0 + q2[qs] * G1 + q2[2 * qs] * G2 + q2[3 * qs] * G3;
      lr3 += q3[0] * G0 + q3[qs] * G1 + q3[2 * qs] * G2 + q3[3 * qs] * G3;
    }
  }
  double xl0 = 0.0, xl1 = 0.0;
#pragma unroll
  for (int r = 0; r < M; ++r) {
    if (r < m) {
      const TabPtr t = tab + r * kRowTab + 8;
      const double u = x[r][0] - (t[0] * lr0 + t[1] * lr1 + t[4] * lr2 + t[5] * lr3);
      const double v = x[r][1] - (t[2] * lr0 + t[3] * lr1 + t[6] * lr2 + t[7] * lr3);
      x[r][0] = u;
      x[r][1] = v;
      if (r == m - 1) {
        xl0 = u;
        xl1 = v;
      }
    }
  }
  const int off = periodic ? 0 : 1;
  __syncthreads();  // exchange buffer dead; the tile space takes the outputs
#pragma unroll
  for (int comp = 0; comp < 2; ++comp) {
    double* dst = comp == 0 ? d1 : d2;
    const double scale = comp == 0 ? g.inv_h : g.inv_h2;
    if (dst == nullptr) continue;  // uniform
    if (active) {
#pragma unroll
      for (int r = 0; r < M; ++r)
        if (r < m) lds[tix(pitch, w, s0 + r + off)] = x[r][comp] * scale;
      if (periodic && p == 0) lds[tix(pitch, w, g.N)] = x[0][comp] * scale;
      if (wall_first)
        lds[tix(pitch, w, 0)] = (comp == 0 ? g.Ml[0] * x[0][0] + g.Ml[1] * x[0][1] + bl0
                                           : g.Ml[2] * x[0][0] + g.Ml[3] * x[0][1] + bl1) * scale;
      if (wall_last)
        lds[tix(pitch, w, g.N)] = (comp == 0 ? g.Mr[0] * xl0 + g.Mr[1] * xl1 + br0
                                             : g.Mr[2] * xl0 + g.Mr[3] * xl1 + br1) * scale;
    }
    __syncthreads();
    wave_tile_io<false>(ln, nax, pitch, line0, nl, nlw, nullptr, dst, lds);
    __syncthreads();
  }
}

// FP32 partition solve of the wavefront mapping (lane = partition): block-LU
// sweeps with the FP32 tables, partition ends -> Gw ([4][64] per wavefront),
// barrier, coupling through the reduced rows, spike correction.  Every
// thread of the workgroup calls it (it holds a barrier).
template <int M, bool UNI>
__device__ __forceinline__ void wave_solve32(const CcdAxisGeom& g, int pp, int lane, int lbase,
                                             int m, bool periodic, float (&v)[M][2], float* Gw) {
  using TabF = typename std::conditional<UNI, kfptr, const float*>::type;
  const TabF tab = UNI ? (TabF)g.rowtab32 : (TabF)(g.rowtab32 + (size_t)pp * M * kRowTab);
#pragma unroll
  for (int r = 1; r < M; ++r) {
    if (r < m) {
      const TabF t = tab + r * kRowTab;
      v[r][0] -= t[0] * v[r - 1][0] + t[1] * v[r - 1][1];
      v[r][1] -= t[2] * v[r - 1][0] + t[3] * v[r - 1][1];
    }
  }
#pragma unroll
  for (int r = M - 1; r >= 0; --r) {
    if (r < m) {
      float a = v[r][0], c = v[r][1];
      if (r + 1 < M && r < m - 1) {
        a -= (float)kAlpha1 * v[r + 1][0] - (float)kB1 * v[r + 1][1];
        c -= -(float)kB2 * v[r + 1][0] + (float)kBeta2 * v[r + 1][1];
      }
      const TabF t = tab + r * kRowTab + 4;
      v[r][0] = t[0] * a + t[1] * c;
      v[r][1] = t[2] * a + t[3] * c;
      if (r == m - 1) {
        Gw[2 * 64 + lane] = v[r][0];
        Gw[3 * 64 + lane] = v[r][1];
      }
    }
  }
  Gw[0 * 64 + lane] = v[0][0];
  Gw[1 * 64 + lane] = v[0][1];
  __syncthreads();
  float lr0 = 0.f, lr1 = 0.f, lr2 = 0.f, lr3 = 0.f;
  const int NQ = g.NQ;
  const TabF q = UNI ? (TabF)g.qtab32 : (TabF)(g.qtab32 + (size_t)pp * 4 * NQ * 4);
  for (int t = 0; t < NQ; ++t) {
    int qq = g.rel ? pp + g.d0 + t : t;
    if (periodic) qq = qq < 0 ? qq + g.P : (qq >= g.P ? qq - g.P : qq);
    else qq = qq < 0 ? 0 : (qq >= g.P ? g.P - 1 : qq);
    const float G0 = Gw[0 * 64 + lbase + qq], G1 = Gw[1 * 64 + lbase + qq];
    const float G2 = Gw[2 * 64 + lbase + qq], G3 = Gw[3 * 64 + lbase + qq];
    const TabF q0 = q + (0 * NQ + t) * 4;
    const TabF q1 = q + (1 * NQ + t) * 4;
    const TabF q2 = q + (2 * NQ + t) * 4;
    const TabF q3 = q + (3 * NQ + t) * 4;
    lr0 += q0[0] * G0 + q0[1] * G1 + q0[2] * G2 + q0[3] * G3;
    lr1 += q1[0] * G0 + q1[1] * G1 + q1[2] * G2 + q1[3] * G3;
    lr2 += q2[0] * G0 + q2[1] * G1 + q2[2] * G2 + q2[3] * G3;
    lr3 += q3[0] * G0 + q3[1] * G1 + q3[2] * G2 + q3[3] * G3;
  }
#pragma unroll
  for (int r = 0; r < M; ++r) {
    if (r < m) {
      const TabF t = tab + r * kRowTab + 8;
      v[r][0] -= t[0] * lr0 + t[1] * lr1 + t[4] * lr2 + t[5] * lr3;
      v[r][1] -= t[2] * lr0 + t[3] * lr1 + t[6] * lr2 + t[7] * lr3;
    }
  }
}

// FP32 solves with FP64 iterative refinement on the wavefront mapping: the
// FP64 right-hand side rounded to FP32 and solved (x1); then twice, the
// residual of the exact block system formed in FP64 from the current iterate
// and the neighbour partitions' end rows (exchanged through LDS) and solved
// in FP32 (d1, d2).  x = x1 + d1 + d2, kept as three FP32 parts and summed
// in FP64, is stored through the same tile as the FP64 kernel (same HBM
// traffic).  The second sweep matters on long wall lines: in the scaled
// unknowns (h f', h^2 f'') one sweep leaves ~(kappa 2^-24)^2 of |x|, which
// the h f' part dominates, and f'' of a smooth field is h k times smaller.
// The line values are re-read from the LDS tile for each residual (the
// exchange buffers lie past it), which keeps the kernel within 128 VGPRs.
template <int M, bool UNI, int LPW>
__device__ __forceinline__ void ccd_wave_tile_ref(const CcdAxisGeom& g, const CcdLines& ln,
                                                  const double* __restrict__ f,
                                                  double* __restrict__ d1,
                                                  double* __restrict__ d2,
                                                  const unsigned block) {
  extern __shared__ double lds[];
  constexpr int PW = 64 / LPW;
  const int lane = threadIdx.x & 63;
  const int p = lane & (PW - 1);
  const int lbase = lane & ~(PW - 1);
  const int nwaves = (int)(blockDim.x >> 6);
  const int nlw = LPW * nwaves;
  const int wv = threadIdx.x >> 6;
  const int w = wv * LPW + lane / PW;
  const int64_t line0 = (int64_t)block * nlw;
  const int64_t remain = ln.n_lines - line0;
  const int nl = (int)(remain < nlw ? remain : nlw);
  const int nax = g.N + 1;
  const int pitch = wave_pitch(nax);
  const bool active = w < nl && p < g.P;
  const int pp = p < g.P ? p : g.P - 1;
  const int mu = (int)((unsigned)g.n / (unsigned)g.P);  // equal partitions (UNI)
  const int s0 = UNI ? pp * mu : (int)((unsigned)(g.n * pp) / (unsigned)g.P);
  const int m = UNI ? mu : (int)((unsigned)(g.n * (pp + 1)) / (unsigned)g.P) - s0;
  const bool periodic = g.periodic != 0;
  const bool wall_first = active && !periodic && p == 0;
  const bool wall_last = active && !periodic && p == g.P - 1;

  wave_tile_io<true>(ln, nax, pitch, line0, nl, nlw, f, nullptr, lds);
  __syncthreads();
  const int lw = w < nl ? w : 0;
  const int j0 = periodic ? s0 - 1 : s0;
  auto fv = [&](int k) {
    int j = j0 + k;
    if (periodic) j = j < 0 ? j + g.N : (j >= g.N ? j - g.N : j);
    return lds[tix(pitch, lw, j)];
  };
  double cl0 = 0.0, cl1 = 0.0, bl0 = 0.0, bl1 = 0.0;
  if (wall_first) {
#pragma unroll
    for (int k = 0; k < 4; ++k) bl0 += g.cIl[k] * fv(k);
#pragma unroll
    for (int k = 0; k < 6; ++k) bl1 += g.cIIl[k] * fv(k);
    cl0 = kAlpha1 * bl0 + kB1 * bl1;
    cl1 = kB2 * bl0 + kBeta2 * bl1;
  }
  double cr0 = 0.0, cr1 = 0.0, br0 = 0.0, br1 = 0.0;
  if (wall_last) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double v = lds[tix(pitch, lw, g.N - k)];
      if (k < 4) br0 += g.cIr[k] * v;
      br1 += g.cIIr[k] * v;
    }
    cr0 = kAlpha1 * br0 - kB1 * br1;
    cr1 = -kB2 * br0 + kBeta2 * br1;
  }
  // exchange buffers past the tile: Gw (FP32 partition ends), Ew (FP64 end rows)
  double* xb = lds + (size_t)nlw * pitch + (size_t)wv * 512;
  float* Gw = reinterpret_cast<float*>(xb);
  double* Ew = xb + 128;
  auto rhs = [&](int r, double& r0, double& r1) {
    const double fa = fv(r), fb = fv(r + 1), fc = fv(r + 2);
    r0 = kA1 * (fc - fa);
    r1 = kA2 * ((fa - 2.0 * fb) + fc);
    if (r == 0) {
      r0 -= cl0;
      r1 -= cl1;
    }
    if (r == m - 1) {
      r0 -= cr0;
      r1 -= cr1;
    }
  };
  float x1[M][2], dv[M][2], dw[M][2];
#pragma unroll
  for (int r = 0; r < M; ++r) {
    x1[r][0] = x1[r][1] = 0.f;
    dv[r][0] = dv[r][1] = 0.f;
    dw[r][0] = dw[r][1] = 0.f;
    if (r < m) {
      double r0, r1;
      rhs(r, r0, r1);
      x1[r][0] = (float)r0;
      x1[r][1] = (float)r1;
    }
  }
  wave_solve32<M, UNI>(g, pp, lane, lbase, m, periodic, x1, Gw);
  auto xo = [&](int r, int c) { return ((double)x1[r][c] + (double)dv[r][c]) + (double)dw[r][c]; };
  const bool has_left = periodic || p > 0, has_right = periodic || p < g.P - 1;
  const int pl = pp > 0 ? pp - 1 : g.P - 1, pr = pp < g.P - 1 ? pp + 1 : 0;
#pragma unroll
  for (int sweep = 0; sweep < 2; ++sweep) {
    // end rows of the current iterate for the neighbours' residual rows
    Ew[0 * 64 + lane] = xo(0, 0);
    Ew[1 * 64 + lane] = xo(0, 1);
#pragma unroll
    for (int r = 0; r < M; ++r)
      if (r == m - 1) {
        Ew[2 * 64 + lane] = xo(r, 0);
        Ew[3 * 64 + lane] = xo(r, 1);
      }
    __syncthreads();
    double lx0 = 0.0, lx1 = 0.0, rx0 = 0.0, rx1 = 0.0;
    if (has_left) {
      lx0 = Ew[2 * 64 + lbase + pl];
      lx1 = Ew[3 * 64 + lbase + pl];
    }
    if (has_right) {
      rx0 = Ew[0 * 64 + lbase + pr];
      rx1 = Ew[1 * 64 + lbase + pr];
    }
    float v[M][2];
#pragma unroll
    for (int r = 0; r < M; ++r) {
      v[r][0] = v[r][1] = 0.f;
      if (r < m) {
        double r0, r1;
        rhs(r, r0, r1);
        const double u = xo(r, 0), wq = xo(r, 1);
        double a0 = u, a1 = wq;
        if (wall_first && r == 0) {
          const double mu = g.Ml[0] * u + g.Ml[1] * wq, mv = g.Ml[2] * u + g.Ml[3] * wq;
          a0 += kAlpha1 * mu + kB1 * mv;
          a1 += kB2 * mu + kBeta2 * mv;
        }
        if (wall_last && r == m - 1) {
          const double mu = g.Mr[0] * u + g.Mr[1] * wq, mv = g.Mr[2] * u + g.Mr[3] * wq;
          a0 += kAlpha1 * mu - kB1 * mv;
          a1 += -kB2 * mu + kBeta2 * mv;
        }
        const bool lo = r > 0 || has_left, hi = r < m - 1 || has_right;
        const double l0 = r > 0 ? xo(r - 1, 0) : lx0, l1 = r > 0 ? xo(r - 1, 1) : lx1;
        const double h0 = r < m - 1 ? xo(r + 1, 0) : rx0;
        const double h1 = r < m - 1 ? xo(r + 1, 1) : rx1;
        if (lo) {
          a0 += kAlpha1 * l0 + kB1 * l1;
          a1 += kB2 * l0 + kBeta2 * l1;
        }
        if (hi) {
          a0 += kAlpha1 * h0 - kB1 * h1;
          a1 += -kB2 * h0 + kBeta2 * h1;
        }
        v[r][0] = (float)(r0 - a0);
        v[r][1] = (float)(r1 - a1);
      }
    }
    __syncthreads();  // every lane has read Ew before the next sweep rewrites it
    wave_solve32<M, UNI>(g, pp, lane, lbase, m, periodic, v, Gw);
#pragma unroll
    for (int r = 0; r < M; ++r) {
      if (sweep == 0) {
        dv[r][0] = v[r][0];
        dv[r][1] = v[r][1];
      } else {
        dw[r][0] = v[r][0];
        dw[r][1] = v[r][1];
      }
    }
  }
  const double x00 = xo(0, 0), x01 = xo(0, 1);
  double xl0 = 0.0, xl1 = 0.0;
#pragma unroll
  for (int r = 0; r < M; ++r)
    if (r == m - 1) {
      xl0 = xo(r, 0);
      xl1 = xo(r, 1);
    }
  const int off = periodic ? 0 : 1;
  __syncthreads();  // the tile's inputs are no longer read; it takes the outputs
#pragma unroll
  for (int comp = 0; comp < 2; ++comp) {
    double* dst = comp == 0 ? d1 : d2;
    const double scale = comp == 0 ? g.inv_h : g.inv_h2;
    if (dst == nullptr) continue;  // uniform
    if (active) {
#pragma unroll
      for (int r = 0; r < M; ++r)
        if (r < m) lds[tix(pitch, w, s0 + r + off)] = xo(r, comp) * scale;
      if (periodic && p == 0) lds[tix(pitch, w, g.N)] = (comp == 0 ? x00 : x01) * scale;
      if (wall_first)
        lds[tix(pitch, w, 0)] = (comp == 0 ? g.Ml[0] * x00 + g.Ml[1] * x01 + bl0
                                           : g.Ml[2] * x00 + g.Ml[3] * x01 + bl1) * scale;
      if (wall_last)
        lds[tix(pitch, w, g.N)] = (comp == 0 ? g.Mr[0] * xl0 + g.Mr[1] * xl1 + br0
                                             : g.Mr[2] * xl0 + g.Mr[3] * xl1 + br1) * scale;
    }
    __syncthreads();
    wave_tile_io<false>(ln, nax, pitch, line0, nl, nlw, nullptr, dst, lds);
    __syncthreads();
  }
}

template <int M, int LPW>
__global__ __launch_bounds__(1024) void ccd_wave_kernel_ref(CcdLaunch L) {
  const unsigned blk = blockIdx.x;
  const int sl = (L.nslots > 1 && blk >= L.block_begin[1]) + (L.nslots > 2 && blk >= L.block_begin[2]);
  const CcdSlot& S = L.slot[sl];
  if (S.g.uniform)
    ccd_wave_tile_ref<M, true, LPW>(S.g, S.ln, S.f, S.d1, S.d2, blk - L.block_begin[sl]);
  else
    ccd_wave_tile_ref<M, false, LPW>(S.g, S.ln, S.f, S.d1, S.d2, blk - L.block_begin[sl]);
}

template <int M, int LPW>
__global__ __launch_bounds__(1024) void ccd_wave_kernel(CcdLaunch L) {
  const unsigned blk = blockIdx.x;
  const int sl = (L.nslots > 1 && blk >= L.block_begin[1]) + (L.nslots > 2 && blk >= L.block_begin[2]);
  const CcdSlot& S = L.slot[sl];
  if (S.g.uniform)
    ccd_wave_tile<M, true, LPW>(S.g, S.ln, S.f, S.d1, S.d2, blk - L.block_begin[sl]);
  else
    ccd_wave_tile<M, false, LPW>(S.g, S.ln, S.f, S.d1, S.d2, blk - L.block_begin[sl]);
}

// FP32-refined wavefront kernel: the tile, then 4 KB of exchange buffers
// per wavefront (the tile stays readable for the residuals).
size_t wave_lds_bytes_ref(const CcdAxisGeom& g, int lines, int waves) {
  return sizeof(double) * ((size_t)lines * (size_t)wave_pitch(g.N + 1) + (size_t)waves * 512);
}

// LDS of the wavefront kernel: the [lines][pitch] tile (it also holds the
// 2 KB-per-wavefront exchange buffer once the inputs are read), then the
// staged partition tables of a non-uniform axis.
size_t wave_lds_bytes(const CcdAxisGeom& g, int lines, int waves) {
  return sizeof(double) * ((size_t)wave_region(lines, wave_pitch(g.N + 1), waves) +
                           (wave_tab_lds(g) ? (size_t)g.ntypes * wave_tab_stride(g.M) : 0));
}

// Slot and slot-local block of workgroup blk.  The slots' first min-count
// blocks are interleaved (slot = blk % nslots) so workgroups of different
// axes, which are in different phases (loads, solve, stores), share the
// chip; measured 2 % faster than slot-by-slot order on the 257^3 3-axis
// grid.  Orders that put the axis-1 and axis-2 workgroups reading the same
// planes next to each other (on one XCD, for L2 reuse) were measured slower.
__host__ __device__ inline void slot_of(const CcdLaunch& L, unsigned blk, int& sl, unsigned& local) {
  const unsigned* B = L.block_begin;
  const unsigned ns = (unsigned)L.nslots;
  unsigned mn = B[1] - B[0];
  for (unsigned i = 1; i < ns; ++i) mn = B[i + 1] - B[i] < mn ? B[i + 1] - B[i] : mn;
  if (blk < ns * mn) {
    sl = (int)(blk % ns);
    local = blk / ns;
    return;
  }
  unsigned j = blk - ns * mn;  // the rest, slot by slot
  for (unsigned i = 0; i < ns; ++i) {
    const unsigned r = B[i + 1] - B[i] - mn;
    if (j < r) {
      sl = (int)i;
      local = mn + j;
      return;
    }
    j -= r;
  }
  sl = 0;  // not reached: blk < B[ns]
  local = 0;
}

// Strided axes: renumber a slot's workgroups so that the ones an XCD runs
// together (slot-local indices congruent mod 8 share an XCD under the
// round-robin dispatch of the interleaved order) take adjacent line groups.
// Their 512 B rows are not 128 B aligned on odd node counts, and the cache
// lines two neighbouring groups share are then fetched and written back once,
// through one L2, instead of by two XCDs.
__host__ __device__ inline unsigned xcd_local(unsigned local, unsigned count) {
  const unsigned per = count / 8;
  return local < per * 8 ? (local % 8) * per + local / 8 : local;
}

template <int M, bool REF>
__global__ __launch_bounds__(1024) void ccd_line_kernel(CcdLaunch L) {
  int sl;
  unsigned local;
  slot_of(L, blockIdx.x, sl, local);
  const CcdSlot& S = L.slot[sl];
  if (!S.contig) local = xcd_local(local, L.block_begin[sl + 1] - L.block_begin[sl]);
  if constexpr (REF)
    ccd_lines_ref<M>(S.g, S.ln, S.contig != 0, S.tiled != 0, S.f, S.d1, S.d2, local);
  else
    ccd_lines<M>(S.g, S.ln, S.contig != 0, S.tiled != 0, S.f, S.d1, S.d2, local);
}


// ---- explicit partition inverse (lane per line, strided axes) --------------
// Same mapping as ccd_lines (lane = line, wavefront = partition), but the
// local solve is the precomputed partition inverse applied to the first
// differences of the line values: y_r = S_r . (fv[k+1] - fv[k]), 2 x (M+1)
// per row, wall closure folded in (differences keep f'' free of the
// cancellation of a sum over raw values).  No
// recurrence state is kept: a lane holds only fv, so the kernel fits 64
// VGPRs and two workgroups share a CU, one loading or storing while the
// other computes.  Tables are wavefront-uniform and read through the scalar
// cache (constant address space).  Rows are produced in order and stored
// directly (strided axes: every store instruction writes 64 consecutive
// lines).

// Aligned store ownership (AL > 0): on 2^k + 1 node axes the rows of a line
// group start at 8 B offsets, so a wavefront store of 64 consecutive lines
// covers two 128 B lines that the neighbouring groups also write, and those
// partial-line writes halve the store rate (4.1 against 6.4 TB/s for the
// bare pattern on 257^3).  With AL > 0 workgroup w solves the 64 lines
// [w*(64-AL) - AL, w*(64-AL) - AL + 64) and stores element (line b, row r)
// only when it owns the AL-element aligned chunk of that address: the
// chunk's first element (clipped to the start of b's run of consecutive
// lines) is line `lead`, and the owner is (lead + AL) / (64 - AL).  Every
// element has exactly one owner, the owner's window holds every line of the
// chunk, and each aligned chunk inside a run is written by one store
// instruction.  The AL lines shared with the previous window are solved
// twice (same arithmetic, so the stored values do not depend on AL).
template <int AL>
struct ExplicitOwner {
  static constexpr int OWN = AL ? 64 - AL : 64;
  static constexpr uint32_t MASK = AL ? AL - 1 : 0;
  uint32_t lane, kk;  // lane = b - (OWN*w - AL); kk = b's index inside its run
  bool active;
  // Whether this lane stores the element whose address (in elements) has
  // low bits e: lead = b - min(e mod AL, kk) and (lead + AL) / OWN == w,
  // i.e. 0 <= lane - min(e mod AL, kk) < OWN.
  __device__ __forceinline__ bool owns(uint32_t e) const {
    if (AL == 0) return active;
    uint32_t off = e & MASK;
    off = off < kk ? off : kk;
    return active && lane - off < (uint32_t)OWN;
  }
};

__device__ __forceinline__ uint32_t elem_bits(const double* p) {
  return (uint32_t)((uintptr_t)p >> 3);
}

// Memory access of the explicit kernel.  BUF: raw buffer loads and stores
// with 32-bit byte offsets, the row part of the offset in the scalar
// operand; a masked lane has kOOR set in its offset, so its load returns 0
// and its store is dropped, without a branch.  Every array must be smaller
// than 2 GiB (checked on the host).  !BUF: 64-bit global addresses and
// exec-masked stores.
constexpr uint32_t kOOR = 0x80000000u;
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t ccd_rsrc(const double* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double ccd_bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
// non-temporal store (outputs stream past the caches, f stays resident)
__device__ __forceinline__ void ccd_bstore(__amdgpu_buffer_rsrc_t r, double v, uint32_t voff,
                                           uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, v), r, voff, soff, 2);
}

// One lane's line in a line group of the explicit kernel.
template <int AL>
struct ExplicitLine {
  ExplicitOwner<AL> own;
  int64_t base;  // element offset of the line's node 0
  uint32_t e1;   // low address bits (in elements) of d1's node 0
  uint32_t vb;   // byte offset of node 0 (kOOR: the lane has no line)
};

template <int AL>
__device__ __forceinline__ ExplicitLine<AL> explicit_line(const CcdLines& ln, const double* d1,
                                                          unsigned block) {
  ExplicitLine<AL> L;
  const int lane = threadIdx.x & 63;
  const int64_t b = (int64_t)block * ExplicitOwner<AL>::OWN - AL + lane;
  const bool active = b >= 0 && b < ln.n_lines;
  L.own.active = active;
  L.own.lane = (uint32_t)lane;
  L.own.kk = 0;
  L.base = 0;
  if (active) {
    const uint32_t bi = (uint32_t)b, in = (uint32_t)ln.inner;
    const uint32_t q = bi / in;
    L.own.kk = bi - q * in;
    L.base = (int64_t)q * ln.outer + (int64_t)L.own.kk * ln.inner_stride;
  }
  // Store ownership follows d1's aligned chunks for both outputs (d2 has the
  // same alignment in the usual case; any chunking is a valid partition).
  L.e1 = elem_bits(d1) + (uint32_t)L.base;
  L.vb = active ? (uint32_t)L.base * 8u : kOOR;
  return L;
}

// Node j (wavefront-uniform) of a lane's line.
template <int AL, bool BUF>
__device__ __forceinline__ double explicit_node(const CcdAxisGeom& g, const CcdLines& ln,
                                                const double* f, const ExplicitLine<AL>& L, int j) {
  if constexpr (BUF) {
    const uint32_t bytes = (uint32_t)(ln.n_lines * (int64_t)(g.N + 1) * 8);
    return ccd_bload(ccd_rsrc(f, bytes), L.vb, (uint32_t)((int64_t)j * ln.stride) * 8u);
  } else {
    return L.own.active ? f[L.base + (int64_t)j * ln.stride] : 0.0;
  }
}

// Issues the loads of partition p of a lane's line: fv[k] = node (s0 - 1 + k)
// periodic (wrapped), (s0 + k) wall; nodes past the partition (short
// partitions) are clamped into the line and meet zero table columns.
template <int M, int AL, bool BUF>
__device__ __forceinline__ void explicit_load(const CcdAxisGeom& g, const CcdLines& ln,
                                              const double* f, const ExplicitLine<AL>& L, int p,
                                              double (&fv)[M + 2]) {
  const int N = g.N;
  const bool periodic = g.periodic != 0;
  const int s0 = g.s[p];
  const int j0 = periodic ? s0 - 1 : s0;
#pragma unroll
  for (int k = 0; k < M + 2; ++k) {
    int j = j0 + k;
    j = j < 0 ? j + N : j;
    j = periodic ? (j >= N ? j - N : j) : (j > N ? N : j);
    fv[k] = explicit_node<AL, BUF>(g, ln, f, L, j);
  }
}

// Solves partition p of the lanes' lines from the loaded fv and stores the
// rows; G is the workgroup's [4P][64] exchange buffer (one barrier).
template <int M, int AL, bool BUF>
__device__ __forceinline__ void explicit_solve(const CcdAxisGeom& g, const CcdLines& ln,
                                               const double* f, double* d1, double* d2,
                                               const ExplicitLine<AL>& L, int p,
                                               double (&fv)[M + 2], double* G) {
  constexpr int K = M + 2;
  const int lane = threadIdx.x & 63;
  const int64_t st = ln.stride;
  const int N = g.N;
  const int s0 = g.s[p];
  const int m = g.s[p + 1] - s0;
  const bool periodic = g.periodic != 0;
  const bool wall_first = !periodic && p == 0;
  const bool wall_last = !periodic && p == g.P - 1;
  // (unused without BUF) a missing output gets an empty range: its stores drop
  const uint32_t bytes = (uint32_t)(ln.n_lines * (int64_t)(N + 1) * 8);
  const __amdgpu_buffer_rsrc_t r1 = ccd_rsrc(d1, d1 ? bytes : 0u), r2 = ccd_rsrc(d2, d2 ? bytes : 0u);
  // stores node jn of both outputs (scaled back from (h f', h^2 f'')) when
  // `keep` (uniform) holds and this lane owns the node's chunk
  auto store_node = [&](int jn, double u, double v, bool keep) {
    const int64_t ro = (int64_t)jn * st;
    const bool o = keep && L.own.owns(L.e1 + (uint32_t)ro);
    if constexpr (BUF) {  // the whole offset in the vector operand
      const uint32_t voff = o ? L.vb + (uint32_t)ro * 8u : kOOR;
      ccd_bstore(r1, u * g.inv_h, voff, 0);
      ccd_bstore(r2, v * g.inv_h2, voff, 0);
    } else if (o) {
      if (d1) __builtin_nontemporal_store(u * g.inv_h, d1 + L.base + ro);
      if (d2) __builtin_nontemporal_store(v * g.inv_h2, d2 + L.base + ro);
    }
  };

  // wall closure right-hand sides (ccd_boundary_helpers.py:64-116) from the
  // raw values: the first partition holds nodes 0..5 in fv[0..5], the last
  // one nodes N-5..N in fv[m-4..m+1]
  double bl0 = 0.0, bl1 = 0.0, br0 = 0.0, br1 = 0.0;
  if (wall_first) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (k < 4) bl0 += g.cIl[k] * fv[k];
      bl1 += g.cIIl[k] * fv[k];
    }
  }
  if (wall_last) {
    if (m < 4) {
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const double v = explicit_node<AL, BUF>(g, ln, f, L, N - k);
        if (k < 4) br0 += g.cIr[k] * v;
        br1 += g.cIIr[k] * v;
      }
    }
#pragma unroll
    for (int mm = 4; mm <= M; ++mm) {
      if (m == mm) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          if (k < 4) br0 += g.cIr[k] * fv[mm + 1 - k];
          br1 += g.cIIr[k] * fv[mm + 1 - k];
        }
      }
    }
  }
  // first differences in place (fv[K-1] is unused; the table column is 0)
#pragma unroll
  for (int k = 0; k + 1 < K; ++k) fv[k] = fv[k + 1] - fv[k];
  const kdptr C = (kdptr)(g.ctab) + (size_t)g.ctype[p] * M * K * 2;
  const kdptr T = (kdptr)(g.rowtab) + (size_t)p * M * kRowTab;

  // ---- local solution at the partition ends (rows 0 and m-1) -> LDS
  double a0 = 0.0, a1 = 0.0, z0 = 0.0, z1 = 0.0;
  {
    const kdptr c0 = C, c1 = C + M * K;  // row 0, components 0 and 1
    const kdptr l0 = C + (m - 1) * K, l1 = C + (M + m - 1) * K;  // row m-1
#pragma unroll
    for (int k = 0; k + 1 < K; ++k) {
      a0 += c0[k] * fv[k];
      a1 += c1[k] * fv[k];
      z0 += l0[k] * fv[k];
      z1 += l1[k] * fv[k];
    }
  }
  G[(p * 4 + 0) * 64 + lane] = a0;
  G[(p * 4 + 1) * 64 + lane] = a1;
  G[(p * 4 + 2) * 64 + lane] = z0;
  G[(p * 4 + 3) * 64 + lane] = z1;
  __syncthreads();

  // ---- coupling: (left, right) end values of the neighbours = Q_p . G
  double lr0 = 0.0, lr1 = 0.0, lr2 = 0.0, lr3 = 0.0;
  {
    const int NQ = g.NQ;
    const kdptr q = (kdptr)(g.qtab) + (size_t)p * 4 * NQ * 4;
    for (int t = 0; t < NQ; ++t) {
      int qq = g.rel ? p + g.d0 + t : t;
      if (periodic) qq = qq < 0 ? qq + g.P : (qq >= g.P ? qq - g.P : qq);
      else qq = qq < 0 ? 0 : (qq >= g.P ? g.P - 1 : qq);
      const double G0 = G[(qq * 4 + 0) * 64 + lane];
      const double G1 = G[(qq * 4 + 1) * 64 + lane];
      const double G2 = G[(qq * 4 + 2) * 64 + lane];
      const double G3 = G[(qq * 4 + 3) * 64 + lane];
      const kdptr q0 = q + (0 * NQ + t) * 4;
      const kdptr q1 = q + (1 * NQ + t) * 4;
      const kdptr q2 = q + (2 * NQ + t) * 4;
      const kdptr q3 = q + (3 * NQ + t) * 4;
      lr0 += q0[0] * G0 + q0[1] * G1 + q0[2] * G2 + q0[3] * G3;
      lr1 += q1[0] * G0 + q1[1] * G1 + q1[2] * G2 + q1[3] * G3;
      lr2 += q2[0] * G0 + q2[1] * G1 + q2[2] * G2 + q2[3] * G3;
      lr3 += q3[0] * G0 + q3[1] * G1 + q3[2] * G2 + q3[3] * G3;
    }
  }

  // ---- rows: x_r = C_r . fv - V_r (lr0, lr1) - W_r (lr2, lr3), stored at once
  const int off = periodic ? 0 : 1;  // unknown r -> node s0 + r + off
  auto spike = [&](int r, double& u, double& v) {
    const kdptr t = T + r * kRowTab + 8;
    u = __builtin_fma(-t[0], lr0, u);
    u = __builtin_fma(-t[1], lr1, u);
    u = __builtin_fma(-t[4], lr2, u);
    u = __builtin_fma(-t[5], lr3, u);
    v = __builtin_fma(-t[2], lr0, v);
    v = __builtin_fma(-t[3], lr1, v);
    v = __builtin_fma(-t[6], lr2, v);
    v = __builtin_fma(-t[7], lr3, v);
  };
  // the end rows reuse their local solution from above
  spike(0, a0, a1);
  store_node(s0 + off, a0, a1, true);
  // interior rows: one scalar-table row per iteration (rows past a short
  // partition are computed from zero table rows and not stored)
#pragma unroll 1
  for (int r = 1; r < M - 1; ++r) {
    const kdptr cu = C + r * K, cv = C + (M + r) * K;
    double u = 0.0, v = 0.0;
#pragma unroll
    for (int k = 0; k + 1 < K; ++k) {
      u += cu[k] * fv[k];
      v += cv[k] * fv[k];
    }
    spike(r, u, v);
    store_node(s0 + r + off, u, v, r < m - 1);
  }
  if (m > 1) spike(m - 1, z0, z1);
  else {
    z0 = a0;
    z1 = a1;
  }
  store_node(s0 + m - 1 + off, z0, z1, m > 1);
  if (periodic && p == 0) store_node(N, a0, a1, true);  // image node N = node 0
  if (wall_first)
    store_node(0, g.Ml[0] * a0 + g.Ml[1] * a1 + bl0, g.Ml[2] * a0 + g.Ml[3] * a1 + bl1, true);
  if (wall_last)
    store_node(N, g.Mr[0] * z0 + g.Mr[1] * z1 + br0, g.Mr[2] * z0 + g.Mr[3] * z1 + br1, true);
}

template <int M, int AL, bool BUF>
__device__ __forceinline__ void ccd_lines_explicit(const CcdAxisGeom& g, const CcdLines& ln,
                                                   const double* f, double* d1, double* d2,
                                                   const unsigned block) {
  extern __shared__ double lds[];
  const int p = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const ExplicitLine<AL> L = explicit_line<AL>(ln, d1, block);
  double fv[M + 2];
  explicit_load<M, AL, BUF>(g, ln, f, L, p, fv);
  explicit_solve<M, AL, BUF>(g, ln, f, d1, d2, L, p, fv, lds);
}

template <int M, int AL, bool BUF>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8)))
void ccd_explicit_kernel(CcdLaunch L) {
  int sl;
  unsigned local;
  slot_of(L, blockIdx.x, sl, local);
  const CcdSlot& S = L.slot[sl];
  local = xcd_local(local, L.block_begin[sl + 1] - L.block_begin[sl]);
  ccd_lines_explicit<M, AL, BUF>(S.g, S.ln, S.f, S.d1, S.d2, local);
}



// Batched launch over many (problem, axis) slots kept in device memory:
// workgroups [begin[s], begin[s+1]) belong to slot s.  All slots share the
// rows cap M and the partition count (one block size per launch).  The
// per-call array pointers of a group of <= kArgSlots slots travel in the
// kernel arguments (BatchPtrs), so a call copies nothing to the device;
// larger groups read them from the device slot table.
constexpr int kArgSlots = 128;
struct BatchPtrs {
  int n;  // 0: pointers from the slot table
  const double* f[kArgSlots];
  double* d1[kArgSlots];
  double* d2[kArgSlots];
};

__device__ __forceinline__ int batch_slot(const unsigned* __restrict__ begin, int nslots,
                                          unsigned blk) {
  int lo = 0, hi = nslots - 1;  // largest s with begin[s] <= blk (uniform)
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (begin[mid] <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int M>
__global__ __launch_bounds__(1024) void ccd_batch_kernel(const CcdSlot* __restrict__ slots,
                                                         const unsigned* __restrict__ begin,
                                                         int nslots, BatchPtrs P) {
  const unsigned blk = blockIdx.x;
  const int lo = batch_slot(begin, nslots, blk);
  const CcdSlot& S = slots[lo];
  ccd_lines<M>(S.g, S.ln, S.contig != 0, S.tiled != 0, P.n ? P.f[lo] : S.f,
               P.n ? P.d1[lo] : S.d1, P.n ? P.d2[lo] : S.d2, blk - begin[lo]);
}

template <int M>
__global__ __launch_bounds__(64 * kWaveLines) void ccd_wave_batch_kernel(
    const CcdSlot* __restrict__ slots, const unsigned* __restrict__ begin, int nslots,
    BatchPtrs P) {
  const unsigned blk = blockIdx.x;
  const int lo = batch_slot(begin, nslots, blk);
  const CcdSlot& S = slots[lo];
  const double* f = P.n ? P.f[lo] : S.f;
  double* d1 = P.n ? P.d1[lo] : S.d1;
  double* d2 = P.n ? P.d2[lo] : S.d2;
  if (S.g.uniform)
    ccd_wave_tile<M, true, 1>(S.g, S.ln, f, d1, d2, blk - begin[lo]);
  else
    ccd_wave_tile<M, false, 1>(S.g, S.ln, f, d1, d2, blk - begin[lo]);
}

template <int M, bool BUF>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8)))
void ccd_explicit_batch_kernel(const CcdSlot* __restrict__ slots, const unsigned* __restrict__ begin,
                               int nslots, BatchPtrs P) {
  const unsigned blk = blockIdx.x;
  const int lo = batch_slot(begin, nslots, blk);
  const CcdSlot& S = slots[lo];
  ccd_lines_explicit<M, 0, BUF>(S.g, S.ln, P.n ? P.f[lo] : S.f, P.n ? P.d1[lo] : S.d1,
                                P.n ? P.d2[lo] : S.d2,
                                xcd_local(blk - begin[lo], begin[lo + 1] - begin[lo]));
}

using BatchFn = void (*)(const CcdSlot*, const unsigned*, int, BatchPtrs);
using WaveFn = void (*)(CcdLaunch);

BatchFn pick_wave_batch_kernel(int M) {
  switch (M) {
    case 8: return ccd_wave_batch_kernel<8>;
    case 16: return ccd_wave_batch_kernel<16>;
  }
  return nullptr;
}

BatchFn pick_explicit_batch_kernel(int M, bool buf) {
  switch (M) {
    case 8: return buf ? ccd_explicit_batch_kernel<8, true> : ccd_explicit_batch_kernel<8, false>;
    case 16: return buf ? ccd_explicit_batch_kernel<16, true> : ccd_explicit_batch_kernel<16, false>;
  }
  return nullptr;
}

// Whether the explicit kernel may address a slot's arrays with 32-bit buffer
// offsets (every array of the slot < 2 GiB).
bool slot_buffer_ok(const CcdSlot& S) {
  return S.ln.n_lines * (int64_t)(S.g.N + 1) * 8 < ((int64_t)1 << 31);
}

BatchFn pick_batch_kernel(int M) {
  switch (M) {
    case 8: return ccd_batch_kernel<8>;
    case 16: return ccd_batch_kernel<16>;
    case 32: return ccd_batch_kernel<32>;
    case 64: return ccd_batch_kernel<64>;
  }
  return nullptr;
}

using KernelFn = void (*)(CcdLaunch);

WaveFn pick_wave_kernel(int M, int lpw, bool ref = false) {
  if (ref) {
    if (M == 8) return lpw == 4 ? ccd_wave_kernel_ref<8, 4> : lpw == 2 ? ccd_wave_kernel_ref<8, 2>
                                                                       : ccd_wave_kernel_ref<8, 1>;
    if (M == 16) return lpw == 4 ? ccd_wave_kernel_ref<16, 4>
                                 : lpw == 2 ? ccd_wave_kernel_ref<16, 2> : ccd_wave_kernel_ref<16, 1>;
    return nullptr;
  }
  if (M == 8) return lpw == 4 ? ccd_wave_kernel<8, 4> : lpw == 2 ? ccd_wave_kernel<8, 2>
                                                                 : ccd_wave_kernel<8, 1>;
  if (M == 16) return lpw == 4 ? ccd_wave_kernel<16, 4> : lpw == 2 ? ccd_wave_kernel<16, 2>
                                                                   : ccd_wave_kernel<16, 1>;
  return nullptr;
}

KernelFn pick_kernel(int M, bool ref = false) {
  if (ref) {
    switch (M) {
      case 8: return ccd_line_kernel<8, true>;
      case 16: return ccd_line_kernel<16, true>;
    }
    return nullptr;
  }
  switch (M) {
    case 8: return ccd_line_kernel<8, false>;
    case 16: return ccd_line_kernel<16, false>;
    case 32: return ccd_line_kernel<32, false>;
    case 64: return ccd_line_kernel<64, false>;
  }
  return nullptr;
}

// Contiguous lines are staged as a [64][nax] tile when it fits LDS and the
// per-thread staging registers (kTilePer).
bool tile_fits(const CcdAxisGeom& g, int64_t nax) {
  return g.M == 16 && 64 * nax <= (int64_t)kTilePer * 64 * g.P && 64 * nax * 8 <= 150 * 1024;
}

size_t lds_bytes(const CcdAxisGeom& g, bool contig, int64_t nax) {
  const bool tiled = contig && tile_fits(g, nax);
  const bool tab = g.M <= 32;  // matches kLdsTab in the kernel
  if (tiled) {
    size_t t = sizeof(double) * 64 * (size_t)nax;
    size_t st = sizeof(double) * g.P * 64 * kStagePitch;
    size_t gb = sizeof(double) * ((tab ? (size_t)g.P * g.M * kRowTab + (size_t)g.P * 16 * g.NQ : 0) +
                                  (size_t)4 * g.P * 64);
    return std::max(t, std::max(st, gb));
  }
  size_t gbytes = sizeof(double) * ((tab ? (size_t)g.P * g.M * kRowTab + (size_t)g.P * 16 * g.NQ : 0) +
                                    (size_t)4 * g.P * 64);
  size_t sbytes = contig ? sizeof(double) * g.P * 64 * kStagePitch : 0;
  return gbytes > sbytes ? gbytes : sbytes;
}

CcdLines lines_for(const tpf_ccd_plan* plan, int axis) {
  const int nd = plan->ndim;
  int64_t stride = 1;
  for (int a = axis + 1; a < nd; ++a) stride *= plan->n_nodes[a];
  int64_t outer_count = 1;
  for (int a = 0; a < axis; ++a) outer_count *= plan->n_nodes[a];
  const int64_t n_ax = plan->n_nodes[axis];
  CcdLines ln{};
  ln.stride = stride;
  if (stride == 1) {  // last axis: line b starts at b * n_ax
    ln.n_lines = outer_count;
    ln.inner = outer_count;
    ln.outer = 0;
    ln.inner_stride = n_ax;
  } else {  // line b = (outer o, inner i): start o*n_ax*stride + i
    ln.n_lines = outer_count * stride;
    ln.inner = stride;
    ln.outer = n_ax * stride;
    ln.inner_stride = 1;
  }
  return ln;
}

std::mutex g_attr_mu;
// (kernel, bytes) pairs whose dynamic-LDS limit has been raised already
struct AttrSet {
  const void* k[16];
  size_t bytes[16];
  int n = 0;
};
AttrSet g_attr;

tpf_status raise_lds_limit(const void* k, size_t bytes) {
  std::lock_guard<std::mutex> lk(g_attr_mu);
  for (int i = 0; i < g_attr.n; ++i)
    if (g_attr.k[i] == k && g_attr.bytes[i] >= bytes) return TPF_OK;
  TPF_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  for (int i = 0; i < g_attr.n; ++i)
    if (g_attr.k[i] == k) {
      g_attr.bytes[i] = bytes;
      return TPF_OK;
    }
  if (g_attr.n < 16) {
    g_attr.k[g_attr.n] = k;
    g_attr.bytes[g_attr.n++] = bytes;
  }
  return TPF_OK;
}

// Launches the given axes in one grid (all must share the rows cap M).
tpf_status launch_axes(tpf_ccd_plan* plan, const int* axes, int naxes, const double* const* f,
                       double* const* d1, double* const* d2, hipStream_t stream,
                       bool force_lane = false) {
  CcdLaunch L{};
  L.nslots = naxes;
  size_t shmem = 0;
  unsigned blocks = 0;
  int M = -1;
  const bool wave = !force_lane && plan->axes[axes[0]].wave == 1;
  for (int i = 0; i < naxes; ++i) {
    CcdAxisHost& ax = plan->axes[axes[i]];
    CcdSlot& S = L.slot[i];
    S.g = ax.geom;
    S.ln = lines_for(plan, axes[i]);
    S.contig = S.ln.stride == 1;
    S.tiled = !wave && S.contig && tile_fits(ax.geom, plan->n_nodes[axes[i]]);
    S.f = f[i];
    S.d1 = d1[i];
    S.d2 = d2[i];
    if (M < 0) M = ax.geom.M;
    if (ax.geom.M != M ||
        (!force_lane && (ax.geom.P != plan->axes[axes[0]].geom.P ||
                         (ax.wave == 1) != (plan->axes[axes[0]].wave == 1))))
      return fail(TPF_ERR_VALUE, "fused CCD launch needs one kernel, rows cap and partition count");
    const size_t sh = wave ? 0 : lds_bytes(ax.geom, S.contig, plan->n_nodes[axes[i]]);
    shmem = sh > shmem ? sh : shmem;
    if (S.ln.n_lines > (int64_t)0x3fffffff)
      return fail(TPF_ERR_VALUE, "too many CCD lines for one launch");
    L.block_begin[i] = blocks;
    const CcdAxisHost& a0 = plan->axes[axes[0]];
    const int64_t per = wave ? (int64_t)a0.lpw * a0.wpg : 64;
    blocks += (unsigned)((S.ln.n_lines + per - 1) / per);
  }
  L.block_begin[naxes] = blocks;
  if (wave) {
    const CcdAxisHost& a0 = plan->axes[axes[0]];
    WaveFn k = pick_wave_kernel(M, a0.lpw, plan->precision == TPF_PREC_FP32_REFINED);
    if (!k) return fail(TPF_ERR_VALUE, "unsupported CCD rows cap " + std::to_string(M));
    const bool refined = plan->precision == TPF_PREC_FP32_REFINED;
    size_t sh = 0;
    for (int i = 0; i < naxes; ++i)
      sh = std::max(sh, refined ? wave_lds_bytes_ref(L.slot[i].g, a0.lpw * a0.wpg, a0.wpg)
                                : wave_lds_bytes(L.slot[i].g, a0.lpw * a0.wpg, a0.wpg));
    TPF_TRY(raise_lds_limit((const void*)k, sh));
    hipLaunchKernelGGL(k, dim3(blocks), dim3(64 * a0.wpg), sh, stream, L);
    TPF_HIP(hipGetLastError());
    return TPF_OK;
  }
  bool all_explicit = !force_lane;  // explicit-mode axes fused with lane-mode ones run as lane
  for (int i = 0; i < naxes; ++i) all_explicit = all_explicit && plan->axes[axes[i]].wave == 2;
  if (all_explicit) {  // strided axes, explicit partition inverse
    if (M != 8 && M != 16) return fail(TPF_ERR_VALUE, "explicit CCD kernel takes rows cap 8 or 16");
    // aligned store ownership: one chunk size per launch (the first axis')
    const int al = plan->store_align[axes[0]] < 0 ? 8 : plan->store_align[axes[0]];
    const int own = al ? 64 - al : 64;
    blocks = 0;
    for (int i = 0; i < naxes; ++i) {
      L.block_begin[i] = blocks;
      blocks += (unsigned)((L.slot[i].ln.n_lines + al + own - 1) / own);
    }
    L.block_begin[naxes] = blocks;
    bool buf = true;
    for (int i = 0; i < naxes; ++i) buf = buf && slot_buffer_ok(L.slot[i]);
    KernelFn k = nullptr;
    if (buf) {
      if (M == 16) k = al == 16 ? ccd_explicit_kernel<16, 16, true>
                   : al == 8    ? ccd_explicit_kernel<16, 8, true>
                                : ccd_explicit_kernel<16, 0, true>;
      else k = al == 16 ? ccd_explicit_kernel<8, 16, true>
             : al == 8  ? ccd_explicit_kernel<8, 8, true>
                        : ccd_explicit_kernel<8, 0, true>;
    } else {
      if (M == 16) k = al == 16 ? ccd_explicit_kernel<16, 16, false>
                   : al == 8    ? ccd_explicit_kernel<16, 8, false>
                                : ccd_explicit_kernel<16, 0, false>;
      else k = al == 16 ? ccd_explicit_kernel<8, 16, false>
             : al == 8  ? ccd_explicit_kernel<8, 8, false>
                        : ccd_explicit_kernel<8, 0, false>;
    }
    const size_t sh = sizeof(double) * 4 * L.slot[0].g.P * 64;
    TPF_TRY(raise_lds_limit((const void*)k, sh));
    hipLaunchKernelGGL(k, dim3(blocks), dim3(64 * L.slot[0].g.P), sh, stream, L);
    TPF_HIP(hipGetLastError());
    return TPF_OK;
  }
  const bool ref = plan->precision == TPF_PREC_FP32_REFINED;
  KernelFn k = pick_kernel(M, ref);
  if (!k) return fail(TPF_ERR_VALUE, "unsupported CCD rows cap " + std::to_string(M));
  TPF_TRY(raise_lds_limit((const void*)k, shmem));
  hipLaunchKernelGGL(k, dim3(blocks), dim3(64 * L.slot[0].g.P), shmem, stream, L);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

// One grid for the given axes when they share a launch shape (kernel, rows
// cap, partition count).  Otherwise: on large grids whose strided axes take
// the explicit kernel and whose other axes take the wavefront kernel, one
// explicit grid for the strided axes and then the others (257^3 periodic:
// 0.30 ms against 0.35 ms for one lane-kernel grid of all three; two streams
// measured slower; with a wall contiguous axis, which runs the tiled lane
// kernel, the split measured slower); else, when the axes share one rows cap
// M <= 16 and <= 16 partitions, one lane-kernel grid for all of them (the
// lane kernel's tables serve every axis); else one launch per axis.
constexpr int64_t kSplitFusedNodes = (int64_t)1 << 22;

// The kernel each axis of one tpf_ccd_apply call runs (0 lane recurrences,
// 1 wavefront per line, 2 explicit inverse) under launch_fused's rules.
void fused_kinds(const tpf_ccd_plan* plan, const int* axes, int n, int* kind) {
  bool same = true, lane_ok = true, rest_wave = true;
  int ne = 0;
  for (int i = 0; i < n; ++i) {
    const CcdAxisHost& x = plan->axes[axes[i]];
    const CcdAxisHost& x0 = plan->axes[axes[0]];
    kind[i] = x.wave;
    same = same && x.geom.M == x0.geom.M && x.geom.P == x0.geom.P &&
           (x.wave == 1) == (x0.wave == 1) && (x.wave == 2) == (x0.wave == 2);
    lane_ok = lane_ok && x.geom.M == x0.geom.M && x.geom.M <= 16 && x.geom.P <= kMaxParts &&
              x.geom.P == x0.geom.P;
    if (x.wave == 2) ++ne;
    else rest_wave = rest_wave && x.wave == 1;
  }
  if (n <= 1 || same) return;
  int64_t total = 1;
  for (int a = 0; a < plan->ndim; ++a) total *= plan->n_nodes[a];
  const bool refined = plan->precision == TPF_PREC_FP32_REFINED;
  if (ne > 0 && rest_wave && total >= kSplitFusedNodes && !refined) return;
  if (lane_ok && !refined)
    for (int i = 0; i < n; ++i) kind[i] = 0;
}

tpf_status launch_fused(tpf_ccd_plan* plan, const int* axes, int n, const double* const* f,
                        double* const* d1, double* const* d2, hipStream_t stream) {
  bool same = true;
  for (int i = 1; i < n; ++i)
    same = same && plan->axes[axes[i]].geom.M == plan->axes[axes[0]].geom.M &&
           plan->axes[axes[i]].geom.P == plan->axes[axes[0]].geom.P &&
           (plan->axes[axes[i]].wave == 1) == (plan->axes[axes[0]].wave == 1) &&
           (plan->axes[axes[i]].wave == 2) == (plan->axes[axes[0]].wave == 2);
  if (same) return launch_axes(plan, axes, n, f, d1, d2, stream);
  int64_t total = 1;
  for (int a = 0; a < plan->ndim; ++a) total *= plan->n_nodes[a];
  int ex[3], ne = 0, rest[3], nr = 0;
  bool rest_wave = true;  // the other axes take the wavefront kernel
  for (int i = 0; i < n; ++i) {
    if (plan->axes[axes[i]].wave == 2 && (ne == 0 || (plan->axes[axes[i]].geom.M ==
                                                       plan->axes[axes[ex[0]]].geom.M &&
                                                   plan->axes[axes[i]].geom.P ==
                                                       plan->axes[axes[ex[0]]].geom.P))) {
      ex[ne++] = i;
    } else {
      rest[nr++] = i;
      rest_wave = rest_wave && plan->axes[axes[i]].wave == 1;
    }
  }
  if (ne > 0 && rest_wave && total >= kSplitFusedNodes &&
      plan->precision != TPF_PREC_FP32_REFINED) {
    int a1[3], a2[3];
    const double *f1[3], *f2[3];
    double *o11[3], *o12[3], *o21[3], *o22[3];
    for (int i = 0; i < ne; ++i) {
      a1[i] = axes[ex[i]];
      f1[i] = f[ex[i]];
      o11[i] = d1[ex[i]];
      o12[i] = d2[ex[i]];
    }
    for (int i = 0; i < nr; ++i) {
      a2[i] = axes[rest[i]];
      f2[i] = f[rest[i]];
      o21[i] = d1[rest[i]];
      o22[i] = d2[rest[i]];
    }
    TPF_TRY(launch_axes(plan, a1, ne, f1, o11, o12, stream));
    return nr ? launch_fused(plan, a2, nr, f2, o21, o22, stream) : TPF_OK;
  }
  if (n > 1 && plan->precision != TPF_PREC_FP32_REFINED) {
    bool lane_ok = true;
    for (int i = 0; i < n; ++i) {
      const CcdAxisGeom& g = plan->axes[axes[i]].geom;
      lane_ok = lane_ok && g.M == plan->axes[axes[0]].geom.M && g.M <= 16 && g.P <= kMaxParts &&
                g.P == plan->axes[axes[0]].geom.P;
    }
    if (lane_ok) return launch_axes(plan, axes, n, f, d1, d2, stream, true);
  }
  for (int i = 0; i < n; ++i)
    TPF_TRY(launch_axes(plan, &axes[i], 1, &f[i], &d1[i], &d2[i], stream));
  return TPF_OK;
}

// ---- grid metric of non-uniform axes (ccd_solver_helpers.py:509-519) ------
// On a non-uniform axis the line solves run in the computational coordinate
// xi (spacing 1/N) and the metric maps the results to x at node j of the
// axis: d1 = J d1_xi, d2 = (J J) d2_xi + (J dJ) d1_xi, with J = dxi/dx and
// dJ = dJ/dxi.  The products J J and J dJ are formed on the host, as the
// reference forms them elementwise.  In place on d1 / d2; xi1 holds d1_xi
// (d1 itself, or scratch when only d2 is wanted).  One read and one write of
// each output per node: the uniform path never runs it.
__global__ __launch_bounds__(256) void k_ccd_metric(int64_t total, int64_t stride, int n,
                                                    const double* __restrict__ m,
                                                    const double* xi1, double* d1, double* d2) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += step) {
    const int j = (int)((i / stride) % n);
    const double a = xi1[i];
    if (d2) d2[i] = m[n + j] * d2[i] + m[2 * (int64_t)n + j] * a;
    if (d1) d1[i] = m[j] * a;
  }
}

// The line solves of `axes` (fused launch policy when `fused`), then the
// metric of every non-uniform one.  A d2-only slot of a non-uniform axis
// solves d1_xi into the plan's scratch (grown on demand; calls of one plan
// on different streams that use it race).
tpf_status run_axes(tpf_ccd_plan* plan, bool fused, const int* axes, int n,
                    const double* const* f, double* const* d1, double* const* d2,
                    hipStream_t stream) {
  int64_t total = 1;
  for (int a = 0; a < plan->ndim; ++a) total *= plan->n_nodes[a];
  double* x1[3];
  int need = 0;
  bool any = false;
  for (int i = 0; i < n; ++i) {
    x1[i] = d1[i];
    if (plan->d_metric[axes[i]] && (d1[i] || d2[i])) {
      any = true;
      if (!d1[i]) ++need;
    }
  }
  if (need) {
    if (plan->scratch_n < need * total) {
      if (plan->scratch) TPF_HIP(hipFree(plan->scratch));
      plan->scratch = nullptr;
      plan->scratch_n = 0;
      TPF_HIP(hipMalloc(&plan->scratch, sizeof(double) * (size_t)(need * total)));
      plan->scratch_n = need * total;
    }
    int k = 0;
    for (int i = 0; i < n; ++i)
      if (plan->d_metric[axes[i]] && !d1[i] && d2[i]) x1[i] = plan->scratch + (k++) * total;
  }
  if (fused) TPF_TRY(launch_fused(plan, axes, n, f, x1, d2, stream));
  else TPF_TRY(launch_axes(plan, axes, n, f, x1, d2, stream));
  if (!any) return TPF_OK;
  const int blocks = (int)std::min<int64_t>((total + 255) / 256, 8192);
  for (int i = 0; i < n; ++i) {
    const int a = axes[i];
    if (!plan->d_metric[a] || !(d1[i] || d2[i])) continue;
    int64_t stride = 1;
    for (int b = a + 1; b < plan->ndim; ++b) stride *= plan->n_nodes[b];
    hipLaunchKernelGGL(k_ccd_metric, dim3(blocks), dim3(256), 0, stream, total, stride,
                       (int)plan->n_nodes[a], (const double*)plan->d_metric[a],
                       (const double*)x1[i], d1[i], d2[i]);
    TPF_HIP(hipGetLastError());
  }
  return TPF_OK;
}

}  // namespace

bool ccd_plan_has_metric(const tpf_ccd_plan* plan) {
  if (!plan) return false;
  for (int a = 0; a < 3; ++a)
    if (plan->d_metric[a]) return true;
  return false;
}

}  // namespace tpf

using namespace tpf;

extern "C" tpf_status tpf_ccd_plan_create(tpf_ccd_plan** out, int ndim, const int64_t* n_nodes,
                                          const double* h, const int* bc, int precision) {
  if (!out || !n_nodes || !h || !bc) return fail(TPF_ERR_VALUE, "null argument");
  *out = nullptr;
  if (ndim < 1 || ndim > 3) return fail(TPF_ERR_VALUE, "ndim must be 1, 2 or 3");
  if (precision != TPF_PREC_FP64 && precision != TPF_PREC_FP32_REFINED)
    return fail(TPF_ERR_VALUE, "unknown precision mode");
  auto* plan = new tpf_ccd_plan();
  plan->ndim = ndim;
  plan->precision = precision;
  for (int a = 0; a < ndim; ++a) {
    plan->n_nodes[a] = n_nodes[a];
    plan->h[a] = h[a];
    plan->bc[a] = bc[a];
    if (n_nodes[a] < 1 || n_nodes[a] > (1 << 20) || !(h[a] > 0.0)) {
      delete plan;
      return fail(TPF_ERR_VALUE, "bad axis size or spacing");
    }
  }
  *out = plan;
  return TPF_OK;
}

namespace tpf {
// Lazily builds and uploads the tables of one axis (the reference builds
// wall solvers lazily per used axis, ccd_solver.py:108-111, 294-299).
// Kernel of one axis: a wavefront per line when the lines are few and long
// (auto: < 4096 lines of > 256 unknowns, e.g. 2-D 512^2 and 1024^2 grids,
// where the lane-per-line kernel needs 32 or 64 rows per partition: 2-D
// 513^2 wall, 0.031-0.035 ms per axis against 0.070-0.081), else a lane per
// line.  The wavefront kernel uses 8 rows per partition when the line fits
// 64 partitions, else 16.
tpf_status build_axis(tpf_ccd_plan* plan, int axis, CcdAxisHost& ax) {
  const int n_ax = (int)plan->n_nodes[axis];
  const bool periodic = plan->bc[axis] == TPF_BC_PERIODIC;
  const int n = periodic ? n_ax - 1 : n_ax - 2;
  int64_t total = 1;
  for (int a = 0; a < plan->ndim; ++a) total *= plan->n_nodes[a];
  const int64_t lines = total / n_ax;
  const int mode = plan->mode[axis];
  const int cap = plan->rows_cap[axis];
  const bool ref = plan->precision == TPF_PREC_FP32_REFINED;
  if (ref && (mode == 3 || cap > 16))
    return fail(TPF_ERR_VALUE, "FP32-refined CCD runs the lane-per-line or wavefront kernel with "
                               "rows cap <= 16");
  const bool strided = lines_for(plan, axis).stride != 1;
  if (mode == 3) {  // lane per line, explicit partition inverse (strided axes)
    if (!strided)
      return fail(TPF_ERR_VALUE, "the explicit-inverse CCD kernel runs strided axes only");
    if (cap != 0 && cap != 8 && cap != 16)
      return fail(TPF_ERR_VALUE, "the explicit-inverse CCD kernel takes rows cap 8 or 16");
    TPF_TRY(build_axis_tables(ax, n_ax, plan->h[axis], periodic, cap, 0));
    if (ax.geom.M > kMaxExplicitRows || ax.geom.P > kMaxParts)
      return fail(TPF_ERR_VALUE, "the explicit-inverse CCD kernel supports lines of <= 256 unknowns "
                                 "(16 partitions x 16 rows); this axis has " + std::to_string(n));
    ax.wave = 2;
    return TPF_OK;
  }
  // Auto picks the wavefront kernel for few long lines, and for many periodic
  // lines of 129-256 unknowns on the contiguous axis (four lines of 16
  // partitions per wavefront: 0.117-0.124 against 0.163 ms for the tiled
  // lane kernel on 257^3; on 129^3 the tiled lane kernel is faster, 0.027
  // against 0.075 ms, and on wall lines everywhere; FP32-refined plans keep
  // the lane kernel, whose fused 3-axis grid is 0.6 ms on 257^3 against
  // 0.8 ms for per-axis launches with the refined wavefront kernel).
  const bool contig_many = !strided && periodic && !ref && lines >= 4096 && n > 128 && n <= 256;
  bool wave = mode == 2 || (mode == 0 && ((lines < 4096 && n > 256) || contig_many ||
                                           (ref && n > 256)));
  if (wave && cap != 0 && cap != 8 && cap != 16)
    return fail(TPF_ERR_VALUE, "the wavefront-per-line CCD kernel takes rows cap 8 or 16");
  if (wave && n >= 1) {
    const int M = cap > 0 ? cap : (mode == 0 && contig_many ? 16
                                   : ((n + 7) / 8 <= kMaxWaveParts ? 8 : 16));
    const int P = (n + M - 1) / M;
    if (P > kMaxWaveParts) {
      if (mode == 2) return fail(TPF_ERR_VALUE, "line too long for 64 partitions of the rows cap");
      wave = false;
    } else {
      TPF_TRY(build_axis_tables(ax, n_ax, plan->h[axis], periodic, M, P));
    }
  }
  if (!wave) {
    TPF_TRY(build_axis_tables(ax, n_ax, plan->h[axis], periodic, cap, 0));
    if (ax.geom.P > kMaxParts)
      return fail(TPF_ERR_VALUE, "rows cap " + std::to_string(cap) +
                                     " needs more than 16 partitions for the lane-per-line kernel");
    if (ref && ax.geom.M > 16)
      return fail(TPF_ERR_VALUE, "FP32-refined CCD supports lines of <= 1024 unknowns (64 "
                                 "partitions x 16 rows); this axis has " + std::to_string(n));
  }
  // Wavefront kernel shape: as many lines per wavefront as its partitions
  // leave lanes for, four wavefronts per workgroup.
  // One line per wavefront on >= 1024 strided lines (1025^2 axis 0): 16
  // wavefronts per workgroup, so each node's tile row is 16 lines = 128 B
  // (8 and 16 measured alike, 4 at 32 B rows 1.3x slower); contiguous lines
  // keep 4 so that four times as many CUs take part (1025^2 axis 1: 0.021
  // against 0.027 ms periodic, 0.035 against 0.048 ms wall).
  if (wave) {
    ax.lpw = ax.geom.P <= 16 ? 4 : (ax.geom.P <= 32 ? 2 : 1);
    // FP32-refined: 4 wavefronts (the tile stays live next to 4 KB of exchange
    // buffers per wavefront: <= 51 KB at 1025 nodes)
    ax.wpg = (ax.lpw == 1 && !ref && lines >= 1024 && strided) ? 16 : 4;
  }
  // Auto: strided FP64 axes of <= 16 partitions x 16 rows take the explicit
  // partition inverse (no recurrence state: two workgroups per CU).
  const bool expl = !wave && mode == 0 && !ref && strided && ax.geom.M <= kMaxExplicitRows &&
                    ax.geom.P <= kMaxParts && !ax.ctab.empty();
  ax.wave = wave ? 1 : (expl ? 2 : 0);
  return TPF_OK;
}

tpf_status ensure_axis(tpf_ccd_plan* plan, int axis) {
  CcdAxisHost& ax = plan->axes[axis];
  if (ax.d_rowtab) return TPF_OK;
  TPF_TRY(build_axis(plan, axis, ax));
  TPF_HIP(hipMalloc(&ax.d_rowtab, ax.rowtab.size() * sizeof(double)));
  TPF_HIP(hipMalloc(&ax.d_qtab, ax.qtab.size() * sizeof(double)));
  TPF_HIP(hipMemcpy(ax.d_rowtab, ax.rowtab.data(), ax.rowtab.size() * sizeof(double),
                    hipMemcpyHostToDevice));
  TPF_HIP(hipMemcpy(ax.d_qtab, ax.qtab.data(), ax.qtab.size() * sizeof(double),
                    hipMemcpyHostToDevice));
  ax.geom.rowtab = ax.d_rowtab;
  ax.geom.qtab = ax.d_qtab;
  TPF_HIP(hipMalloc(&ax.d_rowtab_u, ax.rowtab_u.size() * sizeof(double)));
  TPF_HIP(hipMemcpy(ax.d_rowtab_u, ax.rowtab_u.data(), ax.rowtab_u.size() * sizeof(double),
                    hipMemcpyHostToDevice));
  ax.geom.rowtab_u = ax.d_rowtab_u;
  if (ax.wave == 1 && !ax.geom.uniform) {
    const int P = ax.geom.P, W = 4 * ax.geom.NQ * 4;
    std::vector<double> qt((size_t)P * W);
    for (int p = 0; p < P; ++p)
      for (int i = 0; i < W; ++i) qt[(size_t)i * P + p] = ax.qtab[(size_t)p * W + i];
    TPF_HIP(hipMalloc(&ax.d_qtab_t, qt.size() * sizeof(double)));
    TPF_HIP(hipMemcpy(ax.d_qtab_t, qt.data(), qt.size() * sizeof(double), hipMemcpyHostToDevice));
    ax.geom.qtab_t = ax.d_qtab_t;
  }
  if (ax.wave == 2) {
    TPF_HIP(hipMalloc(&ax.d_ctab, ax.ctab.size() * sizeof(double)));
    TPF_HIP(hipMemcpy(ax.d_ctab, ax.ctab.data(), ax.ctab.size() * sizeof(double),
                      hipMemcpyHostToDevice));
    ax.geom.ctab = ax.d_ctab;
  }
  if (plan->precision == TPF_PREC_FP32_REFINED) {
    std::vector<float> r32(ax.rowtab.begin(), ax.rowtab.end()), q32(ax.qtab.begin(), ax.qtab.end());
    TPF_HIP(hipMalloc(&ax.d_rowtab32, r32.size() * sizeof(float)));
    TPF_HIP(hipMalloc(&ax.d_qtab32, q32.size() * sizeof(float)));
    TPF_HIP(hipMemcpy(ax.d_rowtab32, r32.data(), r32.size() * sizeof(float), hipMemcpyHostToDevice));
    TPF_HIP(hipMemcpy(ax.d_qtab32, q32.data(), q32.size() * sizeof(float), hipMemcpyHostToDevice));
    ax.geom.rowtab32 = ax.d_rowtab32;
    ax.geom.qtab32 = ax.d_qtab32;
  }
  return TPF_OK;
}
}  // namespace tpf

extern "C" tpf_status tpf_ccd_plan_destroy(tpf_ccd_plan* plan) {
  if (!plan) return TPF_OK;
  for (auto& ax : plan->axes) {
    if (ax.d_rowtab) (void)hipFree(ax.d_rowtab);
    if (ax.d_rowtab_u) (void)hipFree(ax.d_rowtab_u);
    if (ax.d_qtab) (void)hipFree(ax.d_qtab);
    if (ax.d_qtab_t) (void)hipFree(ax.d_qtab_t);
    if (ax.d_rowtab32) (void)hipFree(ax.d_rowtab32);
    if (ax.d_qtab32) (void)hipFree(ax.d_qtab32);
    if (ax.d_ctab) (void)hipFree(ax.d_ctab);
  }
  for (double* m : plan->d_metric)
    if (m) (void)hipFree(m);
  if (plan->scratch) (void)hipFree(plan->scratch);
  delete plan;
  return TPF_OK;
}

extern "C" tpf_status tpf_ccd_plan_set_metric(tpf_ccd_plan* plan, int axis, const double* J,
                                              const double* dJ) {
  if (!plan || axis < 0 || axis >= plan->ndim) return fail(TPF_ERR_VALUE, "bad plan/axis");
  if (!J != !dJ) return fail(TPF_ERR_VALUE, "J and dJ must both be given or both be NULL");
  if (!J) {
    if (plan->d_metric[axis]) TPF_HIP(hipFree(plan->d_metric[axis]));
    plan->d_metric[axis] = nullptr;
    return TPF_OK;
  }
  const int64_t n = plan->n_nodes[axis];
  std::vector<double> m((size_t)(3 * n));
  for (int64_t j = 0; j < n; ++j) {
    if (!std::isfinite(J[j]) || !std::isfinite(dJ[j]))
      return fail(TPF_ERR_VALUE, "grid metric must be finite");
    m[(size_t)j] = J[j];
    m[(size_t)(n + j)] = J[j] * J[j];
    m[(size_t)(2 * n + j)] = J[j] * dJ[j];
  }
  if (!plan->d_metric[axis])
    TPF_HIP(hipMalloc(&plan->d_metric[axis], sizeof(double) * (size_t)(3 * n)));
  TPF_HIP(hipMemcpy(plan->d_metric[axis], m.data(), sizeof(double) * m.size(),
                    hipMemcpyHostToDevice));
  return TPF_OK;
}

extern "C" tpf_status tpf_ccd_differentiate(tpf_ccd_plan* plan, const double* f, double* d1,
                                            double* d2, int axis, void* stream) {
  if (!plan || !f) return fail(TPF_ERR_VALUE, "null plan or input");
  if (axis < 0 || axis >= plan->ndim) return fail(TPF_ERR_VALUE, "axis out of range");
  if (!d1 && !d2) return TPF_OK;
  if ((const double*)d1 == f || (const double*)d2 == f || (d1 && d1 == d2))
    return fail(TPF_ERR_VALUE, "CCD outputs must not alias the input or each other");
  TPF_TRY(ensure_axis(plan, axis));
  return run_axes(plan, false, &axis, 1, &f, &d1, &d2, (hipStream_t)stream);
}

extern "C" tpf_status tpf_ccd_apply(tpf_ccd_plan* plan, const double* f, double* d1, double* d2,
                                    uint32_t axis_mask, void* stream) {
  if (!plan || !f) return fail(TPF_ERR_VALUE, "null plan or input");
  if (axis_mask >> plan->ndim) return fail(TPF_ERR_VALUE, "axis_mask names a missing axis");
  int64_t total = 1;
  for (int a = 0; a < plan->ndim; ++a) total *= plan->n_nodes[a];
  int axes[3], n = 0;
  const double* fs[3];
  double *o1[3], *o2[3];
  for (int a = 0; a < plan->ndim; ++a) {
    if (!(axis_mask & (1u << a))) continue;
    TPF_TRY(ensure_axis(plan, a));
    o1[n] = d1 ? d1 + n * total : nullptr;
    o2[n] = d2 ? d2 + n * total : nullptr;
    if ((const double*)o1[n] == f || (const double*)o2[n] == f)
      return fail(TPF_ERR_VALUE, "CCD outputs must not alias the input");
    fs[n] = f;
    axes[n++] = a;
  }
  if (n == 0 || (!d1 && !d2)) return TPF_OK;
  return run_axes(plan, true, axes, n, fs, o1, o2, (hipStream_t)stream);
}

extern "C" tpf_status tpf_ccd_apply_fields(tpf_ccd_plan* plan, const double* const* f,
                                           double* const* d1, double* const* d2,
                                           uint32_t axis_mask, void* stream) {
  if (!plan || !f) return fail(TPF_ERR_VALUE, "null plan or inputs");
  if (axis_mask >> plan->ndim) return fail(TPF_ERR_VALUE, "axis_mask names a missing axis");
  int axes[3], n = 0;
  const double* fs[3];
  double *o1[3], *o2[3];
  for (int a = 0; a < plan->ndim; ++a) {
    if (!(axis_mask & (1u << a))) continue;
    if (!f[a]) return fail(TPF_ERR_VALUE, "null input for a masked axis");
    TPF_TRY(ensure_axis(plan, a));
    fs[n] = f[a];
    o1[n] = d1 ? d1[a] : nullptr;
    o2[n] = d2 ? d2[a] : nullptr;
    axes[n++] = a;
  }
  // the slots of one grid run concurrently: no output may be any slot's
  // input or another output
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j)
      if ((const double*)o1[i] == fs[j] || (const double*)o2[i] == fs[j])
        return fail(TPF_ERR_VALUE, "CCD outputs must not alias any input");
    for (int j = 0; j < n; ++j) {
      if (o1[i] && ((j != i && o1[i] == o1[j]) || o1[i] == o2[j]))
        return fail(TPF_ERR_VALUE, "CCD outputs must be distinct arrays");
      if (o2[i] && j != i && o2[i] == o2[j])
        return fail(TPF_ERR_VALUE, "CCD outputs must be distinct arrays");
    }
  }
  bool any = false;
  for (int i = 0; i < n; ++i) any = any || o1[i] || o2[i];
  if (n == 0 || !any) return TPF_OK;
  return run_axes(plan, true, axes, n, fs, o1, o2, (hipStream_t)stream);
}

extern "C" tpf_status tpf_ccd_axis_info(tpf_ccd_plan* plan, int axis, int* partitions,
                                        int* max_rows, int* window, double* trunc_coeff) {
  if (!plan || axis < 0 || axis >= plan->ndim) return fail(TPF_ERR_VALUE, "bad plan/axis");
  CcdAxisHost& ax = plan->axes[axis];
  if (ax.rowtab.empty()) TPF_TRY(build_axis(plan, axis, ax));
  if (partitions) *partitions = ax.geom.P;
  if (max_rows) *max_rows = ax.geom.M;
  if (window) *window = ax.geom.NQ;
  if (trunc_coeff) *trunc_coeff = ax.trunc_coeff;
  return TPF_OK;
}

extern "C" tpf_status tpf_ccd_axis_kernel(tpf_ccd_plan* plan, int axis, int* kernel) {
  if (!plan || !kernel || axis < 0 || axis >= plan->ndim) return fail(TPF_ERR_VALUE, "bad plan/axis");
  CcdAxisHost& ax = plan->axes[axis];
  if (ax.rowtab.empty()) TPF_TRY(build_axis(plan, axis, ax));
  *kernel = ax.wave == 1 ? 2 : (ax.wave == 2 ? 3 : 1);
  return TPF_OK;
}

// Test hook (host only, no device): runs the scalar model of the kernel on
// host arrays so the SPIKE tables can be checked without a GPU.
extern "C" tpf_status tpf_test_ccd_host_model(tpf_ccd_plan* plan, const double* f, double* d1,
                                              double* d2, int axis) {
  if (!plan || !f || axis < 0 || axis >= plan->ndim) return fail(TPF_ERR_VALUE, "bad args");
  CcdAxisHost& ax = plan->axes[axis];
  if (ax.rowtab.empty()) TPF_TRY(build_axis(plan, axis, ax));
  const CcdLines ln = lines_for(plan, axis);
  for (int64_t b = 0; b < ln.n_lines; ++b) {
    const int64_t base = (b / ln.inner) * ln.outer + (b % ln.inner) * ln.inner_stride;
    (ax.wave == 2 ? host_model_line_explicit : host_model_line)(
        ax, f + base, ln.stride, d1 ? d1 + base : nullptr, d2 ? d2 + base : nullptr);
  }
  return TPF_OK;
}

// Test hook (host only): the workgroup -> (slot, slot-local block) map of a
// fused launch whose slots have counts[0..nslots) workgroups.
extern "C" tpf_status tpf_test_ccd_block_map(int nslots, const unsigned* counts,
                                             unsigned strided_mask, int* slot, unsigned* local) {
  if (nslots < 1 || nslots > 3 || !counts || !slot || !local) return fail(TPF_ERR_VALUE, "bad args");
  CcdLaunch L{};
  L.nslots = nslots;
  unsigned total = 0;
  for (int i = 0; i < nslots; ++i) {
    L.block_begin[i] = total;
    total += counts[i];
  }
  L.block_begin[nslots] = total;
  for (unsigned b = 0; b < total; ++b) {
    slot_of(L, b, slot[b], local[b]);
    if (strided_mask >> slot[b] & 1u) local[b] = xcd_local(local[b], counts[slot[b]]);
  }
  return TPF_OK;
}

extern "C" tpf_status tpf_ccd_plan_set_launch(tpf_ccd_plan* plan, int axis, int mode,
                                              int rows_cap) {
  if (!plan || axis < 0 || axis >= plan->ndim) return fail(TPF_ERR_VALUE, "bad plan/axis");
  if (mode < 0 || mode > 3) return fail(TPF_ERR_VALUE, "mode must be 0 (auto), 1, 2 or 3");
  if (rows_cap != 0 && rows_cap != 8 && rows_cap != 16 && rows_cap != 32 && rows_cap != 64)
    return fail(TPF_ERR_VALUE, "rows cap must be 0 (default), 8, 16, 32 or 64");
  // validate on the host before dropping the current tables
  const int old_mode = plan->mode[axis], old_cap = plan->rows_cap[axis];
  plan->mode[axis] = mode;
  plan->rows_cap[axis] = rows_cap;
  CcdAxisHost trial;
  const tpf_status st = build_axis(plan, axis, trial);
  if (st != TPF_OK) {
    plan->mode[axis] = old_mode;
    plan->rows_cap[axis] = old_cap;
    return st;
  }
  CcdAxisHost& ax = plan->axes[axis];
  if (ax.d_rowtab) TPF_HIP(hipFree(ax.d_rowtab));
  if (ax.d_rowtab_u) TPF_HIP(hipFree(ax.d_rowtab_u));
  if (ax.d_qtab) TPF_HIP(hipFree(ax.d_qtab));
  if (ax.d_qtab_t) TPF_HIP(hipFree(ax.d_qtab_t));
  if (ax.d_rowtab32) TPF_HIP(hipFree(ax.d_rowtab32));
  if (ax.d_qtab32) TPF_HIP(hipFree(ax.d_qtab32));
  if (ax.d_ctab) TPF_HIP(hipFree(ax.d_ctab));
  ax = CcdAxisHost();
  return TPF_OK;
}

extern "C" tpf_status tpf_ccd_plan_set_store_align(tpf_ccd_plan* plan, int axis, int align) {
  if (!plan || axis < 0 || axis >= plan->ndim) return fail(TPF_ERR_VALUE, "bad plan/axis");
  if (align != -1 && align != 0 && align != 8 && align != 16)
    return fail(TPF_ERR_VALUE, "store alignment must be -1 (default), 0, 8 or 16 elements");
  plan->store_align[axis] = align;
  return TPF_OK;
}

// ---- batches of independent problems ---------------------------------------
struct tpf_ccd_batch {
  struct Group {
    int M, P, wave;  // wave: 0 lane recurrences, 1 wavefront per line, 2 explicit inverse
    bool buf;          // explicit inverse: every slot's arrays fit 32-bit buffer offsets
    int first, count;  // slots [first, first + count)
    unsigned blocks;
    size_t shmem;
  };
  int nproblems = 0;
  std::vector<tpf::CcdSlot> slots;      // geometry; pointers filled per call
  std::vector<int> slot_problem, slot_k;
  std::vector<int64_t> totals;          // nodes per problem
  std::vector<unsigned> begin;          // per group: count + 1 entries
  std::vector<int> begin_off;           // group -> offset into begin
  std::vector<Group> groups;
  tpf::CcdSlot* d_slots = nullptr;
  unsigned* d_begin = nullptr;
  tpf::CcdSlot* h_slots = nullptr;      // pinned staging
  hipEvent_t staged = nullptr;          // the last upload of h_slots
};

extern "C" tpf_status tpf_ccd_batch_create(tpf_ccd_batch** out, tpf_ccd_plan* const* plans,
                                           const uint32_t* axis_masks, int count) {
  if (!out || !plans || !axis_masks || count < 1) return fail(TPF_ERR_VALUE, "bad batch arguments");
  *out = nullptr;
  auto* b = new tpf_ccd_batch();
  b->nproblems = count;
  struct Item {
    int problem, k, axis, kind;  // kind: the kernel tpf_ccd_apply would run for this axis
  };
  std::vector<Item> items;
  for (int j = 0; j < count; ++j) {
    tpf_ccd_plan* pl = plans[j];
    if (!pl || (axis_masks[j] >> pl->ndim)) {
      delete b;
      return fail(TPF_ERR_VALUE, "batch entry " + std::to_string(j) + ": bad plan or axis mask");
    }
    if (tpf::ccd_plan_has_metric(pl)) {
      delete b;
      return fail(TPF_ERR_VALUE, "batches take uniform-grid plans (entry " + std::to_string(j) +
                                     " has a grid metric)");
    }
    if (pl->precision != TPF_PREC_FP64) {
      delete b;
      return fail(TPF_ERR_VALUE, "batches take FP64 plans (entry " + std::to_string(j) + ")");
    }
    int64_t total = 1;
    for (int a = 0; a < pl->ndim; ++a) total *= pl->n_nodes[a];
    b->totals.push_back(total);
    int k = 0;
    const size_t first = items.size();
    for (int a = 0; a < pl->ndim; ++a) {
      if (!(axis_masks[j] & (1u << a))) continue;
      const tpf_status st = ensure_axis(pl, a);
      if (st != TPF_OK) {
        delete b;
        return st;
      }
      items.push_back({j, k++, a, pl->axes[a].wave});
    }
    // the kernels tpf_ccd_apply runs these axes with (bit-identical results)
    int ax[3], kinds[3];
    const int na = (int)(items.size() - first);
    for (int i = 0; i < na; ++i) ax[i] = items[first + i].axis;
    fused_kinds(pl, ax, na, kinds);
    for (int i = 0; i < na; ++i) items[first + i].kind = kinds[i];
  }
  // group by kernel and launch shape (M, P), keeping the input order inside a group
  struct Shape {
    int wave, M, P;
    bool operator==(const Shape& o) const { return wave == o.wave && M == o.M && P == o.P; }
  };
  std::vector<Shape> shapes;
  for (const Item& it : items) {
    const CcdAxisHost& ax = plans[it.problem]->axes[it.axis];
    const Shape sh{it.kind, ax.geom.M, ax.geom.P};
    if (std::find(shapes.begin(), shapes.end(), sh) == shapes.end()) shapes.push_back(sh);
  }
  for (const auto& sh : shapes) {
    tpf_ccd_batch::Group G{sh.M, sh.P, sh.wave, true, (int)b->slots.size(), 0, 0, 0};
    b->begin_off.push_back((int)b->begin.size());
    unsigned blocks = 0;
    for (const Item& it : items) {
      tpf_ccd_plan* pl = plans[it.problem];
      const CcdAxisHost& ax = pl->axes[it.axis];
      const CcdAxisGeom& g = ax.geom;
      if (!(Shape{it.kind, g.M, g.P} == sh)) continue;
      CcdSlot S{};
      S.g = g;
      S.ln = lines_for(pl, it.axis);
      S.contig = S.ln.stride == 1;
      S.tiled = sh.wave == 0 && S.contig && tile_fits(g, pl->n_nodes[it.axis]);
      b->slots.push_back(S);
      b->slot_problem.push_back(it.problem);
      b->slot_k.push_back(it.k);
      b->begin.push_back(blocks);
      const int64_t per = sh.wave == 1 ? kWaveLines : 64;
      blocks += (unsigned)((S.ln.n_lines + per - 1) / per);
      if (sh.wave == 0) G.shmem = std::max(G.shmem, lds_bytes(g, S.contig, pl->n_nodes[it.axis]));
      else if (sh.wave == 1) G.shmem = std::max(G.shmem, wave_lds_bytes(g, kWaveLines, kWaveLines));
      else G.shmem = std::max(G.shmem, sizeof(double) * 4 * (size_t)g.P * 64);
      G.buf = G.buf && slot_buffer_ok(S);
      G.count += 1;
    }
    b->begin.push_back(blocks);
    G.blocks = blocks;
    b->groups.push_back(G);
  }
  const size_t ns = b->slots.size();
  hipError_t e = hipMalloc(&b->d_slots, sizeof(CcdSlot) * ns);
  if (e == hipSuccess) e = hipMalloc(&b->d_begin, sizeof(unsigned) * b->begin.size());
  if (e == hipSuccess) e = hipHostMalloc(&b->h_slots, sizeof(CcdSlot) * ns);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->staged, hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_begin, b->begin.data(), sizeof(unsigned) * b->begin.size(),
                  hipMemcpyHostToDevice);
  // geometry of every slot; the array pointers come per call (kernel
  // arguments, or a restaged table for groups above kArgSlots)
  if (e == hipSuccess)
    e = hipMemcpy(b->d_slots, b->slots.data(), sizeof(CcdSlot) * ns, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    tpf_ccd_batch_destroy(b);
    return fail(TPF_ERR_RUNTIME, std::string("batch allocation: ") + hipGetErrorString(e));
  }
  std::copy(b->slots.begin(), b->slots.end(), b->h_slots);
  for (const auto& G : b->groups) {
    const void* k = G.wave == 1   ? (const void*)pick_wave_batch_kernel(G.M)
                    : G.wave == 2 ? (const void*)pick_explicit_batch_kernel(G.M, G.buf)
                                  : (const void*)pick_batch_kernel(G.M);
    if (G.shmem == 0) continue;
    const tpf_status st = raise_lds_limit(k, G.shmem);
    if (st != TPF_OK) {
      tpf_ccd_batch_destroy(b);
      return st;
    }
  }
  *out = b;
  return TPF_OK;
}

extern "C" tpf_status tpf_ccd_batch_destroy(tpf_ccd_batch* b) {
  if (!b) return TPF_OK;
  if (b->staged) {
    (void)hipEventSynchronize(b->staged);
    (void)hipEventDestroy(b->staged);
  }
  if (b->d_slots) (void)hipFree(b->d_slots);
  if (b->d_begin) (void)hipFree(b->d_begin);
  if (b->h_slots) (void)hipHostFree(b->h_slots);
  delete b;
  return TPF_OK;
}

extern "C" tpf_status tpf_ccd_batch_apply(tpf_ccd_batch* b, const double* const* f,
                                          double* const* d1, double* const* d2, void* stream) {
  if (!b || !f || (!d1 && !d2)) return fail(TPF_ERR_VALUE, "null batch argument");
  hipStream_t st = (hipStream_t)stream;
  const size_t ns = b->slots.size();
  std::vector<const double*> sf(ns);
  std::vector<double*> s1(ns), s2(ns);
  bool table = false;  // some group reads its pointers from the device slot table
  for (const auto& G : b->groups) table = table || G.count > kArgSlots;
  for (size_t i = 0; i < ns; ++i) {
    const int j = b->slot_problem[i], k = b->slot_k[i];
    if (!f[j]) return fail(TPF_ERR_VALUE, "batch entry " + std::to_string(j) + ": null input");
    double* o1 = (d1 && d1[j]) ? d1[j] + (int64_t)k * b->totals[j] : nullptr;
    double* o2 = (d2 && d2[j]) ? d2[j] + (int64_t)k * b->totals[j] : nullptr;
    if ((const double*)o1 == f[j] || (const double*)o2 == f[j])
      return fail(TPF_ERR_VALUE, "CCD outputs must not alias the input");
    sf[i] = f[j];
    s1[i] = o1;
    s2[i] = o2;
  }
  if (table) {
    TPF_HIP(hipEventSynchronize(b->staged));  // the previous upload has left h_slots
    for (size_t i = 0; i < ns; ++i) {
      b->h_slots[i].f = sf[i];
      b->h_slots[i].d1 = s1[i];
      b->h_slots[i].d2 = s2[i];
    }
    TPF_HIP(hipMemcpyAsync(b->d_slots, b->h_slots, sizeof(CcdSlot) * ns, hipMemcpyHostToDevice, st));
    TPF_HIP(hipEventRecord(b->staged, st));
  }
  for (size_t gi = 0; gi < b->groups.size(); ++gi) {
    const auto& G = b->groups[gi];
    BatchPtrs P;
    P.n = 0;
    if (G.count <= kArgSlots) {
      P.n = G.count;
      for (int i = 0; i < G.count; ++i) {
        P.f[i] = sf[G.first + i];
        P.d1[i] = s1[G.first + i];
        P.d2[i] = s2[G.first + i];
      }
    }
    const CcdSlot* slots = b->d_slots + G.first;
    const unsigned* begin = b->d_begin + b->begin_off[gi];
    if (G.wave == 1)
      hipLaunchKernelGGL(pick_wave_batch_kernel(G.M), dim3(G.blocks), dim3(64 * kWaveLines),
                         G.shmem, st, slots, begin, G.count, P);
    else
      hipLaunchKernelGGL(G.wave == 2 ? pick_explicit_batch_kernel(G.M, G.buf) : pick_batch_kernel(G.M),
                         dim3(G.blocks), dim3(64 * G.P), G.shmem, st, slots, begin, G.count, P);
    TPF_HIP(hipGetLastError());
  }
  return TPF_OK;
}
