// CCD line-solve kernels for gfx950 and the C-ABI entry points that launch
// them (replace ccd/ccd_solver_helpers.py:38-77 and :426-465).
//
// Mapping: one workgroup = 64 lines x P partitions.  Lane = line, wavefront
// = partition, so every table value a wavefront needs is uniform and comes
// through the scalar cache.  A lane keeps its partition (<= M rows, 2 FP64
// unknowns per row) in registers: RHS -> local block LU forward/back sweep
// -> partition end values to LDS -> coupling (precomputed reduced-system
// rows) -> spike correction -> store.
//
// Strided axes: lanes read/write consecutive addresses (coalesced).
// Contiguous (last) axis: loads hit the same cache lines per lane row and
// are absorbed by L1; stores are re-staged through LDS per wavefront so
// each store instruction writes whole 128 B segments.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "ccd_plan.hpp"
#include "../../include/tpf_testing.h"

namespace tpf {

namespace {

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
  const bool kLdsTab = M <= 32;
  const int NQ = g.NQ;
  double* ltab = lds;
  double* lq = lds + (kLdsTab ? g.P * M * kRowTab : 0);
  double* G = lq + (kLdsTab ? g.P * 16 * NQ : 0);  // [4P][64]
  if (kLdsTab) {
    const int nt = g.P * M * kRowTab, nq = g.P * 16 * NQ;
    for (int i = threadIdx.x; i < nt; i += blockDim.x) ltab[i] = g.rowtab[i];
    for (int i = threadIdx.x; i < nq; i += blockDim.x) lq[i] = g.qtab[i];
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
  const double* tab = (kLdsTab ? ltab : g.rowtab) + (size_t)p * M * kRowTab;
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
    const double* q = (kLdsTab ? lq : g.qtab) + (size_t)p * 4 * NQ * 4;
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

template <int M>
__global__ __launch_bounds__(1024) void ccd_line_kernel(CcdLaunch L) {
  // Block order = slot order (axis 0 lines first).  An XCD-contiguous
  // remapping of blocks was measured slower for the fused 3-axis grid.
  const unsigned blk = blockIdx.x;
  const int sl = (L.nslots > 1 && blk >= L.block_begin[1]) + (L.nslots > 2 && blk >= L.block_begin[2]);
  const CcdSlot& S = L.slot[sl];
  ccd_lines<M>(S.g, S.ln, S.contig != 0, S.tiled != 0, S.f, S.d1, S.d2, blk - L.block_begin[sl]);
}

using KernelFn = void (*)(CcdLaunch);

KernelFn pick_kernel(int M) {
  switch (M) {
    case 8: return ccd_line_kernel<8>;
    case 16: return ccd_line_kernel<16>;
    case 32: return ccd_line_kernel<32>;
    case 64: return ccd_line_kernel<64>;
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
tpf_status launch_axes(tpf_ccd_plan* plan, const int* axes, int naxes, const double* f,
                       double* const* d1, double* const* d2, hipStream_t stream) {
  CcdLaunch L{};
  L.nslots = naxes;
  size_t shmem = 0;
  unsigned blocks = 0;
  int M = -1;
  for (int i = 0; i < naxes; ++i) {
    CcdAxisHost& ax = plan->axes[axes[i]];
    CcdSlot& S = L.slot[i];
    S.g = ax.geom;
    S.ln = lines_for(plan, axes[i]);
    S.contig = S.ln.stride == 1;
    S.tiled = S.contig && tile_fits(ax.geom, plan->n_nodes[axes[i]]);
    S.f = f;
    S.d1 = d1[i];
    S.d2 = d2[i];
    if (M < 0) M = ax.geom.M;
    if (ax.geom.M != M || ax.geom.P != plan->axes[axes[0]].geom.P)
      return fail(TPF_ERR_VALUE, "fused CCD launch needs one rows cap and partition count");
    const size_t sh = lds_bytes(ax.geom, S.contig, plan->n_nodes[axes[i]]);
    shmem = sh > shmem ? sh : shmem;
    if (S.ln.n_lines > (int64_t)0x3fffffff)
      return fail(TPF_ERR_VALUE, "too many CCD lines for one launch");
    L.block_begin[i] = blocks;
    blocks += (unsigned)((S.ln.n_lines + 63) / 64);
  }
  L.block_begin[naxes] = blocks;
  KernelFn k = pick_kernel(M);
  if (!k) return fail(TPF_ERR_VALUE, "unsupported CCD rows cap " + std::to_string(M));
  TPF_TRY(raise_lds_limit((const void*)k, shmem));
  hipLaunchKernelGGL(k, dim3(blocks), dim3(64 * L.slot[0].g.P), shmem, stream, L);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

}  // namespace

}  // namespace tpf

using namespace tpf;

extern "C" tpf_status tpf_ccd_plan_create(tpf_ccd_plan** out, int ndim, const int64_t* n_nodes,
                                          const double* h, const int* bc, int precision) {
  if (!out || !n_nodes || !h || !bc) return fail(TPF_ERR_VALUE, "null argument");
  *out = nullptr;
  if (ndim < 1 || ndim > 3) return fail(TPF_ERR_VALUE, "ndim must be 1, 2 or 3");
  if (precision == TPF_PREC_FP32_REFINED)
    return fail(TPF_ERR_VALUE, "FP32-refined CCD is not built; use TPF_PREC_FP64");
  if (precision != TPF_PREC_FP64) return fail(TPF_ERR_VALUE, "unknown precision mode");
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
tpf_status ensure_axis(tpf_ccd_plan* plan, int axis) {
  CcdAxisHost& ax = plan->axes[axis];
  if (ax.d_rowtab) return TPF_OK;
  TPF_TRY(build_axis_tables(ax, (int)plan->n_nodes[axis], plan->h[axis],
                            plan->bc[axis] == TPF_BC_PERIODIC, 0, 0));
  TPF_HIP(hipMalloc(&ax.d_rowtab, ax.rowtab.size() * sizeof(double)));
  TPF_HIP(hipMalloc(&ax.d_qtab, ax.qtab.size() * sizeof(double)));
  TPF_HIP(hipMemcpy(ax.d_rowtab, ax.rowtab.data(), ax.rowtab.size() * sizeof(double),
                    hipMemcpyHostToDevice));
  TPF_HIP(hipMemcpy(ax.d_qtab, ax.qtab.data(), ax.qtab.size() * sizeof(double),
                    hipMemcpyHostToDevice));
  ax.geom.rowtab = ax.d_rowtab;
  ax.geom.qtab = ax.d_qtab;
  return TPF_OK;
}
}  // namespace tpf

extern "C" tpf_status tpf_ccd_plan_destroy(tpf_ccd_plan* plan) {
  if (!plan) return TPF_OK;
  for (auto& ax : plan->axes) {
    if (ax.d_rowtab) (void)hipFree(ax.d_rowtab);
    if (ax.d_qtab) (void)hipFree(ax.d_qtab);
  }
  delete plan;
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
  return launch_axes(plan, &axis, 1, f, &d1, &d2, (hipStream_t)stream);
}

extern "C" tpf_status tpf_ccd_apply(tpf_ccd_plan* plan, const double* f, double* d1, double* d2,
                                    uint32_t axis_mask, void* stream) {
  if (!plan || !f) return fail(TPF_ERR_VALUE, "null plan or input");
  if (axis_mask >> plan->ndim) return fail(TPF_ERR_VALUE, "axis_mask names a missing axis");
  int64_t total = 1;
  for (int a = 0; a < plan->ndim; ++a) total *= plan->n_nodes[a];
  int axes[3], n = 0;
  double *o1[3], *o2[3];
  for (int a = 0; a < plan->ndim; ++a) {
    if (!(axis_mask & (1u << a))) continue;
    TPF_TRY(ensure_axis(plan, a));
    o1[n] = d1 ? d1 + n * total : nullptr;
    o2[n] = d2 ? d2 + n * total : nullptr;
    if ((const double*)o1[n] == f || (const double*)o2[n] == f)
      return fail(TPF_ERR_VALUE, "CCD outputs must not alias the input");
    axes[n++] = a;
  }
  if (n == 0 || (!d1 && !d2)) return TPF_OK;
  bool same = true;  // one grid when every axis has the same launch shape
  for (int i = 1; i < n; ++i)
    same = same && plan->axes[axes[i]].geom.M == plan->axes[axes[0]].geom.M &&
           plan->axes[axes[i]].geom.P == plan->axes[axes[0]].geom.P;
  if (same) return launch_axes(plan, axes, n, f, o1, o2, (hipStream_t)stream);
  for (int i = 0; i < n; ++i)
    TPF_TRY(launch_axes(plan, &axes[i], 1, f, &o1[i], &o2[i], (hipStream_t)stream));
  return TPF_OK;
}

extern "C" tpf_status tpf_ccd_axis_info(tpf_ccd_plan* plan, int axis, int* partitions,
                                        int* max_rows, int* window, double* trunc_coeff) {
  if (!plan || axis < 0 || axis >= plan->ndim) return fail(TPF_ERR_VALUE, "bad plan/axis");
  CcdAxisHost& ax = plan->axes[axis];
  if (ax.rowtab.empty())
    TPF_TRY(build_axis_tables(ax, (int)plan->n_nodes[axis], plan->h[axis],
                              plan->bc[axis] == TPF_BC_PERIODIC, 0, 0));
  if (partitions) *partitions = ax.geom.P;
  if (max_rows) *max_rows = ax.geom.M;
  if (window) *window = ax.geom.NQ;
  if (trunc_coeff) *trunc_coeff = ax.trunc_coeff;
  return TPF_OK;
}

// Test hook (host only, no device): runs the scalar model of the kernel on
// host arrays so the SPIKE tables can be checked without a GPU.
extern "C" tpf_status tpf_test_ccd_host_model(tpf_ccd_plan* plan, const double* f, double* d1,
                                              double* d2, int axis) {
  if (!plan || !f || axis < 0 || axis >= plan->ndim) return fail(TPF_ERR_VALUE, "bad args");
  CcdAxisHost& ax = plan->axes[axis];
  if (ax.rowtab.empty())
    TPF_TRY(build_axis_tables(ax, (int)plan->n_nodes[axis], plan->h[axis],
                              plan->bc[axis] == TPF_BC_PERIODIC, 0, 0));
  const CcdLines ln = lines_for(plan, axis);
  for (int64_t b = 0; b < ln.n_lines; ++b) {
    const int64_t base = (b / ln.inner) * ln.outer + (b % ln.inner) * ln.inner_stride;
    host_model_line(ax, f + base, ln.stride, d1 ? d1 + base : nullptr, d2 ? d2 + base : nullptr);
  }
  return TPF_OK;
}
