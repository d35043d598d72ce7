// Operator-split CLS reinitialisation (SplitReinitializer,
// levelset/reinit_split.py:22-102; levelset/reinit_ops.py:36-204):
// n_steps pseudo-time steps of
//   compression  q* = clip(q - dtau div[q (1 - q) n], 0, 1),  n = grad q / max(sqrt(max(|grad q|^2, 1e-12)), 1e-6),
//                div per axis = the CCD derivative of the flux + eps_d (g+ - 2 g + g-) with
//                the level-set boundary's ghosts (core/boundary.py:249-293);
//   diffusion    per axis (M2 - mu B2) q_new = (M2 + mu B2) q, M2 = tridiag(-1/8, 1, -1/8),
//                B2 = 3 tridiag(1, -2, 1), mirrored end rows, mu = eps dtau / (2 h^2);
//   then clip to [0, 1]; periodic images follow their roots after every stage;
// then the interface-weighted mass correction (levelset/heaviside.py:118-149).
// dtau = min(dx^2 / (4 ndim eps), dx / 2) (reinit_ops.py:36-51).
//
// The host loop below drives the CCD plan (gradient and flux derivatives)
// and three kernels per step: the compression flux, the filtered divergence
// with the clipped update, and one Thomas sweep per line and axis (one lane
// per line, walked in chunks; constant factors computed once on the host as
// the reference's CPU path does).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "reduce.hpp"
#include "tpf_common.hpp"

extern "C" tpf_status tpf_sync_periodic_images(int ndim, const int64_t* n_nodes,
                                               const int* periodic, const double* in, double* out,
                                               void* stream);
extern "C" tpf_status tpf_dccd_mass_correct(int ndim, const int64_t* n_nodes, const double* L,
                                            const int* periodic, const double* q,
                                            double m_target, double* out, double* mass_before,
                                            void* stream);
extern "C" tpf_status tpf_fields_mass(int ndim, const int64_t* n_nodes, const double* L,
                                      const int* bc, const double* q, double* mass, void* stream);

namespace tpf {
namespace {

constexpr double kBeta2 = -1.0 / 8.0;  // M2 off-diagonal (ccd_solver.py:74)

struct RGeom {
  int n[3];
  int bc[3];  // TPF_REINIT_BC_* of the slot
  int k[3];
  int64_t total;
};

__device__ __forceinline__ int64_t stride_of(const RGeom& g, int a) {
  return a == 0 ? (int64_t)g.n[1] * g.n[2] : (a == 1 ? (int64_t)g.n[2] : 1);
}

__device__ __forceinline__ void coords(const RGeom& g, int64_t idx, int (&c)[3]) {
  const uint32_t plane = (uint32_t)g.n[1] * (uint32_t)g.n[2];
  c[0] = (int)((uint32_t)idx / plane);
  const uint32_t rem = (uint32_t)idx - (uint32_t)c[0] * plane;
  c[1] = (int)(rem / (uint32_t)g.n[2]);
  c[2] = (int)(rem - (uint32_t)c[1] * (uint32_t)g.n[2]);
}

__device__ __forceinline__ int64_t root_of(const RGeom& g, int64_t idx, int (&c)[3]) {
  int64_t r = idx;
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (g.bc[a] == TPF_REINIT_BC_PERIODIC && g.n[a] > 1 && c[a] == g.n[a] - 1) {
      r -= (int64_t)c[a] * stride_of(g, a);
      c[a] = 0;
    }
  return r;
}

// flux_a = q (1 - q) g_a / safe_grad (reinit_ops.py:54-75, 106-114)
__global__ __launch_bounds__(kBlock) void k_reinit_flux(RGeom g, int ndim,
                                                        const double* __restrict__ q,
                                                        const double* __restrict__ d1,
                                                        double* __restrict__ flux) {
  const int64_t T = g.total;
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < T;
       idx += (int64_t)gridDim.x * kBlock) {
    double gs = 0.0;
    for (int a = 0; a < ndim; ++a) gs += d1[a * T + idx] * d1[a * T + idx];
    const double safe = fmax(sqrt(fmax(gs, 1e-12)), 1e-6);
    const double w = q[idx] * (1.0 - q[idx]);
    for (int a = 0; a < ndim; ++a) flux[a * T + idx] = w * (d1[a * T + idx] / safe);
  }
}

// out = clip(q - dtau sum_a filtered(fp_a), 0, 1) at the root of every node
__global__ __launch_bounds__(kBlock) void k_reinit_compress(RGeom g, const double* __restrict__ q,
                                                            const double* __restrict__ fp,
                                                            double eps_d, double dtau,
                                                            double* __restrict__ out) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords(g, idx, c);
    const int64_t r = root_of(g, idx, c);
    double div = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (g.n[a] == 1) continue;
      const int N = g.n[a] - 1;
      const int i = c[a];
      const int64_t st = stride_of(g, a);
      const double* fa = fp + (int64_t)g.k[a] * g.total;
      const double f0 = fa[r];
      double fm, fpl;
      if (i > 0) fm = fa[r - st];
      else if (g.bc[a] == TPF_REINIT_BC_PERIODIC) fm = fa[r + (int64_t)(N - 1) * st];
      else if (g.bc[a] == TPF_REINIT_BC_ZERO) fm = 0.0;
      else fm = f0;  // neumann (mirror) and outflow (repeat) with one ghost
      if (i < N) fpl = fa[r + st];
      else if (g.bc[a] == TPF_REINIT_BC_PERIODIC) fpl = fa[r - (int64_t)(N - 1) * st];
      else if (g.bc[a] == TPF_REINIT_BC_ZERO) fpl = 0.0;
      else fpl = f0;
      div = div + (f0 + eps_d * (fpl - 2.0 * f0 + fm));
    }
    out[idx] = fmin(fmax(q[r] - dtau * div, 0.0), 1.0);
  }
}

// One CN-ADI sweep along slot a: rhs from q, forward elimination into out,
// back substitution in place (reinit_ops.py:164-204, CPU Thomas path).
// One Thomas sweep per line (forward elimination with the host factors tf,
// back substitution with m and sup), 64 consecutive lines per workgroup (one
// lane each), the line walked in chunks of kCnChunk nodes so a chunk's loads
// are all in flight before its recurrence steps run.  Strided axes: the
// lanes' nodes are adjacent in memory, so each lane loads its chunk directly
// (coalesced across lanes).  Contiguous axis: the workgroup's lines are one
// contiguous block; every chunk is staged through an LDS tile with loads and
// stores that walk memory in order, and the lanes run their recurrences on
// the tile.  The forward pass parks the eliminated right-hand sides in out.
constexpr int kCnChunk = 16;
constexpr int kCnPitch = kCnChunk + 3;  // chunk + 2 halo nodes + 1 pad

__global__ __launch_bounds__(64) void k_reinit_cn(RGeom g, int a, double d_R, double c_R,
                                                  const double* __restrict__ tf,
                                                  const double* __restrict__ m,
                                                  const double* __restrict__ sup,
                                                  const double* __restrict__ q,
                                                  double* __restrict__ out) {
  __shared__ double tile[64 * kCnPitch];
  const int n = g.n[a];
  const int64_t st = stride_of(g, a);
  const int64_t nlines = g.total / n;
  const int lane = threadIdx.x;
  const int64_t l0 = (int64_t)blockIdx.x * 64;
  const int nl = (int)(nlines - l0 < 64 ? nlines - l0 : 64);
  const bool active = lane < nl;
  const int64_t l = l0 + (active ? lane : 0);
  int64_t base;
  if (a == 0) base = l;
  else if (a == 1) base = (l / g.n[2]) * ((int64_t)g.n[1] * g.n[2]) + (l % g.n[2]);
  else base = l * n;
  const bool contig = st == 1;
  double* trow = tile + lane * kCnPitch;
  // ---- forward elimination
  double dprev = 0.0;
  for (int i0 = 0; i0 < n; i0 += kCnChunk) {
    const int cnt = n - i0 < kCnChunk ? n - i0 : kCnChunk;
    // v[k] = q at node i0 - 1 + k, k < cnt + 2 (nodes outside [0, n) unused)
    double v[kCnChunk + 2];
    if (contig) {
      __syncthreads();  // previous chunk's tile consumed
      const int w = cnt + 2;
      const double* blk = q + l0 * (int64_t)n;
      for (int e = lane; e < nl * w; e += 64) {
        const int j = e / w, k = e - j * w, node = i0 - 1 + k;
        if (node >= 0 && node < n) tile[j * kCnPitch + k] = blk[(int64_t)j * n + node];
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kCnChunk + 2; ++k) {
        const int node = i0 - 1 + k;
        v[k] = (k < cnt + 2 && node >= 0 && node < n) ? trow[k] : 0.0;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kCnChunk + 2; ++k) {
        const int node = i0 - 1 + k;
        v[k] = (active && k < cnt + 2 && node >= 0 && node < n) ? q[base + (int64_t)node * st] : 0.0;
      }
    }
    double dv[kCnChunk];
#pragma unroll
    for (int k = 0; k < kCnChunk; ++k) {
      if (k < cnt) {
        const int i = i0 + k;
        double rhs;
        if (i == 0) rhs = d_R * v[k + 1] + 2.0 * c_R * v[k + 2];
        else if (i == n - 1) rhs = 2.0 * c_R * v[k] + d_R * v[k + 1];
        else rhs = c_R * v[k] + d_R * v[k + 1] + c_R * v[k + 2];
        const double d = i == 0 ? rhs : rhs - tf[i - 1] * dprev;
        dv[k] = d;
        dprev = d;
      }
    }
    if (contig) {
      __syncthreads();  // every lane has read its q values
#pragma unroll
      for (int k = 0; k < kCnChunk; ++k)
        if (k < cnt) trow[k] = dv[k];
      __syncthreads();
      double* blk = out + l0 * (int64_t)n;
      for (int e = lane; e < nl * cnt; e += 64) {
        const int j = e / cnt, k = e - j * cnt;
        blk[(int64_t)j * n + i0 + k] = tile[j * kCnPitch + k];
      }
    } else if (active) {
#pragma unroll
      for (int k = 0; k < kCnChunk; ++k)
        if (k < cnt) out[base + (int64_t)(i0 + k) * st] = dv[k];
    }
  }
  // ---- back substitution, chunks from the end of the line
  double x = 0.0;
  const int last0 = ((n - 1) / kCnChunk) * kCnChunk;
  for (int i0 = last0; i0 >= 0; i0 -= kCnChunk) {
    const int cnt = n - i0 < kCnChunk ? n - i0 : kCnChunk;
    double dv[kCnChunk];
    if (contig) {
      __syncthreads();
      const double* blk = out + l0 * (int64_t)n;
      for (int e = lane; e < nl * cnt; e += 64) {
        const int j = e / cnt, k = e - j * cnt;
        tile[j * kCnPitch + k] = blk[(int64_t)j * n + i0 + k];
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kCnChunk; ++k) dv[k] = k < cnt ? trow[k] : 0.0;
    } else {
#pragma unroll
      for (int k = 0; k < kCnChunk; ++k)
        dv[k] = (active && k < cnt) ? out[base + (int64_t)(i0 + k) * st] : 0.0;
    }
#pragma unroll
    for (int k = kCnChunk - 1; k >= 0; --k) {
      if (k < cnt) {
        const int i = i0 + k;
        x = i == n - 1 ? dv[k] / m[n - 1] : (dv[k] - sup[i] * x) / m[i];
        dv[k] = x;
      }
    }
    if (contig) {
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kCnChunk; ++k)
        if (k < cnt) trow[k] = dv[k];
      __syncthreads();
      double* blk = out + l0 * (int64_t)n;
      for (int e = lane; e < nl * cnt; e += 64) {
        const int j = e / cnt, k = e - j * cnt;
        blk[(int64_t)j * n + i0 + k] = tile[j * kCnPitch + k];
      }
    } else if (active) {
#pragma unroll
      for (int k = 0; k < kCnChunk; ++k)
        if (k < cnt) out[base + (int64_t)(i0 + k) * st] = dv[k];
    }
  }
}

// The same sweep for few lines (2-D grids, where k_reinit_cn would run a few
// wavefronts whose recurrences wait on global memory at every chunk): a
// workgroup of 64 threads stages `lb` whole lines and the axis factors in LDS
// with coalesced loads, lanes 0..lb-1 run the recurrences of k_reinit_cn on
// LDS with the same arithmetic (rolling q values in registers, eliminated
// right-hand sides and then the solution written over the line), and the
// workgroup writes the lines back.  LDS: (3 n + lb pitch) doubles.
__global__ __launch_bounds__(64) void k_reinit_cn_lds(RGeom g, int a, int lb, int pitch,
                                                      double d_R, double c_R,
                                                      const double* __restrict__ tf,
                                                      const double* __restrict__ m,
                                                      const double* __restrict__ sup,
                                                      const double* __restrict__ q,
                                                      double* __restrict__ out) {
  extern __shared__ double lds[];
  const int n = g.n[a];
  double* ftf = lds;         // n - 1 used
  double* fm = lds + n;      // n
  double* fsup = lds + 2 * n;  // n - 1 used
  double* tile = lds + 3 * n;  // lb x pitch
  const int64_t st = stride_of(g, a);
  const int64_t nlines = g.total / n;
  const int64_t l0 = (int64_t)blockIdx.x * lb;
  const int nl = (int)(nlines - l0 < lb ? nlines - l0 : lb);
  const int t = threadIdx.x;
  for (int i = t; i < n - 1; i += 64) {
    ftf[i] = tf[i];
    fsup[i] = sup[i];
  }
  for (int i = t; i < n; i += 64) fm[i] = m[i];
  // node i of line l sits at base(l) + i * st
  auto base_of = [&](int64_t l) -> int64_t {
    if (a == 0) return l;
    if (a == 1) return (l / g.n[2]) * ((int64_t)g.n[1] * g.n[2]) + (l % g.n[2]);
    return l * n;
  };
  const int cnt = nl * n;
  if (a == 2) {  // the workgroup's lines are one contiguous block
    const double* src = q + l0 * n;
    for (int e = t; e < cnt; e += 64) {
      const int j = e / n, i = e - j * n;
      tile[j * pitch + i] = src[e];
    }
  } else {  // node-major: adjacent lines are adjacent in memory
    for (int e = t; e < cnt; e += 64) {
      const int i = e / nl, j = e - i * nl;
      tile[j * pitch + i] = q[base_of(l0 + j) + (int64_t)i * st];
    }
  }
  __syncthreads();
  if (t < nl) {
    double* row = tile + t * pitch;
    double qm = 0.0, qc = row[0], dprev = 0.0;
    for (int i = 0; i < n; ++i) {
      const double qp = i + 1 < n ? row[i + 1] : 0.0;
      double rhs;
      if (i == 0) rhs = d_R * qc + 2.0 * c_R * qp;
      else if (i == n - 1) rhs = 2.0 * c_R * qm + d_R * qc;
      else rhs = c_R * qm + d_R * qc + c_R * qp;
      const double d = i == 0 ? rhs : rhs - ftf[i - 1] * dprev;
      row[i] = d;
      dprev = d;
      qm = qc;
      qc = qp;
    }
    double x = 0.0;
    for (int i = n - 1; i >= 0; --i) {
      x = i == n - 1 ? row[i] / fm[n - 1] : (row[i] - fsup[i] * x) / fm[i];
      row[i] = x;
    }
  }
  __syncthreads();
  if (a == 2) {
    double* dst = out + l0 * n;
    for (int e = t; e < cnt; e += 64) {
      const int j = e / n, i = e - j * n;
      dst[e] = tile[j * pitch + i];
    }
  } else {
    for (int e = t; e < cnt; e += 64) {
      const int i = e / nl, j = e - i * nl;
      out[base_of(l0 + j) + (int64_t)i * st] = tile[j * pitch + i];
    }
  }
}

// Lines per workgroup of k_reinit_cn_lds (0: use k_reinit_cn).  Taken when
// k_reinit_cn would launch fewer than 256 workgroups and up to 8 lines fit
// 64 KB of LDS next to the factors.
int cn_lds_lines(int n, int64_t nlines) {
  if (nlines >= 256 * 64) return 0;
  const int pitch = n | 1;
  const int64_t room = 65536 / 8 - 3 * (int64_t)n;
  if (room < pitch) return 0;
  const int64_t lb = room / pitch;
  return (int)(lb < 8 ? lb : 8);
}

__global__ __launch_bounds__(kBlock) void k_clip01(int64_t n, const double* in, double* out) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    out[i] = fmin(fmax(in[i], 0.0), 1.0);
}

unsigned blocks_for(int64_t total) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((total + kBlock - 1) / kBlock, 2048));
}

}  // namespace
}  // namespace tpf

using namespace tpf;

extern "C" tpf_status tpf_reinit_split(tpf_ccd_plan* ccd, int ndim, const int64_t* n_nodes,
                                       const double* L, const int* bc, double eps, int n_steps,
                                       double eps_d_comp, int mass_correction,
                                       const double* psi, double* out, void* stream) {
  if (!ccd || !n_nodes || !L || !bc || !psi || !out) return fail(TPF_ERR_VALUE, "null argument");
  if (ccd_plan_has_metric(ccd)) return fail(TPF_ERR_VALUE, "uniform grids only (plan has a metric)");
  if (ndim < 1 || ndim > 3) return fail(TPF_ERR_VALUE, "ndim must be 1, 2 or 3");
  if (!(eps > 0.0) || n_steps < 0) return fail(TPF_ERR_VALUE, "need eps > 0 and n_steps >= 0");
  if (out == psi) return fail(TPF_ERR_VALUE, "psi and out must be distinct arrays");
  RGeom g{};
  for (int a = 0; a < 3; ++a) g.n[a] = 1;
  int periodic[3] = {0, 0, 0};
  double h[3] = {1.0, 1.0, 1.0};
  for (int a = 0; a < ndim; ++a) {
    const int s = 3 - ndim + a;
    if (n_nodes[a] < 3 || !(L[a] > 0.0)) return fail(TPF_ERR_VALUE, "axis needs >= 3 nodes, L > 0");
    if (bc[a] < TPF_REINIT_BC_ZERO || bc[a] > TPF_REINIT_BC_OUTFLOW)
      return fail(TPF_ERR_VALUE, "level-set bc must be zero, periodic, neumann or outflow");
    g.n[s] = (int)n_nodes[a];
    g.bc[s] = bc[a];
    g.k[s] = a;
    periodic[a] = bc[a] == TPF_REINIT_BC_PERIODIC;
    h[a] = L[a] / (double)(n_nodes[a] - 1);
  }
  g.total = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (g.total >= (int64_t)0xffffffffLL) return fail(TPF_ERR_VALUE, "grid too large (2^32 nodes)");
  hipStream_t st = (hipStream_t)stream;
  const int64_t T = g.total;
  // dtau (reinit_ops.py:36-51) and the CN factors per axis (:117-161)
  double dx = h[0];
  for (int a = 1; a < ndim; ++a) dx = std::min(dx, h[a]);
  const double dtau = std::min(0.5 * dx * dx / (2.0 * ndim * eps), 0.5 * dx);
  std::vector<double> hostf;
  std::vector<int64_t> off_tf(ndim), off_m(ndim), off_sup(ndim);
  std::vector<double> dR(ndim), cR(ndim);
  for (int a = 0; a < ndim; ++a) {
    const int n = (int)n_nodes[a];
    const double mu = eps * dtau / (2.0 * h[a] * h[a]);
    const double dL = 1.0 + 6.0 * mu, cL = kBeta2 - 3.0 * mu;
    dR[a] = 1.0 - 6.0 * mu;
    cR[a] = kBeta2 + 3.0 * mu;
    std::vector<double> sup(n - 1, cL), sub(n - 1, cL), m(n, dL), tf(n - 1);
    sup[0] = 2.0 * cL;
    sub[n - 2] = 2.0 * cL;
    for (int i = 1; i < n; ++i) {
      tf[i - 1] = sub[i - 1] / m[i - 1];
      m[i] -= tf[i - 1] * sup[i - 1];
    }
    off_tf[a] = (int64_t)hostf.size();
    hostf.insert(hostf.end(), tf.begin(), tf.end());
    off_m[a] = (int64_t)hostf.size();
    hostf.insert(hostf.end(), m.begin(), m.end());
    off_sup[a] = (int64_t)hostf.size();
    hostf.insert(hostf.end(), sup.begin(), sup.end());
  }
  double* buf = nullptr;
  const size_t nbuf = (size_t)T * (2 + 3 * (size_t)ndim) + hostf.size();
  TPF_HIP(hipMallocAsync(&buf, sizeof(double) * nbuf, st));
  double* q = buf;
  double* qs = q + T;
  double* d1 = qs + T;
  double* flux = d1 + (int64_t)ndim * T;
  double* fpd = flux + (int64_t)ndim * T;
  double* fac = fpd + (int64_t)ndim * T;
  auto run = [&]() -> tpf_status {
    TPF_HIP(hipMemcpyAsync(fac, hostf.data(), sizeof(double) * hostf.size(),
                           hipMemcpyHostToDevice, st));
    TPF_HIP(hipMemcpyAsync(q, psi, sizeof(double) * T, hipMemcpyDeviceToDevice, st));
    TPF_TRY(tpf_sync_periodic_images(ndim, n_nodes, periodic, q, q, st));
    double m_old = 0.0;
    if (mass_correction) {
      int bcs[3];
      for (int a = 0; a < ndim; ++a) bcs[a] = periodic[a] ? TPF_BC_PERIODIC : TPF_BC_WALL;
      TPF_TRY(tpf_fields_mass(ndim, n_nodes, L, bcs, q, &m_old, st));
    }
    const uint32_t mask = (1u << ndim) - 1u;
    for (int step = 0; step < n_steps; ++step) {
      TPF_TRY(tpf_ccd_apply(ccd, q, d1, nullptr, mask, st));
      hipLaunchKernelGGL(k_reinit_flux, dim3(blocks_for(T)), dim3(kBlock), 0, st, g, ndim,
                         (const double*)q, (const double*)d1, flux);
      TPF_HIP(hipGetLastError());
      {  // every axis differentiates its own flux, one launch
        const double* fa[3];
        double* fpa[3];
        for (int a = 0; a < ndim; ++a) {
          fa[a] = flux + (int64_t)a * T;
          fpa[a] = fpd + (int64_t)a * T;
        }
        TPF_TRY(tpf_ccd_apply_fields(ccd, fa, fpa, nullptr, mask, st));
      }
      hipLaunchKernelGGL(k_reinit_compress, dim3(blocks_for(T)), dim3(kBlock), 0, st, g,
                         (const double*)q, (const double*)fpd, eps_d_comp, dtau, qs);
      TPF_HIP(hipGetLastError());
      // CN-ADI sweeps, ping-pong qs <-> q, images after every sweep
      double* src = qs;
      double* dst = q;
      for (int a = 0; a < ndim; ++a) {
        const int s = 3 - ndim + a;
        const int64_t nlines = T / g.n[s];
        const int lb = cn_lds_lines(g.n[s], nlines);
        if (lb > 0) {
          const int pitch = g.n[s] | 1;
          const size_t sh = sizeof(double) * (3 * (size_t)g.n[s] + (size_t)lb * pitch);
          hipLaunchKernelGGL(k_reinit_cn_lds, dim3((unsigned)((nlines + lb - 1) / lb)), dim3(64),
                             sh, st, g, s, lb, pitch, dR[a], cR[a],
                             (const double*)(fac + off_tf[a]), (const double*)(fac + off_m[a]),
                             (const double*)(fac + off_sup[a]), (const double*)src, dst);
        } else {
          hipLaunchKernelGGL(k_reinit_cn, dim3((unsigned)((nlines + 63) / 64)), dim3(64), 0, st,
                             g, s, dR[a], cR[a], (const double*)(fac + off_tf[a]),
                             (const double*)(fac + off_m[a]), (const double*)(fac + off_sup[a]),
                             (const double*)src, dst);
        }
        TPF_HIP(hipGetLastError());
        TPF_TRY(tpf_sync_periodic_images(ndim, n_nodes, periodic, dst, dst, st));
        std::swap(src, dst);
      }
      hipLaunchKernelGGL(k_clip01, dim3(blocks_for(T)), dim3(kBlock), 0, st, T,
                         (const double*)src, q);
      TPF_HIP(hipGetLastError());
      TPF_TRY(tpf_sync_periodic_images(ndim, n_nodes, periodic, q, q, st));
    }
    if (mass_correction)
      TPF_TRY(tpf_dccd_mass_correct(ndim, n_nodes, L, periodic, q, m_old, out, nullptr, st));
    else
      TPF_HIP(hipMemcpyAsync(out, q, sizeof(double) * T, hipMemcpyDeviceToDevice, st));
    return TPF_OK;
  };
  const tpf_status s = run();
  (void)hipFreeAsync(buf, st);
  return s;
}
