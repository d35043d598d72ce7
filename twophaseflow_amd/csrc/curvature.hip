// Interface curvature straight from the CLS field psi
// (CurvatureCalculatorPsi, levelset/curvature_psi.py:45-196):
//   2-D  kappa = -[psi_y^2 psi_xx - 2 psi_x psi_y psi_xy + psi_x^2 psi_yy] / (|grad psi|^2 + 1e-30)^1.5
//   3-D  kappa = -(lap - grad.H.grad / (|grad|^2 + 1e-30)) / sqrt(|grad|^2 + 1e-30)
// zero where psi <= psi_min or psi >= 1 - psi_min; the derivatives come from
// the CCD (caller) and the optional DCCD filter below (levelset/curvature.py:
// 60-128): successive 3-point passes f + eps (2 psi - 1)^2 (f+ - 2 f + f-)
// along every axis (periodic axes over the unique nodes, image = node 0; wall
// axes interior only), then the original values restored on every wall
// boundary plane.  Plus the optional 3x3 Gaussian (2-D, edge-replicated,
// curvature_psi.py:171-196).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "reduce.hpp"
#include "tpf_common.hpp"

namespace tpf {
namespace {

struct KGeom {
  int n[3];
  int periodic[3];
  int64_t total;
};

__device__ __forceinline__ int64_t stride_of(const KGeom& g, int a) {
  return a == 0 ? (int64_t)g.n[1] * g.n[2] : (a == 1 ? (int64_t)g.n[2] : 1);
}

__device__ __forceinline__ void coords(const KGeom& g, int64_t idx, int (&c)[3]) {
  const uint32_t plane = (uint32_t)g.n[1] * (uint32_t)g.n[2];
  c[0] = (int)((uint32_t)idx / plane);
  const uint32_t rem = (uint32_t)idx - (uint32_t)c[0] * plane;
  c[1] = (int)(rem / (uint32_t)g.n[2]);
  c[2] = (int)(rem - (uint32_t)c[1] * (uint32_t)g.n[2]);
}

// One 1-D filter pass along slot a (curvature.py:60-103); the last pass also
// restores the wall boundary planes from the unfiltered field (:119-127).
__global__ __launch_bounds__(kBlock) void k_filter_pass(KGeom g, int a, double eps,
                                                        const double* __restrict__ psi,
                                                        const double* __restrict__ in,
                                                        const double* __restrict__ orig,
                                                        int restore, double* __restrict__ out) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords(g, idx, c);
    const int N = g.n[a] - 1;
    const int64_t st = stride_of(g, a);
    const int i = c[a];
    double v;
    if (g.periodic[a]) {
      // filtered unique nodes; the image N repeats node 0
      const int k = i == N ? 0 : i;
      const int64_t o = idx - (int64_t)(i - k) * st;
      const int64_t om = k == 0 ? o + (int64_t)(N - 1) * st : o - st;
      const int64_t op = k == N - 1 ? o - (int64_t)(N - 1) * st : o + st;
      const double s = 2.0 * psi[o] - 1.0;
      v = in[o] + (eps * (s * s)) * (in[op] - 2.0 * in[o] + in[om]);
    } else if (i == 0 || i == N) {
      v = in[idx];
    } else {
      const double s = 2.0 * psi[idx] - 1.0;
      v = in[idx] + (eps * (s * s)) * (in[idx + st] - 2.0 * in[idx] + in[idx - st]);
    }
    if (restore) {
#pragma unroll
      for (int b = 0; b < 3; ++b)
        if (g.n[b] > 1 && !g.periodic[b] && (c[b] == 0 || c[b] == g.n[b] - 1)) v = orig[idx];
    }
    out[idx] = v;
  }
}

// The same pass over up to kMaxFields fields sharing psi: the switch
// eps (2 psi - 1)^2 is formed once per node and each field takes exactly
// k_filter_pass's arithmetic.
constexpr int kMaxFields = 8;
struct FieldSet {
  const double* in[kMaxFields];
  const double* orig[kMaxFields];
  double* out[kMaxFields];
};

__global__ __launch_bounds__(kBlock) void k_filter_pass_multi(KGeom g, int a, double eps,
                                                              const double* __restrict__ psi,
                                                              FieldSet F, int nf, int restore) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords(g, idx, c);
    const int N = g.n[a] - 1;
    const int64_t st = stride_of(g, a);
    const int i = c[a];
    bool keep = false;  // wall boundary plane restored from the original
    if (restore) {
#pragma unroll
      for (int b = 0; b < 3; ++b)
        if (g.n[b] > 1 && !g.periodic[b] && (c[b] == 0 || c[b] == g.n[b] - 1)) keep = true;
    }
    int64_t o = idx, om = 0, op = 0;
    bool filt = true;
    if (g.periodic[a]) {
      const int k = i == N ? 0 : i;
      o = idx - (int64_t)(i - k) * st;
      om = k == 0 ? o + (int64_t)(N - 1) * st : o - st;
      op = k == N - 1 ? o - (int64_t)(N - 1) * st : o + st;
    } else if (i == 0 || i == N) {
      filt = false;
    } else {
      om = idx - st;
      op = idx + st;
    }
    double w = 0.0;
    if (filt) {
      const double s = 2.0 * psi[o] - 1.0;
      w = eps * (s * s);
    }
    for (int f = 0; f < nf; ++f) {
      double v;
      if (keep) {
        v = F.orig[f][idx];
      } else if (filt) {
        const double* in = F.in[f];
        v = in[o] + w * (in[op] - 2.0 * in[o] + in[om]);
      } else {
        v = F.in[f][idx];
      }
      F.out[f][idx] = v;
    }
  }
}

// d1, d2 stacked [ndim][total]; mixed: 2-D [psi_xy], 3-D [xy, xz, yz].
__global__ __launch_bounds__(kBlock) void k_curvature(KGeom g, int ndim,
                                                      const double* __restrict__ psi,
                                                      const double* __restrict__ d1,
                                                      const double* __restrict__ d2,
                                                      const double* __restrict__ mixed,
                                                      double psi_min, double* __restrict__ out) {
  const int64_t T = g.total;
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < T;
       idx += (int64_t)gridDim.x * kBlock) {
    double grad_sq = 0.0;
    for (int a = 0; a < ndim; ++a) grad_sq += d1[a * T + idx] * d1[a * T + idx];
    double kappa;
    if (ndim == 2) {
      const double px = d1[idx], py = d1[T + idx];
      const double num = py * py * d2[idx] - 2.0 * px * py * mixed[idx] + px * px * d2[T + idx];
      kappa = -num / pow(grad_sq + 1e-30, 1.5);
    } else {
      const double lap = ((0.0 + d2[idx]) + d2[T + idx]) + d2[2 * T + idx];
      double con = 0.0;
      for (int a = 0; a < 3; ++a) con += d1[a * T + idx] * d1[a * T + idx] * d2[a * T + idx];
      con = con + 2.0 * d1[idx] * d1[T + idx] * mixed[idx];            // (0, 1)
      con = con + 2.0 * d1[idx] * d1[2 * T + idx] * mixed[T + idx];     // (0, 2)
      con = con + 2.0 * d1[T + idx] * d1[2 * T + idx] * mixed[2 * T + idx];  // (1, 2)
      const double gs = grad_sq + 1e-30;
      kappa = -(lap - con / gs) / sqrt(gs);
    }
    const double p = psi[idx];
    out[idx] = (p <= psi_min || p >= 1.0 - psi_min) ? 0.0 : kappa;
  }
}

// 3x3 Gaussian [1 2 1; 2 4 2; 1 2 1] / 16 with edge replication (2-D).
__global__ __launch_bounds__(kBlock) void k_gauss3(int n0, int n1, const double* __restrict__ in,
                                                   double* __restrict__ out) {
  const int64_t T = (int64_t)n0 * n1;
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < T;
       idx += (int64_t)gridDim.x * kBlock) {
    const int i = (int)(idx / n1), j = (int)(idx - (int64_t)i * n1);
    auto at = [&](int di, int dj) {
      const int ii = min(max(i + di, 0), n0 - 1), jj = min(max(j + dj, 0), n1 - 1);
      return in[(int64_t)ii * n1 + jj];
    };
    const double r = 1.0 * at(-1, -1) + 2.0 * at(-1, 0) + 1.0 * at(-1, 1) + 2.0 * at(0, -1) +
                     4.0 * at(0, 0) + 2.0 * at(0, 1) + 1.0 * at(1, -1) + 2.0 * at(1, 0) +
                     1.0 * at(1, 1);
    out[idx] = r / 16.0;
  }
}

tpf_status make_geom(KGeom& g, int ndim, const int64_t* n_nodes, const int* periodic) {
  if (!n_nodes || !periodic) return fail(TPF_ERR_VALUE, "null argument");
  if (ndim < 1 || ndim > 3) return fail(TPF_ERR_VALUE, "ndim must be 1, 2 or 3");
  g = KGeom{};
  for (int a = 0; a < 3; ++a) g.n[a] = 1;
  for (int a = 0; a < ndim; ++a) {
    const int s = 3 - ndim + a;
    if (n_nodes[a] < 3) return fail(TPF_ERR_VALUE, "axis needs >= 3 nodes");
    g.n[s] = (int)n_nodes[a];
    g.periodic[s] = periodic[a] != 0;
  }
  g.total = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (g.total >= (int64_t)0xffffffffLL) return fail(TPF_ERR_VALUE, "grid too large (2^32 nodes)");
  return TPF_OK;
}

unsigned blocks_for(int64_t total) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((total + kBlock - 1) / kBlock, 2048));
}

}  // namespace
}  // namespace tpf

using namespace tpf;

extern "C" tpf_status tpf_dccd_filter_nd(int ndim, const int64_t* n_nodes, const int* periodic,
                                         const double* psi, double eps, const double* f,
                                         double* out, double* scratch, void* stream) {
  KGeom g;
  TPF_TRY(make_geom(g, ndim, n_nodes, periodic));
  if (!psi || !f || !out || !scratch) return fail(TPF_ERR_VALUE, "null argument");
  if (out == f || scratch == f || scratch == out)
    return fail(TPF_ERR_VALUE, "f, out and scratch must be distinct arrays");
  hipStream_t st = (hipStream_t)stream;
  // passes along the axes in order; pass t of ndim writes out when ndim-1-t
  // is even, else scratch, so the last pass lands in out and no pass reads
  // the array it writes
  const double* src = f;
  for (int t = 0; t < ndim; ++t) {
    const int a = 3 - ndim + t;
    double* dst = ((ndim - 1 - t) % 2 == 0) ? out : scratch;
    hipLaunchKernelGGL(k_filter_pass, dim3(blocks_for(g.total)), dim3(kBlock), 0, st, g, a, eps,
                       psi, src, f, t == ndim - 1 ? 1 : 0, dst);
    TPF_HIP(hipGetLastError());
    src = dst;
  }
  return TPF_OK;
}

extern "C" tpf_status tpf_dccd_filter_nd_multi(int ndim, const int64_t* n_nodes,
                                               const int* periodic, const double* psi, double eps,
                                               int nf, const double* const* f,
                                               double* const* out, double* const* scratch,
                                               void* stream) {
  KGeom g;
  TPF_TRY(make_geom(g, ndim, n_nodes, periodic));
  if (!psi || !f || !out || !scratch) return fail(TPF_ERR_VALUE, "null argument");
  if (nf < 1 || nf > kMaxFields) return fail(TPF_ERR_VALUE, "1 <= nf <= 8 fields per call");
  const double* all[3 * kMaxFields];
  for (int i = 0; i < nf; ++i) {
    if (!f[i] || !out[i] || !scratch[i]) return fail(TPF_ERR_VALUE, "null field");
    all[3 * i] = f[i];
    all[3 * i + 1] = out[i];
    all[3 * i + 2] = scratch[i];
  }
  for (int i = 0; i < 3 * nf; ++i)
    for (int j = i + 1; j < 3 * nf; ++j)
      if (all[i] == all[j])
        return fail(TPF_ERR_VALUE, "fields, outputs and scratch arrays must all be distinct");
  hipStream_t st = (hipStream_t)stream;
  // the passes of tpf_dccd_filter_nd, every field per launch
  FieldSet F{};
  for (int i = 0; i < nf; ++i) {
    F.in[i] = f[i];
    F.orig[i] = f[i];
  }
  for (int t = 0; t < ndim; ++t) {
    const int a = 3 - ndim + t;
    const bool to_out = (ndim - 1 - t) % 2 == 0;
    for (int i = 0; i < nf; ++i) F.out[i] = to_out ? out[i] : scratch[i];
    hipLaunchKernelGGL(k_filter_pass_multi, dim3(blocks_for(g.total)), dim3(kBlock), 0, st, g, a,
                       eps, psi, F, nf, t == ndim - 1 ? 1 : 0);
    TPF_HIP(hipGetLastError());
    for (int i = 0; i < nf; ++i) F.in[i] = F.out[i];
  }
  return TPF_OK;
}

extern "C" tpf_status tpf_curvature_psi(int ndim, const int64_t* n_nodes, const double* psi,
                                        const double* d1, const double* d2, const double* mixed,
                                        double psi_min, double* out, void* stream) {
  if (ndim != 2 && ndim != 3) return fail(TPF_ERR_VALUE, "curvature from psi is 2-D or 3-D");
  KGeom g;
  const int per[3] = {0, 0, 0};
  TPF_TRY(make_geom(g, ndim, n_nodes, per));
  if (!psi || !d1 || !d2 || !mixed || !out) return fail(TPF_ERR_VALUE, "null argument");
  hipLaunchKernelGGL(k_curvature, dim3(blocks_for(g.total)), dim3(kBlock), 0, (hipStream_t)stream,
                     g, ndim, psi, d1, d2, mixed, psi_min, out);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

extern "C" tpf_status tpf_gaussian_3x3(const int64_t* n_nodes, const double* in, double* out,
                                       void* stream) {
  if (!n_nodes || !in || !out || in == out) return fail(TPF_ERR_VALUE, "bad argument");
  if (n_nodes[0] < 1 || n_nodes[1] < 1) return fail(TPF_ERR_VALUE, "bad n_nodes");
  const int64_t T = n_nodes[0] * n_nodes[1];
  hipLaunchKernelGGL(k_gauss3, dim3(blocks_for(T)), dim3(kBlock), 0, (hipStream_t)stream,
                     (int)n_nodes[0], (int)n_nodes[1], in, out);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}
