// Dissipative-CCD level-set advection (DissipativeCCDAdvection,
// levelset/advection_dccd.py:20-152): TVD-RK3 (time_integration/tvd_rk3.py:
// 26-48) of d psi/dt = -sum_a F~_a with F~_a the adaptive 3-point filter
// fp + eps_d (2 psi - 1)^2 (fp_{+1} - 2 fp + fp_{-1}) of fp = the CCD first
// derivative of psi u_a along a (levelset/advection_kernels.py:31-34; ghost
// values from core/boundary.py:249-286; non-periodic end nodes unfiltered,
// advection_dccd.py:145-152), each stage clipped and its periodic images
// synchronised (core/boundary.py:116-133), with the optional
// interface-weighted mass correction (levelset/heaviside.py:118-149).
//
// Per stage: one flux kernel (psi u_a for every axis), the CCD launches
// (caller), and one kernel that filters, sums the axes, forms the TVD-RK3
// combination, clips and writes the images from their roots (an image node
// takes the value of the node with every periodic image index set to 0, which
// is what the reference's axis-by-axis copies leave there).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "reduce.hpp"
#include "tpf_common.hpp"

// Fixed device scratch of the allocation-free mass correction
// (tpf_dccd_mass_correct_noalloc): 2 partial sums per workgroup of the
// reduction (at most 2048 workgroups), the two finished sums and the factor.
// One copy per device, alive until the process ends; every call on a device
// shares it.  In host code `scratch` names a host-side handle, not device
// memory: kernels get the address from hipGetSymbolAddress.
static constexpr int kMassScratchDoubles = 2 * 2048 + 8;
static __device__ __align__(256) double scratch[kMassScratchDoubles];

namespace tpf {
namespace {

struct DGeom {
  int n[3];
  int periodic[3];  // level-set boundary of the slot is periodic
  int k[3];         // axis index of the slot
  int64_t total;
  double h[3];
};

__device__ __forceinline__ int64_t stride_of(const DGeom& g, int a) {
  return a == 0 ? (int64_t)g.n[1] * g.n[2] : (a == 1 ? (int64_t)g.n[2] : 1);
}

__device__ __forceinline__ void coords(const DGeom& g, int64_t idx, int (&c)[3]) {
  const uint32_t plane = (uint32_t)g.n[1] * (uint32_t)g.n[2];
  c[0] = (int)((uint32_t)idx / plane);
  const uint32_t rem = (uint32_t)idx - (uint32_t)c[0] * plane;
  c[1] = (int)(rem / (uint32_t)g.n[2]);
  c[2] = (int)(rem - (uint32_t)c[1] * (uint32_t)g.n[2]);
}

// The root of a node: every periodic image index (N) replaced by 0.
__device__ __forceinline__ int64_t root_of(const DGeom& g, int64_t idx, int (&c)[3]) {
  int64_t r = idx;
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (g.periodic[a] && g.n[a] > 1 && c[a] == g.n[a] - 1) {
      r -= (int64_t)c[a] * stride_of(g, a);
      c[a] = 0;
    }
  return r;
}

__global__ __launch_bounds__(kBlock) void k_dccd_flux(DGeom g, const double* __restrict__ q,
                                                      const double* __restrict__ u0,
                                                      const double* __restrict__ u1,
                                                      const double* __restrict__ u2,
                                                      double* __restrict__ f, int ndim) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    const double qi = q[idx];
    f[idx] = qi * u0[idx];
    if (ndim > 1) f[g.total + idx] = qi * u1[idx];
    if (ndim > 2) f[2 * g.total + idx] = qi * u2[idx];
  }
}

// out = clip(a q0 + b (q + dt L(q))) at the root of every node.
__global__ __launch_bounds__(kBlock) void k_dccd_stage(DGeom g, const double* __restrict__ q0,
                                                       const double* __restrict__ q,
                                                       const double* __restrict__ fp,
                                                       double eps_d, double dt, double ca,
                                                       double cb, int clip, double lo, double hi,
                                                       double* __restrict__ out) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords(g, idx, c);
    const int64_t r = root_of(g, idx, c);
    const double psi = q[r];
    const double sw = (2.0 * psi - 1.0) * (2.0 * psi - 1.0);
    double L = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (g.n[a] == 1) continue;
      const int N = g.n[a] - 1;
      const int i = c[a];
      const int64_t st = stride_of(g, a);
      const double* fa = fp + (int64_t)g.k[a] * g.total;
      const double f0 = fa[r];
      double Ft;
      if (!g.periodic[a] && (i == 0 || i == N)) {
        Ft = f0;  // unfiltered wall / outflow end node
      } else {
        // periodic ghosts: node -1 -> N-1, node N+1 -> 1 (core/boundary.py:273-276)
        const double fm = i == 0 ? fa[r + (int64_t)(N - 1) * st] : fa[r - st];
        const double fpl = i == N ? fa[r - (int64_t)(N - 1) * st] : fa[r + st];
        Ft = f0 + eps_d * sw * (fpl - 2.0 * f0 + fm);
      }
      L -= Ft;
    }
    double v = ca * q0[r] + cb * (psi + dt * L);
    if (clip) v = fmin(fmax(v, lo), hi);
    out[idx] = v;
  }
}

// images take their roots' values (in place is safe: a root is never an image)
__global__ __launch_bounds__(kBlock) void k_dccd_sync(DGeom g, const double* __restrict__ in,
                                                      double* __restrict__ out) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords(g, idx, c);
    const int64_t r = root_of(g, idx, c);
    if (r != idx || in != out) out[idx] = in[r];
  }
}

// Grid.cell_volumes weight of a node (core/grid.py:1123-1170, uniform):
// h per axis, 0 on the periodic image node.
__device__ __forceinline__ double dvol(const DGeom& g, const int (&c)[3]) {
  double w = 1.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (g.n[a] == 1) continue;
    w *= (g.periodic[a] && c[a] == g.n[a] - 1) ? 0.0 : g.h[a];
  }
  return w;
}

// partials of sum q dV and sum 4 q (1 - q) dV
__global__ __launch_bounds__(kBlock) void k_dccd_mass_sums(DGeom g, const double* __restrict__ q,
                                                           double* partials) {
  __shared__ double sh[2 * kBlock / 64];
  double v[2] = {0.0, 0.0};
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords(g, idx, c);
    const double dv = dvol(g, c);
    const double qi = q[idx];
    v[0] += qi * dv;
    v[1] += (4.0 * qi * (1.0 - qi)) * dv;
  }
  block_partials<2>(v, partials, sh);
}

__global__ __launch_bounds__(kBlock) void k_dccd_mass_fin(const double* partials, int nblocks,
                                                          double* out) {
  __shared__ double sh[2 * kBlock / 64];
  double v[2];
  finish_sum<2>(v, partials, nblocks, sh);
  if (threadIdx.x == 0) {
    out[0] = v[0];
    out[1] = v[1];
  }
}

// The two sums of k_dccd_mass_sums finished as k_dccd_mass_fin does, then the
// correction factor formed on the device (heaviside.py:138-149):
// ratio = (m_target - M_new) / W gated at W > 1e-12.  out[0..1] = (M_new, W),
// out[2] = factor.
__global__ __launch_bounds__(kBlock) void k_dccd_mass_fin_factor(const double* partials,
                                                                 int nblocks, double m_target,
                                                                 double* out) {
  __shared__ double sh[2 * kBlock / 64];
  double v[2];
  finish_sum<2>(v, partials, nblocks, sh);
  if (threadIdx.x == 0) {
    const double W = v[1];
    const double ratio = (m_target - v[0]) / (W > 1e-12 ? W : 1.0);
    out[0] = v[0];
    out[1] = W;
    out[2] = (W > 1e-12 ? 1.0 : 0.0) * ratio;
  }
}

// out = clip(q + factor 4 q (1 - q), 0, 1) at the root of every node, the
// factor read from device memory.
__global__ __launch_bounds__(kBlock) void k_dccd_mass_apply_dev(DGeom g,
                                                                const double* __restrict__ q,
                                                                const double* __restrict__ factor,
                                                                double* __restrict__ out) {
  const double fac = *factor;
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords(g, idx, c);
    const double qi = q[root_of(g, idx, c)];
    const double w = 4.0 * qi * (1.0 - qi);
    out[idx] = fmin(fmax(qi + fac * w, 0.0), 1.0);
  }
}

// out = clip(q + factor 4 q (1 - q), 0, 1) at the root of every node.
__global__ __launch_bounds__(kBlock) void k_dccd_mass_apply(DGeom g, const double* __restrict__ q,
                                                            double factor,
                                                            double* __restrict__ out) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int c[3];
    coords(g, idx, c);
    const double qi = q[root_of(g, idx, c)];
    const double w = 4.0 * qi * (1.0 - qi);
    out[idx] = fmin(fmax(qi + factor * w, 0.0), 1.0);
  }
}

tpf_status make_geom(DGeom& g, int ndim, const int64_t* n_nodes, const double* L,
                     const int* periodic) {
  if (!n_nodes || !periodic) return fail(TPF_ERR_VALUE, "null argument");
  if (ndim < 1 || ndim > 3) return fail(TPF_ERR_VALUE, "ndim must be 1, 2 or 3");
  g = DGeom{};
  for (int a = 0; a < 3; ++a) {
    g.n[a] = 1;
    g.h[a] = 1.0;
  }
  for (int a = 0; a < ndim; ++a) {
    const int s = 3 - ndim + a;
    if (n_nodes[a] < 3) return fail(TPF_ERR_VALUE, "DCCD axis needs >= 3 nodes");
    g.n[s] = (int)n_nodes[a];
    g.periodic[s] = periodic[a] != 0;
    g.k[s] = a;
    g.h[s] = L ? L[a] / (double)(n_nodes[a] - 1) : 1.0;
  }
  g.total = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (g.total >= (int64_t)0xffffffffLL) return fail(TPF_ERR_VALUE, "grid too large (2^32 nodes)");
  return TPF_OK;
}

// In-place image sync: one thread per node of the image planes (c[a] = N on
// a periodic slot a), the planes of the periodic slots enumerated one after
// another.  A node on several image planes is written once per plane, with
// the same root value each time; roots lie on no image plane, so no write
// changes a value another thread reads.
__global__ __launch_bounds__(kBlock) void k_dccd_sync_planes(DGeom g, int64_t nplane,
                                                             double* __restrict__ q) {
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < nplane;
       t += (int64_t)gridDim.x * kBlock) {
    int64_t u = t;
    int a = 0;
    for (; a < 3; ++a) {
      if (!g.periodic[a] || g.n[a] == 1) continue;
      const int64_t pa = g.total / g.n[a];
      if (u < pa) break;
      u -= pa;
    }
    if (a == 3) continue;  // not reached for t < nplane
    // the other two slots of the plane, in order
    const int b0 = a == 0 ? 1 : 0, b1 = a == 2 ? 1 : 2;
    int c[3];
    c[a] = g.n[a] - 1;
    c[b0] = (int)(u / g.n[b1]);
    c[b1] = (int)(u - (int64_t)c[b0] * g.n[b1]);
    const int64_t idx = (int64_t)c[0] * g.n[1] * g.n[2] + (int64_t)c[1] * g.n[2] + c[2];
    const int64_t r = root_of(g, idx, c);
    q[idx] = q[r];
  }
}

constexpr int kMaxBlocks = 2048;

unsigned blocks_for(int64_t total) {
  return (unsigned)std::max<int64_t>(1,
                                     std::min<int64_t>((total + kBlock - 1) / kBlock, kMaxBlocks));
}

static_assert(2 * kMaxBlocks + 3 <= kMassScratchDoubles, "mass scratch too small");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current.  The symbol's name is only a host-side
// handle: kernels receive this address.
tpf_status mass_scratch(double** out) {
  int dev = 0;
  TPF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(TPF_ERR_RUNTIME, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(TPF_ERR_RUNTIME, std::string("hipGetSymbolAddress(scratch): ") +
                                       hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return TPF_OK;
}

}  // namespace
}  // namespace tpf

using namespace tpf;

extern "C" tpf_status tpf_dccd_flux(int ndim, const int64_t* n_nodes, const double* q,
                                    const double* const* vel, double* f, void* stream) {
  DGeom g;
  const int per[3] = {0, 0, 0};
  TPF_TRY(make_geom(g, ndim, n_nodes, nullptr, per));
  if (!q || !vel || !f) return fail(TPF_ERR_VALUE, "null argument");
  for (int a = 0; a < ndim; ++a)
    if (!vel[a]) return fail(TPF_ERR_VALUE, "null velocity component");
  hipLaunchKernelGGL(k_dccd_flux, dim3(blocks_for(g.total)), dim3(kBlock), 0, (hipStream_t)stream,
                     g, q, vel[0], ndim > 1 ? vel[1] : nullptr, ndim > 2 ? vel[2] : nullptr, f,
                     ndim);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

extern "C" tpf_status tpf_dccd_stage(int ndim, const int64_t* n_nodes, const int* periodic,
                                     const double* q0, const double* q, const double* fp,
                                     double eps_d, double dt, double ca, double cb, int clip,
                                     double lo, double hi, double* out, void* stream) {
  DGeom g;
  TPF_TRY(make_geom(g, ndim, n_nodes, nullptr, periodic));
  if (!q0 || !q || !fp || !out) return fail(TPF_ERR_VALUE, "null argument");
  if (out == q || out == q0) return fail(TPF_ERR_VALUE, "out must not alias q or q0");
  hipLaunchKernelGGL(k_dccd_stage, dim3(blocks_for(g.total)), dim3(kBlock), 0,
                     (hipStream_t)stream, g, q0, q, fp, eps_d, dt, ca, cb, clip, lo, hi, out);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

extern "C" tpf_status tpf_sync_periodic_images(int ndim, const int64_t* n_nodes,
                                               const int* periodic, const double* in, double* out,
                                               void* stream) {
  DGeom g;
  TPF_TRY(make_geom(g, ndim, n_nodes, nullptr, periodic));
  if (!in || !out) return fail(TPF_ERR_VALUE, "null argument");
  if (in == out) {  // only the image planes change; nothing without a periodic axis
    int64_t nplane = 0;
    for (int a = 0; a < 3; ++a)
      if (g.periodic[a] && g.n[a] > 1) nplane += g.total / g.n[a];
    if (nplane == 0) return TPF_OK;
    hipLaunchKernelGGL(k_dccd_sync_planes, dim3(blocks_for(nplane)), dim3(kBlock), 0,
                       (hipStream_t)stream, g, nplane, out);
    TPF_HIP(hipGetLastError());
    return TPF_OK;
  }
  hipLaunchKernelGGL(k_dccd_sync, dim3(blocks_for(g.total)), dim3(kBlock), 0, (hipStream_t)stream,
                     g, in, out);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

extern "C" tpf_status tpf_dccd_mass_correct(int ndim, const int64_t* n_nodes, const double* L,
                                            const int* periodic, const double* q,
                                            double m_target, double* out, double* mass_before,
                                            void* stream) {
  DGeom g;
  if (!L) return fail(TPF_ERR_VALUE, "null argument");
  TPF_TRY(make_geom(g, ndim, n_nodes, L, periodic));
  if (!q || !out || out == q) return fail(TPF_ERR_VALUE, "q and out must be distinct arrays");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = blocks_for(g.total);
  double* buf = nullptr;
  double host[2] = {0.0, 0.0};
  TPF_HIP(hipMallocAsync(&buf, sizeof(double) * (2 * (size_t)nb + 2), st));
  hipLaunchKernelGGL(k_dccd_mass_sums, dim3(nb), dim3(kBlock), 0, st, g, q, buf);
  hipLaunchKernelGGL(k_dccd_mass_fin, dim3(1), dim3(kBlock), 0, st, (const double*)buf, (int)nb,
                     buf + 2 * nb);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(host, buf + 2 * nb, sizeof(host), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFreeAsync(buf, st);
  if (e != hipSuccess) return fail(TPF_ERR_RUNTIME, std::string("mass sums: ") + hipGetErrorString(e));
  // heaviside.py:138-149: ratio = (M - M_new) / W, gated at W > 1e-12
  const double M_new = host[0], W = host[1];
  const double ratio = (m_target - M_new) / (W > 1e-12 ? W : 1.0);
  const double factor = (W > 1e-12 ? 1.0 : 0.0) * ratio;
  if (mass_before) *mass_before = M_new;
  hipLaunchKernelGGL(k_dccd_mass_apply, dim3(nb), dim3(kBlock), 0, st, g, q, factor, out);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}

extern "C" tpf_status tpf_dccd_mass_correct_noalloc(int ndim, const int64_t* n_nodes,
                                                    const double* L, const int* periodic,
                                                    const double* q, double m_target, double* out,
                                                    double* mass_before, void* stream) {
  DGeom g;
  if (!L) return fail(TPF_ERR_VALUE, "null argument");
  TPF_TRY(make_geom(g, ndim, n_nodes, L, periodic));
  if (!q || !out || out == q) return fail(TPF_ERR_VALUE, "q and out must be distinct arrays");
  const unsigned nb = blocks_for(g.total);
  if (2 * (int64_t)nb + 3 > kMassScratchDoubles)
    return fail(TPF_ERR_VALUE, "mass correction needs more scratch than the fixed maximum");
  double* sc = nullptr;
  TPF_TRY(mass_scratch(&sc));
  hipStream_t st = (hipStream_t)stream;
  double* sums = sc + 2 * (size_t)nb;  // (M_new, W, factor)
  hipLaunchKernelGGL(k_dccd_mass_sums, dim3(nb), dim3(kBlock), 0, st, g, q, sc);
  hipLaunchKernelGGL(k_dccd_mass_fin_factor, dim3(1), dim3(kBlock), 0, st, (const double*)sc,
                     (int)nb, m_target, sums);
  hipLaunchKernelGGL(k_dccd_mass_apply_dev, dim3(nb), dim3(kBlock), 0, st, g, q,
                     (const double*)(sums + 2), out);
  TPF_HIP(hipGetLastError());
  if (mass_before) {
    TPF_HIP(hipMemcpyAsync(mass_before, sums, sizeof(double), hipMemcpyDeviceToHost, st));
    TPF_HIP(hipStreamSynchronize(st));
  }
  return TPF_OK;
}
