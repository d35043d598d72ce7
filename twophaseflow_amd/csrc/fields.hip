// PPE inputs and post-projection outputs on the device.
//
// * gas spheres -> psi -> rho : simulation/initial_conditions/builder.py:100-172,
//   shape_primitives.py:21-53, levelset/heaviside.py:167-187
// * mass integral             : levelset/heaviside.py:118-149 with
//   core/grid.py:1123-1170 weights
// * velocity correction       : coupling/velocity_corrector.py:39-89
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "reduce.hpp"
#include "tpf_common.hpp"

namespace tpf {
namespace {

constexpr int kMaxSpheres = 512;

struct FieldGeom {
  int ndim;
  int n[3];
  double h[3];
  double L[3];
  int64_t total;
};

tpf_status make_geom(FieldGeom& g, int ndim, const int64_t* n_nodes, const double* L) {
  if (ndim < 1 || ndim > 3 || !n_nodes) return fail(TPF_ERR_VALUE, "bad ndim/shape");
  g.ndim = ndim;
  g.total = 1;
  for (int a = 0; a < 3; ++a) {
    g.n[a] = 1;
    g.h[a] = 1.0;
    g.L[a] = 1.0;
  }
  for (int a = 0; a < ndim; ++a) {
    if (n_nodes[a] < 2) return fail(TPF_ERR_VALUE, "each axis needs >= 2 nodes");
    g.n[a] = (int)n_nodes[a];
    if (L) {
      if (!(L[a] > 0.0)) return fail(TPF_ERR_VALUE, "L must be positive");
      g.L[a] = L[a];
      g.h[a] = L[a] / (double)(n_nodes[a] - 1);
    }
    g.total *= n_nodes[a];
  }
  if (g.total >= (int64_t)0xffffffffLL) return fail(TPF_ERR_VALUE, "grid too large");
  return TPF_OK;
}

__device__ inline void unravel(const FieldGeom& g, int64_t idx, int (&i)[3]) {
  uint32_t u = (uint32_t)idx;
  for (int a = g.ndim - 1; a >= 0; --a) {
    const uint32_t n = (uint32_t)g.n[a];
    const uint32_t q = u / n;
    i[a] = (int)(u - q * n);
    u = q;
  }
}

// node coordinate as numpy.linspace(0, L, N+1) produces it (i*step, last = L)
__device__ inline double coord(const FieldGeom& g, int a, int i) {
  return i == g.n[a] - 1 ? g.L[a] : (double)i * g.h[a];
}

__global__ __launch_bounds__(kBlock) void k_bubbles(FieldGeom g, int ns,
                                                    const double* __restrict__ sph, double eps,
                                                    double rho_l, double rho_g,
                                                    double* __restrict__ psi,
                                                    double* __restrict__ rho) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int i[3] = {0, 0, 0};
    unravel(g, idx, i);
    double x[3];
    for (int a = 0; a < g.ndim; ++a) x[a] = coord(g, a, i[a]);
    double phi_gas = INFINITY;
    for (int k = 0; k < ns; ++k) {
      const double* c = sph + k * 4;
      double r2 = 0.0;
      for (int a = 0; a < g.ndim; ++a) {
        const double d = x[a] - c[a];
        r2 = r2 + d * d;
      }
      phi_gas = fmin(phi_gas, sqrt(r2) - c[3]);
    }
    // liquid background: phi = max(-inf, -phi_gas); psi ~ 1 in liquid
    double t = -phi_gas / eps;
    t = fmin(fmax(t, -500.0), 500.0);
    const double ps = 1.0 / (1.0 + exp(t));
    if (psi) psi[idx] = ps;
    if (rho) rho[idx] = rho_g + (rho_l - rho_g) * ps;
  }
}

// control width of node i along a (Grid.cell_volumes: wall axes keep the
// full nodal h, periodic axes give the image node 0 weight)
__device__ inline double cv_width(const FieldGeom& g, const int* bc, int a, int i) {
  if (bc[a] == TPF_BC_PERIODIC && i == g.n[a] - 1) return 0.0;
  return g.h[a];
}

struct BcPack {
  int bc[3];
};

__global__ __launch_bounds__(kBlock) void k_mass(FieldGeom g, BcPack bp,
                                                 const double* __restrict__ q,
                                                 double* partials) {
  __shared__ double sh[kBlock / 64];
  double v[1] = {0.0};
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int i[3] = {0, 0, 0};
    unravel(g, idx, i);
    double w = 1.0;
    for (int a = 0; a < g.ndim; ++a) w *= cv_width(g, bp.bc, a, i[a]);
    v[0] += q[idx] * w;
  }
  block_partials<1>(v, partials, sh);
}

__global__ __launch_bounds__(kBlock) void k_fin_mass(const double* partials, int nblocks,
                                                     double* out) {
  __shared__ double sh[kBlock / 64];
  double v[1];
  finish_sum<1>(v, partials, nblocks, sh);
  if (threadIdx.x == 0) *out = v[0];
}

__global__ __launch_bounds__(kBlock) void k_correct(FieldGeom g, BcPack bp, int axis, double dt,
                                                    const double* __restrict__ rho,
                                                    const double* __restrict__ grad,
                                                    double* __restrict__ u) {
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < g.total;
       idx += (int64_t)gridDim.x * kBlock) {
    int i[3] = {0, 0, 0};
    unravel(g, idx, i);
    double gr = grad[idx];
    if (bp.bc[axis] == TPF_BC_WALL && (i[axis] == 0 || i[axis] == g.n[axis] - 1)) gr = 0.0;
    u[idx] = u[idx] - (dt / rho[idx]) * gr;
  }
}

int grid_of(int64_t total) {
  int64_t b = (total + kBlock - 1) / kBlock;
  return (int)(b < 4096 ? (b < 1 ? 1 : b) : 4096);
}

}  // namespace
}  // namespace tpf

using namespace tpf;

extern "C" tpf_status tpf_fields_bubbles(int ndim, const int64_t* n_nodes, const double* L,
                                         int n_spheres, const double* centres,
                                         const double* radii, double eps, double rho_l,
                                         double rho_g, double* psi, double* rho, void* stream) {
  FieldGeom g;
  TPF_TRY(make_geom(g, ndim, n_nodes, L));
  if (!L) return fail(TPF_ERR_VALUE, "L is required");
  if (n_spheres < 0 || n_spheres > kMaxSpheres || (n_spheres && (!centres || !radii)))
    return fail(TPF_ERR_VALUE, "bad sphere list (at most 512 spheres)");
  if (!(eps > 0.0)) return fail(TPF_ERR_VALUE, "eps must be positive");
  if (!psi && !rho) return TPF_OK;
  double host[kMaxSpheres * 4];
  for (int k = 0; k < n_spheres; ++k) {
    for (int a = 0; a < 3; ++a) host[k * 4 + a] = a < ndim ? centres[k * ndim + a] : 0.0;
    host[k * 4 + 3] = radii[k];
  }
  double* dsph = nullptr;
  hipStream_t st = (hipStream_t)stream;
  TPF_HIP(hipMalloc(&dsph, sizeof(double) * 4 * (n_spheres > 0 ? n_spheres : 1)));
  if (n_spheres)
    TPF_HIP(hipMemcpyAsync(dsph, host, sizeof(double) * 4 * n_spheres, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_bubbles, dim3(grid_of(g.total)), dim3(kBlock), 0, st, g, n_spheres,
                     (const double*)dsph, eps, rho_l, rho_g, psi, rho);
  hipError_t e = hipGetLastError();
  (void)hipStreamSynchronize(st);
  (void)hipFree(dsph);
  if (e != hipSuccess) return fail(TPF_ERR_RUNTIME, std::string("k_bubbles: ") + hipGetErrorString(e));
  return TPF_OK;
}

extern "C" tpf_status tpf_fields_mass(int ndim, const int64_t* n_nodes, const double* L,
                                      const int* bc, const double* q, double* mass,
                                      void* stream) {
  FieldGeom g;
  TPF_TRY(make_geom(g, ndim, n_nodes, L));
  if (!L || !bc || !q || !mass) return fail(TPF_ERR_VALUE, "null argument");
  BcPack bp{{0, 0, 0}};
  for (int a = 0; a < ndim; ++a) bp.bc[a] = bc[a];
  // For periodic axes Grid.cell_volumes uses the wrapped half cells,
  // 0.5*(h+h) = h on a uniform grid: the same h as wall axes.
  const int grid = grid_of(g.total);
  hipStream_t st = (hipStream_t)stream;
  double* buf = nullptr;
  TPF_HIP(hipMalloc(&buf, sizeof(double) * (grid + 1)));
  hipLaunchKernelGGL(k_mass, dim3(grid), dim3(kBlock), 0, st, g, bp, q, buf);
  hipLaunchKernelGGL(k_fin_mass, dim3(1), dim3(kBlock), 0, st, (const double*)buf, grid, buf + grid);
  hipError_t e = hipGetLastError();
  double out = 0.0;
  if (e == hipSuccess) e = hipMemcpyAsync(&out, buf + grid, sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(buf);
  if (e != hipSuccess) return fail(TPF_ERR_RUNTIME, std::string("k_mass: ") + hipGetErrorString(e));
  *mass = out;
  return TPF_OK;
}

extern "C" tpf_status tpf_fields_correct_velocity(int ndim, const int64_t* n_nodes, const int* bc,
                                                  int axis, double dt, const double* rho,
                                                  const double* grad, double* u, void* stream) {
  FieldGeom g;
  TPF_TRY(make_geom(g, ndim, n_nodes, nullptr));
  if (!bc || !rho || !grad || !u) return fail(TPF_ERR_VALUE, "null argument");
  if (axis < 0 || axis >= ndim) return fail(TPF_ERR_VALUE, "axis out of range");
  BcPack bp{{0, 0, 0}};
  for (int a = 0; a < ndim; ++a) bp.bc[a] = bc[a];
  hipLaunchKernelGGL(k_correct, dim3(grid_of(g.total)), dim3(kBlock), 0, (hipStream_t)stream, g,
                     bp, axis, dt, rho, grad, u);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}
