// FCCD face primitives (FCCDSolver, ccd/fccd.py:165-624) on one axis of a
// node-centred field: 4th-order face gradient, face value and face second
// derivative from u and q = the nodal CCD second derivative (ccd/fccd.py:
// 95-112, 243-317), the R4 Hermite node gradient (:128-131, 392-462), the
// upwind Taylor face state (:115-125, 333-390) and the face-flux divergence
// (:464-521).  Uniform weights as ccd/fccd_helpers.py:17-35.  Face arrays
// hold N faces along the axis (face j between nodes j and j+1; on periodic
// axes face N-1 pairs node N-1 with node 0); nodal arrays hold N+1 nodes.
// Memory-bound streaming stencils: one thread per output element, the axis
// neighbours come through L1/L2.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "tpf_common.hpp"

namespace tpf {
namespace {

constexpr int kBlock = 256;

struct FaceGeom {
  int64_t outer, inner;  // products of the dimensions before / after the axis
  int n;                 // nodes along the axis (N + 1)
  int periodic;
  double inv_H, H24, H2_16, H16, H_half, H2_8;
};

__device__ __forceinline__ double face_grad(const FaceGeom& g, const double* __restrict__ u,
                                            const double* __restrict__ q, int64_t lo,
                                            int64_t hi) {
  return (u[hi] - u[lo]) * g.inv_H - g.H24 * (q[hi] - q[lo]);
}

// ops 0..2, 5: face outputs; 3: node gradient; 4: face divergence.
__global__ __launch_bounds__(kBlock) void k_fccd_face_op(int op, FaceGeom g,
                                                         const double* __restrict__ u,
                                                         const double* __restrict__ q,
                                                         const double* __restrict__ d,
                                                         const double* __restrict__ vel,
                                                         double* __restrict__ out,
                                                         int64_t total) {
  const int N = g.n - 1;
  const bool face_out = op <= 2 || op == 5;
  const int len = face_out ? N : g.n;
  for (int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x; o < total;
       o += (int64_t)gridDim.x * kBlock) {
    const int64_t b = o % g.inner;
    const int64_t t = o / g.inner;
    const int i = (int)(t % len);
    const int64_t a = t / len;
    const int64_t base = a * (int64_t)g.n * g.inner + b;  // node 0 of this line
    auto node = [&](int k) { return base + (int64_t)k * g.inner; };
    auto wrap = [&](int k) { return k < 0 ? k + N : (k >= N ? k - N : k); };
    double r;
    if (face_out) {
      const int64_t lo = node(i), hi = node(g.periodic && i + 1 == N ? 0 : i + 1);
      if (op == 0) {
        r = face_grad(g, u, q, lo, hi);
      } else if (op == 1) {
        r = 0.5 * (u[lo] + u[hi]) - g.H2_16 * (q[lo] + q[hi]);
      } else if (op == 2) {
        r = 0.5 * (q[lo] + q[hi]);
      } else {
        const int64_t f = a * (int64_t)N * g.inner + (int64_t)i * g.inner + b;
        const double left = u[lo] + g.H_half * d[lo] + g.H2_8 * q[lo];
        const double right = u[hi] - g.H_half * d[hi] + g.H2_8 * q[hi];
        r = vel[f] >= 0.0 ? left : right;
      }
    } else if (op == 3) {
      if (g.periodic) {
        const int k = i == N ? 0 : i;  // image node repeats node 0
        const int km = wrap(k - 1), kp = wrap(k + 1);
        const double dL = face_grad(g, u, q, node(km), node(k));
        const double dR = face_grad(g, u, q, node(k), node(kp));
        r = 0.5 * (dL + dR) - g.H16 * (q[node(kp)] - q[node(km)]);
      } else if (i == 0) {
        r = face_grad(g, u, q, node(0), node(1));
      } else if (i == N) {
        r = face_grad(g, u, q, node(N - 1), node(N));
      } else {
        const double dL = face_grad(g, u, q, node(i - 1), node(i));
        const double dR = face_grad(g, u, q, node(i), node(i + 1));
        r = 0.5 * (dL + dR) - g.H16 * (q[node(i + 1)] - q[node(i - 1)]);
      }
    } else {  // op 4: u holds the face array
      const int64_t fb = a * (int64_t)N * g.inner + b;
      auto face = [&](int k) { return u[fb + (int64_t)k * g.inner]; };
      if (g.periodic) {
        const int k = i == N ? 0 : i;
        r = (face(k) - face(wrap(k - 1))) * g.inv_H;
      } else {
        r = (i == 0 || i == N) ? 0.0 : (face(i) - face(i - 1)) * g.inv_H;
      }
    }
    out[o] = r;
  }
}

}  // namespace
}  // namespace tpf

using namespace tpf;

extern "C" tpf_status tpf_fccd_face_op(int op, int ndim, const int64_t* n_nodes, const double* L,
                                       const int* bc, int axis, const double* u, const double* q,
                                       const double* d, const double* vel, double* out,
                                       void* stream) {
  if (!n_nodes || !L || !bc || !u || !out) return fail(TPF_ERR_VALUE, "null argument");
  if (ndim < 1 || ndim > 3 || axis < 0 || axis >= ndim)
    return fail(TPF_ERR_VALUE, "bad ndim / axis");
  if (op < 0 || op > 5) return fail(TPF_ERR_VALUE, "FCCD face op must be 0..5");
  if (op != 4 && !q) return fail(TPF_ERR_VALUE, "this FCCD op needs q (the CCD f'')");
  if (op == 5 && (!d || !vel)) return fail(TPF_ERR_VALUE, "upwind face value needs d and vel");
  if (n_nodes[axis] < 3 || !(L[axis] > 0.0)) return fail(TPF_ERR_VALUE, "axis too short");
  FaceGeom g{};
  g.outer = 1;
  g.inner = 1;
  for (int a = 0; a < ndim; ++a) {
    if (n_nodes[a] < 1) return fail(TPF_ERR_VALUE, "bad n_nodes");
    if (a < axis) g.outer *= n_nodes[a];
    if (a > axis) g.inner *= n_nodes[a];
  }
  g.n = (int)n_nodes[axis];
  g.periodic = bc[axis] == TPF_BC_PERIODIC;
  const double H = L[axis] / (double)(g.n - 1);
  g.inv_H = 1.0 / H;
  g.H24 = H / 24.0;
  g.H2_16 = H * H / 16.0;
  g.H16 = H / 16.0;
  g.H_half = 0.5 * H;
  g.H2_8 = H * H / 8.0;
  const bool face_out = op <= 2 || op == 5;
  const int64_t total = g.outer * g.inner * (face_out ? g.n - 1 : g.n);
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((total + kBlock - 1) / kBlock, 8192));
  hipLaunchKernelGGL(k_fccd_face_op, dim3((unsigned)blocks), dim3(kBlock), 0, (hipStream_t)stream,
                     op, g, u, q, d, vel, out, total);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}
