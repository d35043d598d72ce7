// Closest-point quintic-Hermite field extension in 2-D
// (ClosestPointExtender._extend_2d, levelset/closest_point_extender.py:
// 136-228): q at nodes with phi >= 0 is replaced by the tensor-product
// quintic Hermite interpolant of (q, q', q'') at the closest point
// x - phi n, its cell found by binary search on the node coordinates and
// shifted (three times per axis) until its far node lies on the phi < 0 side.
//
// The closest point and the bracket are computed without contracting
// multiplies into FMAs, in the reference's order: on symmetry lines the
// closest point lies on a grid line to within rounding and the bracket
// depends on its last bit.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>

#include "reduce.hpp"
#include "tpf_common.hpp"

namespace tpf {
namespace {

struct ExtFields {
  const double *q, *qx, *qxx, *qy, *qyy, *qxy, *qxxy, *qxyy, *qxxyy;
};

// largest i with x[i] <= v (numpy searchsorted(..., side="right") - 1)
__device__ inline int bracket(const double* __restrict__ x, int n, double v) {
  int lo = 0, hi = n;  // count of x[i] <= v lies in [lo, hi]
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (x[mid] <= v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - 1;
}

__device__ inline double h5(double t, double Fa, double Ga, double Ha, double Fb, double Gb,
                            double Hb) {
  const double c0 = Fa, c1 = Ga, c2 = 0.5 * Ha;
  const double c3 = 10 * (Fb - Fa) - 6 * Ga - 4 * Gb + 0.5 * (Hb - 3 * Ha);
  const double c4 = 15 * (Fa - Fb) + 8 * Ga + 7 * Gb + 0.5 * (3 * Ha - 2 * Hb);
  const double c5 = 6 * (Fb - Fa) - 3 * (Ga + Gb) + 0.5 * (Hb - Ha);
  return c0 + t * (c1 + t * (c2 + t * (c3 + t * (c4 + t * c5))));
}

__global__ __launch_bounds__(kBlock) void k_extend_2d(int nx, int ny, const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      const double* __restrict__ phi,
                                                      const double* __restrict__ nxv,
                                                      const double* __restrict__ nyv, ExtFields F,
                                                      double* __restrict__ out) {
  const int64_t n = (int64_t)nx * ny;
  for (int64_t idx = (int64_t)blockIdx.x * kBlock + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * kBlock) {
    const int i = (int)(idx / ny), j = (int)(idx - (int64_t)i * ny);
    const double ph = phi[idx];
    if (!(ph >= 0.0)) {
      out[idx] = F.q[idx];
      continue;
    }
    const double pnx = ph * nxv[idx];
    const double pny = ph * nyv[idx];
    const double xc = fmin(fmax(x[i] - pnx, x[0]), x[nx - 1]);
    const double yc = fmin(fmax(y[j] - pny, y[0]), y[ny - 1]);
    int ix = min(max(bracket(x, nx, xc), 0), nx - 2);
    int iy = min(max(bracket(y, ny, yc), 0), ny - 2);
    for (int r = 0; r < 3; ++r) {
      const int ix1 = min(ix + 1, nx - 1);
      if (phi[(int64_t)ix1 * ny + iy] >= 0.0) ix = min(max(ix - 1, 0), nx - 2);
    }
    const int ix1 = min(ix + 1, nx - 1);
    for (int r = 0; r < 3; ++r) {
      const int iy1 = min(iy + 1, ny - 1);
      if (phi[(int64_t)ix * ny + iy1] >= 0.0) iy = min(max(iy - 1, 0), ny - 2);
    }
    const int iy1 = min(iy + 1, ny - 1);
    const double hx = x[ix1] - x[ix];
    const double hy = y[iy1] - y[iy];
    const double tx = (xc - x[ix]) / fmax(hx, 1e-30);
    const double ty = (yc - y[iy]) / fmax(hy, 1e-30);
    const double hx2 = hx * hx, hy2 = hy * hy;
    auto row = [&](const double* Fv, const double* Gv, const double* Hv, int jy) {
      const int64_t a = (int64_t)ix * ny + jy, b = (int64_t)ix1 * ny + jy;
      return h5(tx, Fv[a], hx * Gv[a], hx2 * Hv[a], Fv[b], hx * Gv[b], hx2 * Hv[b]);
    };
    const double v0 = row(F.q, F.qx, F.qxx, iy), d0 = row(F.qy, F.qxy, F.qxxy, iy),
                 dd0 = row(F.qyy, F.qxyy, F.qxxyy, iy);
    const double v1 = row(F.q, F.qx, F.qxx, iy1), d1 = row(F.qy, F.qxy, F.qxxy, iy1),
                 dd1 = row(F.qyy, F.qxyy, F.qxxyy, iy1);
    out[idx] = h5(ty, v0, hy * d0, hy2 * dd0, v1, hy * d1, hy2 * dd1);
  }
}

}  // namespace
}  // namespace tpf

using namespace tpf;

extern "C" tpf_status tpf_ns_extend_2d(tpf_ccd_plan* plan, const int64_t* n_nodes,
                                       const double* x, const double* y, const double* q,
                                       const double* phi, const double* nx, const double* ny,
                                       double* out, double* work, void* stream) {
  if (!plan || !n_nodes || !x || !y || !q || !phi || !nx || !ny || !out || !work)
    return fail(TPF_ERR_VALUE, "null argument");
  if (tpf::ccd_plan_has_metric(plan))
    return tpf::fail(TPF_ERR_VALUE, "uniform grids only (plan has a metric)");
  if (n_nodes[0] < 2 || n_nodes[1] < 2) return fail(TPF_ERR_VALUE, "axes need >= 2 nodes");
  if (out == q || out == phi) return fail(TPF_ERR_VALUE, "out must not alias q or phi");
  const int64_t n = n_nodes[0] * n_nodes[1];
  if (n >= ((int64_t)1 << 31)) return fail(TPF_ERR_VALUE, "grid too large (2^31 nodes)");
  double* D1 = work;          // qx, qy
  double* D2 = work + 2 * n;  // qxx, qyy
  double* qxy = work + 4 * n;
  double* qxxy = work + 5 * n;
  double* qxyy = work + 6 * n;
  double* qxxyy = work + 7 * n;
  TPF_TRY(tpf_ccd_apply(plan, q, D1, D2, 3u, stream));
  TPF_TRY(tpf_ccd_differentiate(plan, D1 + n, qxy, qxxy, 0, stream));
  TPF_TRY(tpf_ccd_differentiate(plan, D2 + n, qxyy, qxxyy, 0, stream));
  ExtFields F{q, D1, D2, D1 + n, D2 + n, qxy, qxxy, qxyy, qxxyy};
  const unsigned nb = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kBlock - 1) / kBlock, 4096));
  hipLaunchKernelGGL(k_extend_2d, dim3(nb), dim3(kBlock), 0, (hipStream_t)stream,
                     (int)n_nodes[0], (int)n_nodes[1], x, y, phi, nx, ny, F, out);
  TPF_HIP(hipGetLastError());
  return TPF_OK;
}
