// Host-side CCD plan objects (tables live on the host and on the device).
#pragma once

#include <vector>

#include "ccd.hpp"
#include "tpf_common.hpp"

namespace tpf {

struct CcdAxisHost {
  int wave = 0;                 // 1: one wavefront per line (lane = partition);
                                // 2: lane per line, explicit partition inverse
  int lpw = 1;                  // wave == 1: lines per wavefront (64 / P rounded)
  int wpg = 16;                 // wave == 1: wavefronts per workgroup
  CcdAxisGeom geom{};           // device pointers filled after upload
  std::vector<double> rowtab;   // [P][M][kRowTab]
  std::vector<double> qtab;     // [P][4][NQ*4]
  std::vector<double> ctab;     // [types][2][M][M+2] (M <= 16), geom.ctype maps partitions
  std::vector<double> rowtab_u; // distinct [M][kRowTab] blocks of rowtab, geom.rtype maps partitions
  double* d_rowtab_u = nullptr;
  double trunc_coeff = 0.0;     // largest dropped reduced-system coefficient (relative)
  double* d_rowtab = nullptr;
  double* d_qtab = nullptr;
  double* d_qtab_t = nullptr;   // wavefront kernel on non-uniform axes: qtab transposed
  float* d_rowtab32 = nullptr;  // FP32-refined plans
  float* d_qtab32 = nullptr;
  double* d_ctab = nullptr;
};

// Smallest rows-per-partition cap in {8,16,32,64} that fits n unknowns into
// at most kMaxParts partitions; -1 when none does.
int choose_rows_cap(int n);

// Builds the SPIKE tables of one axis.  rows_cap / parts_override <= 0 pick
// the defaults.
tpf_status build_axis_tables(CcdAxisHost& ax, int n_nodes, double h, int periodic,
                             int rows_cap, int parts_override);

// Scalar model of the kernel on one line with node stride `stride`.
void host_model_line(const CcdAxisHost& ax, const double* f, int64_t stride, double* d1,
                     double* d2);
// The same for the explicit partition-inverse kernel (tables of mode 3).
void host_model_line_explicit(const CcdAxisHost& ax, const double* f, int64_t stride, double* d1,
                              double* d2);

}  // namespace tpf

struct tpf_ccd_plan {
  int ndim = 0;
  int64_t n_nodes[3] = {1, 1, 1};
  double h[3] = {1, 1, 1};
  int bc[3] = {0, 0, 0};
  int precision = 0;
  int rows_cap[3] = {0, 0, 0};  // per-axis rows-per-partition override (0: default)
  int mode[3] = {0, 0, 0};      // per-axis kernel: 0 auto, 1 lane per line, 2 wavefront per line
  int store_align[3] = {-1, -1, -1};  // explicit kernel's store chunk (elements; -1 default 8)
  tpf::CcdAxisHost axes[3];
  // non-uniform axes: device [3][n_nodes[a]] = J, J*J, J*dJ (J = dxi/dx,
  // dJ = dJ/dxi), applied after the line solves; nullptr: no metric
  double* d_metric[3] = {nullptr, nullptr, nullptr};
  double* scratch = nullptr;  // d1 in xi of d2-only calls on non-uniform axes
  int64_t scratch_n = 0;
};
