// PPE handle, geometry and device-side solver state.
#pragma once

#include <cstdint>

#include "tpf_common.hpp"

namespace tpf {

// Node grid right-aligned into 3 slots (missing leading axes have n = 1);
// slot 2 is the contiguous axis.
struct PpeGeom {
  int n0, n1, n2;
  int64_t total;
  int64_t pin;
  double h[3];   // node spacing per slot
  double ih[3];  // 1 / h
};

// Device-resident PCG scalars (updated by the last block of each kernel).
struct PcgState {
  double rr, rz, bb, pq, alpha, beta, tol2;
  double true_rr, true_bb;
  double scale;             // mixed mode: FP64 residual norm carried into FP32
  int64_t iter, maxiter;
  int done;
  unsigned counter;
};

}  // namespace tpf

struct tpf_ppe {
  tpf::PpeGeom g{};
  hipStream_t user_stream = nullptr;
  hipStream_t stream = nullptr;
  int grid = 0;
  const double* rho = nullptr;
  double *dinv = nullptr, *b = nullptr, *r = nullptr, *p = nullptr, *q = nullptr;
  double* partials = nullptr;
  tpf::PcgState* st = nullptr;
  tpf::PcgState* st_host = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  const double* graph_x = nullptr;
  const double* graph_rho = nullptr;
  int graph_chunk = 0;
  // mixed-precision workspace (FP32 inner solve), allocated on first use
  float *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *x32 = nullptr, *dinv32 = nullptr,
        *rho32 = nullptr;
  tpf::PcgState* st_in = nullptr;       // inner (FP32) solve state
  tpf::PcgState* st_in_host = nullptr;
  const double* rho32_src = nullptr;    // rho the FP32 copies were made from
  hipGraphExec_t graph32_exec = nullptr;
  int graph32_chunk = 0;
  const double* graph32_rho = nullptr;
  void* gmres_work = nullptr;           // GMRES basis and line-solve scratch
  void* mg = nullptr;                   // multigrid hierarchy (FP32)
  void* fcg_st = nullptr;               // multigrid-FCG device scalars
  void* fcg_host = nullptr;             // pinned host copy
};

void tpf_gmres_release(tpf_ppe* s);
void tpf_mg_release(tpf_ppe* s);

namespace tpf {
// Multigrid hooks used by the multigrid-preconditioned CG of ppe.hip
// (implemented in mg.hip).
tpf_status mg_ensure(tpf_ppe* s, int max_levels);  // hierarchy for the current rho; synchronous
void mg_invalidate(tpf_ppe* s);                     // density changed: rebuild on next use
float* mg_fine_b(tpf_ppe* s);                       // level-0 right-hand side (cycle input)
float* mg_fine_x(tpf_ppe* s);                       // level-0 iterate (cycle output)
// x_0 = M(b_0): one cycle on stream `st` (not the null stream), replayed from
// a captured graph per (nu, gamma, coarse_sweeps, w_levels, stream); gamma
// repeats apply on the top w_levels levels (all when w_levels <= 0).
tpf_status mg_cycle(tpf_ppe* s, hipStream_t st, int nu, int gamma, int coarse_sweeps,
                    int w_levels);
// Learned coarse correction (A-DEF2) after a cycle: z += P^L (W^T E^-1 W (R^L (r - A z)))
// with W [k][n_level] on level `level`; E^-1 is uploaded by mg_deflate_setup.
tpf_status mg_deflate_setup(tpf_ppe* s, int k, int level, const double* einv_host);
tpf_status mg_deflate(tpf_ppe* s, hipStream_t st, const float* W, int k, int level);
tpf_status mg_level_total(tpf_ppe* s, int level, int64_t* total);
}  // namespace tpf
tpf_status tpf_ppe_solve_mixed(tpf_ppe* s, const double* rhs, double* x,
                               const tpf_solve_opts* opts, tpf_solve_report* rep);
