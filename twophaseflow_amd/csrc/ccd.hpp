// CCD line-solve plan shared by the host table builder and the HIP kernels.
//
// Every CCD line is the 2x2-block tridiagonal system of the reference
// (ccd/ccd_solver_helpers.py:227-314 wall, :406-423 periodic), written in
// the scaled unknowns (u, v) = (h f', h^2 f''), in which every block is
// independent of h.  A line of n unknowns is cut into P partitions (one
// wavefront each and P <= 16 when a lane holds a line; one lane each and
// P <= 64 when a wavefront holds a line); each partition is solved locally with a
// precomputed unpivoted block LU, and the partitions are coupled through a
// precomputed reduced (SPIKE) system whose inverse rows are stored per
// partition.
#pragma once

#include <cstdint>

namespace tpf {

constexpr int kMaxParts = 16;      // partitions of the lane-per-line kernel (wavefronts per workgroup)
constexpr int kMaxWaveParts = 64;  // partitions of the wavefront-per-line kernel (lanes)
constexpr int kMaxExplicitRows = 16;  // rows cap of the explicit-inverse kernel
constexpr int kRowTab = 16;     // doubles per row: Lf(4) iD(4) V(4) W(4)

// CCD coefficients, Table 1 of the paper (ccd/ccd_solver.py:70-76).
constexpr double kAlpha1 = 7.0 / 16.0;
constexpr double kA1 = 15.0 / 16.0;
constexpr double kB1 = 1.0 / 16.0;
constexpr double kBeta2 = -1.0 / 8.0;
constexpr double kA2 = 3.0;
constexpr double kB2 = -9.0 / 8.0;

// Kernel-side description of one axis (passed by value).
struct CcdAxisGeom {
  const double* rowtab;  // [P][M][kRowTab]
  const double* qtab;    // [P][4][NQ*4]: rows (left0,left1,right0,right1)
  const double* qtab_t;  // [4][NQ*4][P]: qtab with the partition fastest (wavefront
                         // kernel, lane = partition: one wavefront load covers 64
                         // consecutive doubles instead of 64 cache lines)
  const float* rowtab32; // FP32 copies (FP32-refined plans only)
  const float* qtab32;
  const double* ctab;    // [types][2][M][M+2]: partition inverse on first differences (explicit mode)
  int n;                 // unknowns per line (N periodic, N-1 wall)
  int N;                 // cells along the axis
  int P, M, NQ, d0;      // partitions, rows cap, coupling window, window start
  int rel;               // 1: window partition = p + d0 + t; 0: partition = t
  int periodic;
  int s[kMaxWaveParts + 1];  // partition starts (unknown index), s[P] = n
  int ctype[kMaxParts];      // partition -> ctab type (explicit mode)
  const double* rowtab_u;    // [ntypes][M][kRowTab]: the distinct partition row tables
  int ntypes;                // number of distinct partition row tables
  unsigned char rtype[kMaxWaveParts];  // partition -> its table in rowtab_u
  int uniform;               // 1: every partition has the same rows and relative coupling
                             //    rows (periodic, equal partitions): partition 0's tables serve all
  double inv_h, inv_h2;
  // wall closure in scaled unknowns (ccd/ccd_solver.py:408-456)
  double Lb[4], Ub[4], Ml[4], Mr[4];
  double cIl[4], cIIl[6], cIr[4], cIIr[6];
};

// Line addressing: line b starts at (b / inner) * outer + (b % inner) * inner_stride,
// and its nodes are `stride` apart.
struct CcdLines {
  int64_t n_lines, inner, outer, inner_stride, stride;
};

}  // namespace tpf
