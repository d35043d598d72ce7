// Host side of the CCD plan: SPIKE tables in extended precision, device
// upload, and a scalar host model of the kernel algorithm (tests only).
//
// The tables replace the reference's dense per-axis LU factorisations
// (ccd/ccd_solver_helpers.py:227-314 wall, :406-423 periodic); the block LU
// per partition follows the block-Thomas template of ccd/block_tridiag.py:
// 43-104, including its fail-closed singular-pivot check (:167-174).
#include "ccd_plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace tpf {

namespace {

using ld = long double;

struct M2 {
  ld a, b, c, d;  // [[a, b], [c, d]]
};
inline M2 mul(const M2& x, const M2& y) {
  return {x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d,
          x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
}
inline M2 sub(const M2& x, const M2& y) { return {x.a - y.a, x.b - y.b, x.c - y.c, x.d - y.d}; }
inline M2 add(const M2& x, const M2& y) { return {x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d}; }
inline ld det(const M2& x) { return x.a * x.d - x.b * x.c; }
inline M2 inv(const M2& x) {
  ld id = 1.0L / det(x);
  return {x.d * id, -x.b * id, -x.c * id, x.a * id};
}
const M2 kI{1, 0, 0, 1};
const M2 kZ{0, 0, 0, 0};

// Scaled blocks: S X S^-1 with S = diag(h, h^2) (ccd_solver_helpers.py:235-240).
const M2 kLb{(ld)kAlpha1, (ld)kB1, (ld)kB2, (ld)kBeta2};
const M2 kUb{(ld)kAlpha1, -(ld)kB1, -(ld)kB2, (ld)kBeta2};
// Scaled closure matrices (ccd/ccd_solver.py:429-430, 452-453).
const M2 kMl{-1.5L, 1.5L, 0, 0};
const M2 kMr{-1.5L, -1.5L, 0, 0};

// Gauss-Jordan inverse with partial pivoting, row-major n x n, in place.
bool invert(std::vector<ld>& a, int n) {
  std::vector<ld> b(size_t(n) * n, 0.0L);
  for (int i = 0; i < n; ++i) b[size_t(i) * n + i] = 1.0L;
  for (int col = 0; col < n; ++col) {
    int piv = col;
    for (int r = col + 1; r < n; ++r)
      if (std::fabs(a[size_t(r) * n + col]) > std::fabs(a[size_t(piv) * n + col])) piv = r;
    if (std::fabs(a[size_t(piv) * n + col]) < 1e-300L) return false;
    if (piv != col)
      for (int k = 0; k < n; ++k) {
        std::swap(a[size_t(piv) * n + k], a[size_t(col) * n + k]);
        std::swap(b[size_t(piv) * n + k], b[size_t(col) * n + k]);
      }
    ld ip = 1.0L / a[size_t(col) * n + col];
    for (int k = 0; k < n; ++k) {
      a[size_t(col) * n + k] *= ip;
      b[size_t(col) * n + k] *= ip;
    }
    for (int r = 0; r < n; ++r) {
      if (r == col) continue;
      ld f = a[size_t(r) * n + col];
      if (f == 0.0L) continue;
      for (int k = 0; k < n; ++k) {
        a[size_t(r) * n + k] -= f * a[size_t(col) * n + k];
        b[size_t(r) * n + k] -= f * b[size_t(col) * n + k];
      }
    }
  }
  a.swap(b);
  return true;
}

inline void put(double* dst, const M2& m) {
  dst[0] = (double)m.a;
  dst[1] = (double)m.b;
  dst[2] = (double)m.c;
  dst[3] = (double)m.d;
}

}  // namespace

int choose_rows_cap(int n) {
  // Lines of 33-256 unknowns take 16 rows: the contiguous axis is staged
  // through LDS only at 16 rows, and on 65^3 and 129^3 the all-axes call ran
  // 2.1-2.4x faster with 16 than with 8 (the strided axes alone are within
  // 5 % either way).
  if (n > 32 && n <= 16 * kMaxParts) return 16;
  for (int m : {8, 16, 32, 64})
    if ((n + m - 1) / m <= kMaxParts) return m;
  return -1;
}

tpf_status build_axis_tables(CcdAxisHost& ax, int n_nodes, double h, int periodic,
                             int rows_cap, int parts_override) {
  const int N = n_nodes - 1;
  if (periodic) {
    if (N < 3)
      return fail(TPF_ERR_VALUE, "periodic CCD needs >= 3 unique nodes; got " + std::to_string(N));
  } else if (n_nodes < 6) {
    return fail(TPF_ERR_VALUE,
                "CCD wall Eq-II boundary closure requires n_pts >= 6 for the six-point "
                "O(h^4) formula; no lower-order fallback exists (got n_pts=" +
                    std::to_string(n_nodes) + ")");
  }
  const int n = periodic ? N : N - 1;
  int M = rows_cap > 0 ? rows_cap : choose_rows_cap(n);
  if (M <= 0) return fail(TPF_ERR_VALUE, "CCD line too long for one workgroup: n=" + std::to_string(n));
  int P = parts_override > 0 ? parts_override : (n + M - 1) / M;
  if (P > kMaxWaveParts || (n + P - 1) / P > M)
    return fail(TPF_ERR_VALUE, "CCD partitioning does not fit: n=" + std::to_string(n) +
                                   " P=" + std::to_string(P) + " M=" + std::to_string(M));
  const int min_rows = periodic ? 2 : 4;
  if (P > 1 && n / P < min_rows)
    return fail(TPF_ERR_VALUE, "CCD partitions too short for n=" + std::to_string(n));

  CcdAxisGeom& g = ax.geom;
  std::memset(&g, 0, sizeof(g));
  g.n = n;
  g.N = N;
  g.P = P;
  g.M = M;
  g.periodic = periodic;
  for (int p = 0; p <= P; ++p) g.s[p] = (int)((int64_t)n * p / P);
  g.inv_h = 1.0 / h;
  g.inv_h2 = 1.0 / (h * h);
  put(g.Lb, kLb);
  put(g.Ub, kUb);
  put(g.Ml, kMl);
  put(g.Mr, kMr);
  const double cI[4] = {-23.0 / 6.0, 21.0 / 4.0, -1.5, 1.0 / 12.0};
  const double cII[6] = {45.0 / 12.0, -154.0 / 12.0, 214.0 / 12.0, -156.0 / 12.0, 61.0 / 12.0,
                         -10.0 / 12.0};
  for (int k = 0; k < 4; ++k) {
    g.cIl[k] = cI[k];
    g.cIr[k] = -cI[k];
  }
  for (int k = 0; k < 6; ++k) g.cIIl[k] = g.cIIr[k] = cII[k];

  // Local block LU + spikes per partition.
  ax.rowtab.assign(size_t(P) * M * kRowTab, 0.0);
  ax.ctab.clear();
  std::vector<M2> Vfirst(P), Vlast(P), Wfirst(P), Wlast(P);
  for (int p = 0; p < P; ++p) {
    const int s0 = g.s[p], m = g.s[p + 1] - g.s[p];
    std::vector<M2> Lf(m, kZ), iD(m), D(m, kI);
    for (int r = 0; r < m; ++r) {
      int j = s0 + r;
      if (!periodic && j == 0) D[r] = add(D[r], mul(kLb, kMl));
      if (!periodic && j == n - 1) D[r] = add(D[r], mul(kUb, kMr));
    }
    M2 Dp = D[0];
    for (int r = 0; r < m; ++r) {
      if (r > 0) {
        Lf[r] = mul(kLb, iD[r - 1]);
        Dp = sub(D[r], mul(Lf[r], kUb));
      }
      ld dt = det(Dp);
      if (!std::isfinite((double)dt) || std::fabs(dt) < 1e-300L)
        return fail(TPF_ERR_LINALG, "CCD block-tridiagonal factorization encountered a singular "
                                    "pivot at block " + std::to_string(s0 + r) +
                                    "; no pseudo-inverse fallback is permitted");
      iD[r] = inv(Dp);
    }
    // Spike columns: V = A^-1 [Lo;0..], W = A^-1 [0..;Up] (2 columns each).
    const bool has_left = periodic || p > 0;
    const bool has_right = periodic || p < P - 1;
    auto solve_cols = [&](int at_row, const M2& B, std::vector<M2>& X) {
      std::vector<M2> y(m, kZ);
      for (int r = 0; r < m; ++r) {
        M2 rhs = (r == at_row) ? B : kZ;
        y[r] = r ? sub(rhs, mul(Lf[r], y[r - 1])) : rhs;
      }
      X.assign(m, kZ);
      for (int r = m - 1; r >= 0; --r) {
        M2 t = (r == m - 1) ? y[r] : sub(y[r], mul(kUb, X[r + 1]));
        X[r] = mul(iD[r], t);
      }
    };
    std::vector<M2> V, W;
    solve_cols(0, has_left ? kLb : kZ, V);
    solve_cols(m - 1, has_right ? kUb : kZ, W);
    Vfirst[p] = V[0];
    Vlast[p] = V[m - 1];
    Wfirst[p] = W[0];
    Wlast[p] = W[m - 1];
    for (int r = 0; r < m; ++r) {
      double* row = &ax.rowtab[(size_t(p) * M + r) * kRowTab];
      put(row + 0, Lf[r]);
      put(row + 4, iD[r]);
      put(row + 8, V[r]);
      put(row + 12, W[r]);
    }
    if (M <= kMaxExplicitRows && P <= kMaxParts) {
      // Partition inverse applied to the line values, in first differences:
      // y_r = sum_k C[r][k] fv[k] with C[r][.] the local solution (no
      // neighbour coupling) for fv = e_k, wall closures included.  Every row
      // of C sums to 0 (constants have no derivative), so
      // y_r = sum_{k <= m} S[r][k] (fv[k+1] - fv[k]) with S[r][k] =
      // -sum_{j <= k} C[r][j]; the differences of smooth data are small and
      // exact, which keeps f'' free of the cancellation a sum over raw values
      // would carry.  Identical partitions share a type.
      const int K = M + 2;
      std::vector<ld> Cl(size_t(M) * K * 2, 0.0L);
      std::vector<double> Cp(size_t(M) * K * 2, 0.0);
      for (int k = 0; k < m + 2; ++k) {
        std::vector<ld> fv(m + 2, 0.0L), y(2 * m), x(2 * m);
        fv[k] = 1.0L;
        for (int r = 0; r < m; ++r) {
          ld r0 = (ld)kA1 * (fv[r + 2] - fv[r]);
          ld r1 = (ld)kA2 * ((fv[r] - 2.0L * fv[r + 1]) + fv[r + 2]);
          if (!periodic && s0 + r == 0) {
            ld b0 = 0, b1 = 0;
            for (int q = 0; q < 4; ++q) b0 += (ld)g.cIl[q] * fv[q];
            for (int q = 0; q < 6; ++q) b1 += (ld)g.cIIl[q] * fv[q];
            r0 -= kLb.a * b0 + kLb.b * b1;
            r1 -= kLb.c * b0 + kLb.d * b1;
          }
          if (!periodic && s0 + r == n - 1) {
            ld b0 = 0, b1 = 0;
            for (int q = 0; q < 4; ++q) b0 += (ld)g.cIr[q] * fv[m + 1 - q];
            for (int q = 0; q < 6; ++q) b1 += (ld)g.cIIr[q] * fv[m + 1 - q];
            r0 -= kUb.a * b0 + kUb.b * b1;
            r1 -= kUb.c * b0 + kUb.d * b1;
          }
          if (r > 0) {
            r0 -= Lf[r].a * y[2 * r - 2] + Lf[r].b * y[2 * r - 1];
            r1 -= Lf[r].c * y[2 * r - 2] + Lf[r].d * y[2 * r - 1];
          }
          y[2 * r] = r0;
          y[2 * r + 1] = r1;
        }
        for (int r = m - 1; r >= 0; --r) {
          ld a = y[2 * r], c = y[2 * r + 1];
          if (r < m - 1) {
            a -= kUb.a * x[2 * r + 2] + kUb.b * x[2 * r + 3];
            c -= kUb.c * x[2 * r + 2] + kUb.d * x[2 * r + 3];
          }
          x[2 * r] = iD[r].a * a + iD[r].b * c;
          x[2 * r + 1] = iD[r].c * a + iD[r].d * c;
        }
        for (int r = 0; r < m; ++r) {
          Cl[(size_t(r) * K + k) * 2 + 0] = x[2 * r];
          Cl[(size_t(r) * K + k) * 2 + 1] = x[2 * r + 1];
        }
      }
      for (int r = 0; r < m; ++r)  // component-major: [c][r][k]
        for (int c = 0; c < 2; ++c) {
          ld acc = 0.0L;
          for (int k = 0; k <= m; ++k) {
            acc += Cl[(size_t(r) * K + k) * 2 + c];
            Cp[(size_t(c) * M + r) * K + k] = (double)(-acc);
          }
        }
      const size_t tsz = Cp.size();
      int type = -1;
      for (size_t t = 0; t * tsz < ax.ctab.size() && type < 0; ++t)
        if (std::memcmp(&ax.ctab[t * tsz], Cp.data(), tsz * sizeof(double)) == 0) type = (int)t;
      if (type < 0) {
        type = (int)(ax.ctab.size() / tsz);
        ax.ctab.insert(ax.ctab.end(), Cp.begin(), Cp.end());
      }
      g.ctype[p] = type;
    }
  }

  // Distinct partition row tables (wall lines: first, last and one or two
  // interior partition sizes), so kernels with per-partition tables can stage
  // them in LDS.
  ax.rowtab_u.clear();
  {
    const size_t bsz = size_t(M) * kRowTab;
    for (int p = 0; p < P; ++p) {
      const double* blk = &ax.rowtab[size_t(p) * bsz];
      int type = -1;
      for (size_t t = 0; t * bsz < ax.rowtab_u.size() && type < 0; ++t)
        if (std::memcmp(&ax.rowtab_u[t * bsz], blk, bsz * sizeof(double)) == 0) type = (int)t;
      if (type < 0) {
        type = (int)(ax.rowtab_u.size() / bsz);
        ax.rowtab_u.insert(ax.rowtab_u.end(), blk, blk + bsz);
      }
      g.rtype[p] = (unsigned char)type;
    }
    g.ntypes = (int)(ax.rowtab_u.size() / bsz);
  }

  // Reduced system (I + K) Z = G over Z = [first(p), last(p)]_p, 2 comps each.
  const int nz = 4 * P;
  std::vector<ld> A(size_t(nz) * nz, 0.0L);
  auto addblk = [&](int row_blk, int col_blk, const M2& m2) {
    A[size_t(2 * row_blk) * nz + 2 * col_blk] += m2.a;
    A[size_t(2 * row_blk) * nz + 2 * col_blk + 1] += m2.b;
    A[size_t(2 * row_blk + 1) * nz + 2 * col_blk] += m2.c;
    A[size_t(2 * row_blk + 1) * nz + 2 * col_blk + 1] += m2.d;
  };
  for (int p = 0; p < P; ++p) {
    addblk(2 * p, 2 * p, kI);
    addblk(2 * p + 1, 2 * p + 1, kI);
    const bool has_left = periodic || p > 0;
    const bool has_right = periodic || p < P - 1;
    if (has_left) {
      int lb = 2 * ((p - 1 + P) % P) + 1;  // last of previous partition
      addblk(2 * p, lb, Vfirst[p]);
      addblk(2 * p + 1, lb, Vlast[p]);
    }
    if (has_right) {
      int rb = 2 * ((p + 1) % P);  // first of next partition
      addblk(2 * p, rb, Wfirst[p]);
      addblk(2 * p + 1, rb, Wlast[p]);
    }
  }
  if (!invert(A, nz)) return fail(TPF_ERR_LINALG, "CCD reduced system is singular");

  // Coupling window: smallest D whose dropped coefficients are negligible.
  ld amax = 0;
  for (ld v : A) amax = std::max(amax, std::fabs(v));
  auto row_of = [&](int p, int o) -> int {  // o: 0,1 left comps; 2,3 right comps
    if (o < 2) return 2 * (2 * ((p - 1 + P) % P) + 1) + o;
    return 2 * (2 * ((p + 1) % P)) + (o - 2);
  };
  auto dist = [&](int p, int q) {
    int d = q - p;
    if (periodic) {
      d = ((d % P) + P) % P;
      if (d > P / 2) d -= P;
    }
    return d;
  };
  int D = 0;
  double dropped = 0.0;
  for (D = 0; 2 * D + 1 < P; ++D) {
    ld worst = 0;
    for (int p = 0; p < P; ++p)
      for (int o = 0; o < 4; ++o) {
        if (!periodic && ((o < 2 && p == 0) || (o >= 2 && p == P - 1))) continue;  // unused row
        for (int q = 0; q < P; ++q)
          if (std::abs(dist(p, q)) > D)
            for (int c = 0; c < 4; ++c)
              worst = std::max(worst, std::fabs(A[size_t(row_of(p, o)) * nz + 4 * q + c]));
      }
    if (worst <= 1e-19L * amax) {
      dropped = (double)(worst / amax);
      break;
    }
  }
  if (2 * D + 1 >= P) {  // full coupling, absolute partition index
    g.NQ = P;
    g.d0 = 0;
    g.rel = 0;
    dropped = 0.0;
  } else {
    g.NQ = 2 * D + 1;
    g.d0 = -D;
    g.rel = 1;
  }
  ax.trunc_coeff = dropped;
  // Periodic lines cut into equal partitions: the block rows are identical
  // and the reduced system is circulant, so with a relative coupling window
  // every partition's tables equal partition 0's (the reduced inverse's rows
  // agree to its rounding).
  g.uniform = periodic && g.rel && n % P == 0 ? 1 : 0;
  ax.qtab.assign(size_t(P) * 4 * g.NQ * 4, 0.0);
  for (int p = 0; p < P; ++p) {
    const bool has_left = periodic || p > 0;
    const bool has_right = periodic || p < P - 1;
    for (int o = 0; o < 4; ++o) {
      if ((o < 2 && !has_left) || (o >= 2 && !has_right)) continue;
      for (int t = 0; t < g.NQ; ++t) {
        int q = g.rel ? p + g.d0 + t : t;
        if (periodic) q = ((q % P) + P) % P;
        else if (q < 0 || q >= P) continue;
        for (int c = 0; c < 4; ++c)
          ax.qtab[((size_t(p) * 4 + o) * g.NQ + t) * 4 + c] =
              (double)A[size_t(row_of(p, o)) * nz + 4 * q + c];
      }
    }
  }
  return TPF_OK;
}

// Scalar model of the kernel, one line at a time (tests of the tables only).
void host_model_line(const CcdAxisHost& ax, const double* f, int64_t stride, double* d1,
                     double* d2) {
  const CcdAxisGeom& g = ax.geom;
  const int P = g.P, M = g.M, NQ = g.NQ;
  std::vector<double> G(size_t(4) * P), X(size_t(2) * g.n);
  auto node = [&](int j) { return f[int64_t(j) * stride]; };
  for (int p = 0; p < P; ++p) {
    const int s0 = g.s[p], m = g.s[p + 1] - s0;
    std::vector<double> fv(m + 2), y(2 * m), x(2 * m);
    for (int k = 0; k < m + 2; ++k) {
      int j = g.periodic ? ((s0 - 1 + k) % g.N + g.N) % g.N : s0 + k;
      fv[k] = node(j);
    }
    for (int r = 0; r < m; ++r) {
      double r0 = kA1 * (fv[r + 2] - fv[r]);
      double r1 = kA2 * ((fv[r] - 2.0 * fv[r + 1]) + fv[r + 2]);
      if (!g.periodic && s0 + r == 0) {
        double b0 = 0, b1 = 0;
        for (int k = 0; k < 4; ++k) b0 += g.cIl[k] * fv[k];
        for (int k = 0; k < 6; ++k) b1 += g.cIIl[k] * fv[k];
        r0 -= g.Lb[0] * b0 + g.Lb[1] * b1;
        r1 -= g.Lb[2] * b0 + g.Lb[3] * b1;
      }
      if (!g.periodic && s0 + r == g.n - 1) {
        double b0 = 0, b1 = 0;
        for (int k = 0; k < 4; ++k) b0 += g.cIr[k] * fv[m + 1 - k];
        for (int k = 0; k < 6; ++k) b1 += g.cIIr[k] * fv[m + 1 - k];
        r0 -= g.Ub[0] * b0 + g.Ub[1] * b1;
        r1 -= g.Ub[2] * b0 + g.Ub[3] * b1;
      }
      const double* t = &ax.rowtab[(size_t(p) * M + r) * kRowTab];
      if (r > 0) {
        r0 -= t[0] * y[2 * r - 2] + t[1] * y[2 * r - 1];
        r1 -= t[2] * y[2 * r - 2] + t[3] * y[2 * r - 1];
      }
      y[2 * r] = r0;
      y[2 * r + 1] = r1;
    }
    for (int r = m - 1; r >= 0; --r) {
      double a = y[2 * r], b = y[2 * r + 1];
      if (r < m - 1) {
        a -= g.Ub[0] * x[2 * r + 2] + g.Ub[1] * x[2 * r + 3];
        b -= g.Ub[2] * x[2 * r + 2] + g.Ub[3] * x[2 * r + 3];
      }
      const double* t = &ax.rowtab[(size_t(p) * M + r) * kRowTab + 4];
      x[2 * r] = t[0] * a + t[1] * b;
      x[2 * r + 1] = t[2] * a + t[3] * b;
    }
    G[4 * p + 0] = x[0];
    G[4 * p + 1] = x[1];
    G[4 * p + 2] = x[2 * m - 2];
    G[4 * p + 3] = x[2 * m - 1];
    for (int r = 0; r < m; ++r) {
      X[2 * (s0 + r)] = x[2 * r];
      X[2 * (s0 + r) + 1] = x[2 * r + 1];
    }
  }
  for (int p = 0; p < P; ++p) {
    const int s0 = g.s[p], m = g.s[p + 1] - s0;
    double lr[4] = {0, 0, 0, 0};
    for (int o = 0; o < 4; ++o)
      for (int t = 0; t < NQ; ++t) {
        int q = g.rel ? p + g.d0 + t : t;
        q = g.periodic ? ((q % P) + P) % P : std::min(std::max(q, 0), P - 1);
        for (int c = 0; c < 4; ++c)
          lr[o] += ax.qtab[((size_t(p) * 4 + o) * NQ + t) * 4 + c] * G[4 * q + c];
      }
    for (int r = 0; r < m; ++r) {
      const double* t = &ax.rowtab[(size_t(p) * M + r) * kRowTab + 8];
      double& u = X[2 * (s0 + r)];
      double& v = X[2 * (s0 + r) + 1];
      u -= t[0] * lr[0] + t[1] * lr[1] + t[4] * lr[2] + t[5] * lr[3];
      v -= t[2] * lr[0] + t[3] * lr[1] + t[6] * lr[2] + t[7] * lr[3];
    }
  }
  const int off = g.periodic ? 0 : 1;
  for (int j = 0; j < g.n; ++j) {
    if (d1) d1[int64_t(j + off) * stride] = X[2 * j] * g.inv_h;
    if (d2) d2[int64_t(j + off) * stride] = X[2 * j + 1] * g.inv_h2;
  }
  if (g.periodic) {
    if (d1) d1[int64_t(g.N) * stride] = X[0] * g.inv_h;
    if (d2) d2[int64_t(g.N) * stride] = X[1] * g.inv_h2;
  } else {
    double b0 = 0, b1 = 0, c0 = 0, c1 = 0;
    for (int k = 0; k < 4; ++k) {
      b0 += g.cIl[k] * node(k);
      c0 += g.cIr[k] * node(g.N - k);
    }
    for (int k = 0; k < 6; ++k) {
      b1 += g.cIIl[k] * node(k);
      c1 += g.cIIr[k] * node(g.N - k);
    }
    const double *x1 = &X[0], *xl = &X[2 * (g.n - 1)];
    double u0 = g.Ml[0] * x1[0] + g.Ml[1] * x1[1] + b0;
    double v0 = g.Ml[2] * x1[0] + g.Ml[3] * x1[1] + b1;
    double uN = g.Mr[0] * xl[0] + g.Mr[1] * xl[1] + c0;
    double vN = g.Mr[2] * xl[0] + g.Mr[3] * xl[1] + c1;
    if (d1) {
      d1[0] = u0 * g.inv_h;
      d1[int64_t(g.N) * stride] = uN * g.inv_h;
    }
    if (d2) {
      d2[0] = v0 * g.inv_h2;
      d2[int64_t(g.N) * stride] = vN * g.inv_h2;
    }
  }
}

// Scalar model of the explicit-inverse kernel on one line (tests of ctab).
void host_model_line_explicit(const CcdAxisHost& ax, const double* f, int64_t stride, double* d1,
                              double* d2) {
  const CcdAxisGeom& g = ax.geom;
  const int P = g.P, M = g.M, NQ = g.NQ, K = M + 2;
  std::vector<double> G(size_t(4) * P), fvs(size_t(P) * K, 0.0);
  auto node = [&](int j) { return f[int64_t(j) * stride]; };
  auto row = [&](int p, int r, int c) {  // sum_k S[r][k] (fv[k+1] - fv[k])
    const double* S = &ax.ctab[size_t(g.ctype[p]) * M * K * 2];
    const double* fv = &fvs[size_t(p) * K];
    double acc = 0.0;
    for (int k = 0; k + 1 < K; ++k) acc += S[(size_t(c) * M + r) * K + k] * (fv[k + 1] - fv[k]);
    return acc;
  };
  for (int p = 0; p < P; ++p) {
    const int s0 = g.s[p], m = g.s[p + 1] - s0;
    for (int k = 0; k < m + 2; ++k)
      fvs[size_t(p) * K + k] = node(g.periodic ? ((s0 - 1 + k) % g.N + g.N) % g.N : s0 + k);
    G[4 * p + 0] = row(p, 0, 0);
    G[4 * p + 1] = row(p, 0, 1);
    G[4 * p + 2] = row(p, m - 1, 0);
    G[4 * p + 3] = row(p, m - 1, 1);
  }
  std::vector<double> X(size_t(2) * g.n);
  for (int p = 0; p < P; ++p) {
    const int s0 = g.s[p], m = g.s[p + 1] - s0;
    double lr[4] = {0, 0, 0, 0};
    for (int o = 0; o < 4; ++o)
      for (int t = 0; t < NQ; ++t) {
        int q = g.rel ? p + g.d0 + t : t;
        q = g.periodic ? ((q % P) + P) % P : std::min(std::max(q, 0), P - 1);
        for (int c = 0; c < 4; ++c)
          lr[o] += ax.qtab[((size_t(p) * 4 + o) * NQ + t) * 4 + c] * G[4 * q + c];
      }
    for (int r = 0; r < m; ++r) {
      const double* t = &ax.rowtab[(size_t(p) * M + r) * kRowTab + 8];
      X[2 * (s0 + r)] = row(p, r, 0) - (t[0] * lr[0] + t[1] * lr[1] + t[4] * lr[2] + t[5] * lr[3]);
      X[2 * (s0 + r) + 1] =
          row(p, r, 1) - (t[2] * lr[0] + t[3] * lr[1] + t[6] * lr[2] + t[7] * lr[3]);
    }
  }
  const int off = g.periodic ? 0 : 1;
  for (int j = 0; j < g.n; ++j) {
    if (d1) d1[int64_t(j + off) * stride] = X[2 * j] * g.inv_h;
    if (d2) d2[int64_t(j + off) * stride] = X[2 * j + 1] * g.inv_h2;
  }
  if (g.periodic) {
    if (d1) d1[int64_t(g.N) * stride] = X[0] * g.inv_h;
    if (d2) d2[int64_t(g.N) * stride] = X[1] * g.inv_h2;
    return;
  }
  double b0 = 0, b1 = 0, c0 = 0, c1 = 0;
  for (int k = 0; k < 4; ++k) {
    b0 += g.cIl[k] * node(k);
    c0 += g.cIr[k] * node(g.N - k);
  }
  for (int k = 0; k < 6; ++k) {
    b1 += g.cIIl[k] * node(k);
    c1 += g.cIIr[k] * node(g.N - k);
  }
  const double *x1 = &X[0], *xl = &X[2 * (g.n - 1)];
  if (d1) {
    d1[0] = (g.Ml[0] * x1[0] + g.Ml[1] * x1[1] + b0) * g.inv_h;
    d1[int64_t(g.N) * stride] = (g.Mr[0] * xl[0] + g.Mr[1] * xl[1] + c0) * g.inv_h;
  }
  if (d2) {
    d2[0] = (g.Ml[2] * x1[0] + g.Ml[3] * x1[1] + b1) * g.inv_h2;
    d2[int64_t(g.N) * stride] = (g.Mr[2] * xl[0] + g.Mr[3] * xl[1] + c1) * g.inv_h2;
  }
}

}  // namespace tpf
