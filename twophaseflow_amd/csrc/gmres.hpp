// Operator-independent restarted GMRES (SciPy 1.15 semantics), shared by the
// FVM and FCCD PPE handles.  Implemented in ppe_gmres.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>

#include "tpf_common.hpp"

namespace tpf {

// out = Op(in); in and out are distinct device vectors of the solve's length.
using GmresOp = std::function<tpf_status(const double* in, double* out)>;

// Sizes the handle's basis (restart m), vectors of length n, block partials
// for `grid_blocks` blocks and `nscratch` extra n-vectors; keeps an existing
// work area that is large enough.
tpf_status gmres_work(void** slot, int64_t n, int m, int grid_blocks, int nscratch);
void gmres_work_free(void** slot);
// The work area's right-hand-side vector (length n): callers stage b here.
double* gmres_rhs(void* work);
// The first of the work area's `nscratch` extra vectors (length n each).
double* gmres_scratch(void* work);
// Left-preconditioned GMRES(m) on A x = b with SciPy 1.15's stopping rules
// (inner: preconditioned residual estimate <= ptol; outer: true
// ||b - A x|| <= tol ||b||; maxiter counts restart cycles).  x holds the
// initial guess; r.converged = 0 when the budget ends above tol.
tpf_status gmres_core(void** work, hipStream_t st, unsigned grid, int64_t n, int m,
                      const GmresOp& apply, const GmresOp& psolve, const double* b, double* x,
                      double tol, int64_t maxiter, tpf_solve_report& r);

}  // namespace tpf
