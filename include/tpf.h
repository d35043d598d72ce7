/*
 * tpf.h -- C ABI of the MI355X-native CCD operator and variable-density PPE.
 *
 * Plain pointers and sizes only.  Every array argument is a DEVICE pointer
 * (hipMalloc'd or a torch CUDA tensor's data_ptr) holding a C-order FP64
 * field of shape n_nodes[0] x ... x n_nodes[ndim-1] (node-centred, N+1 nodes
 * per axis, as Grid.shape in the reference).  `stream` is a hipStream_t
 * passed as void* (NULL = default stream).  Calls are asynchronous on that
 * stream unless stated otherwise.  Handles are not thread-safe.
 *
 * Every function returns a tpf_status; tpf_last_error() gives the message of
 * the last failure on the calling thread.  Reference file paths below are
 * relative to src/twophase of t0m0h1r0/TwoPhaseFlow.
 */
#ifndef TPF_H
#define TPF_H

#include <stdint.h>

#if defined(__GNUC__)
#define TPF_API __attribute__((visibility("default")))
#else
#define TPF_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  TPF_OK = 0,
  TPF_ERR_VALUE = 1,      /* bad argument/shape   -> ValueError   (ccd_solver.py:459-465, fvm_matrixfree.py:84-89) */
  TPF_ERR_RUNTIME = 2,    /* no convergence, HIP  -> RuntimeError (fvm_matrixfree.py:250-256) */
  TPF_ERR_LINALG = 3      /* singular pivot       -> LinAlgError  (block_tridiag.py:167-174) */
} tpf_status;

enum { TPF_BC_WALL = 0, TPF_BC_PERIODIC = 1 };
enum { TPF_PREC_FP64 = 0, TPF_PREC_FP32_REFINED = 1 };

TPF_API const char* tpf_last_error(void);
TPF_API int tpf_version(void);

/* ---------------------------------------------------------------- CCD --- */
typedef struct tpf_ccd_plan tpf_ccd_plan;

/* Replaces CCDSolver.__init__ + the per-axis factorisations
 * (ccd/ccd_solver.py:99-120, ccd/ccd_solver_helpers.py:227-314, 406-423).
 * h[a] = L[a]/N[a]; bc[a] in {TPF_BC_WALL, TPF_BC_PERIODIC}.
 * Wall axes need n_nodes >= 6 (TPF_ERR_VALUE otherwise); periodic >= 4.
 * precision: TPF_PREC_FP64, or TPF_PREC_FP32_REFINED (FP32 solves under
 * FP64 residual refinement, one sweep on lines of <= 256 unknowns (a lane
 * per line) and two on longer lines (a wavefront per line); FP64 inputs and
 * outputs; lines of <= 1024 unknowns, rows cap <= 16, TPF_ERR_VALUE beyond;
 * no counterpart in the reference). */
TPF_API tpf_status tpf_ccd_plan_create(tpf_ccd_plan** plan, int ndim,
                               const int64_t* n_nodes, const double* h,
                               const int* bc, int precision);
TPF_API tpf_status tpf_ccd_plan_destroy(tpf_ccd_plan* plan);

/* Replaces CCDSolver.differentiate(data, axis) -> (d1, d2)
 * (ccd/ccd_solver.py:134-172; ccd_solver_helpers.py:38-77 wall,
 * :426-465 periodic).  Either output may be NULL (first_derivative /
 * second_derivative, ccd_solver.py:205-237).  f, d1, d2 have the full
 * field shape; d1/d2 must not alias f. */
TPF_API tpf_status tpf_ccd_differentiate(tpf_ccd_plan* plan, const double* f,
                                 double* d1, double* d2, int axis,
                                 void* stream);

/* All axes whose bit is set in axis_mask, in increasing axis order:
 * d1[k*n_total ...] and d2[k*n_total ...] hold the k-th selected axis. */
TPF_API tpf_status tpf_ccd_apply(tpf_ccd_plan* plan, const double* f, double* d1,
                         double* d2, uint32_t axis_mask, void* stream);

/* Replaces the per-axis loops that differentiate a different field along
 * each axis, `fp_a, _ = ccd.differentiate(f_a, axis=a)`
 * (levelset/advection_dccd.py:119-121, levelset/reinit_ops.py:86-89):
 * every axis a whose bit is set in axis_mask differentiates f[a] along a
 * into d1[a] / d2[a], all in one launch when the axes share a launch shape
 * (the axes of tpf_ccd_apply's grid, each with its own input).  f, d1, d2
 * are arrays of ndim pointers indexed by axis; d1 or d2 may be NULL, and so
 * may their entries.  No output may alias any input or another output. */
TPF_API tpf_status tpf_ccd_apply_fields(tpf_ccd_plan* plan, const double* const* f,
                                        double* const* d1, double* const* d2,
                                        uint32_t axis_mask, void* stream);

/* Launch geometry / reduced-system facts of one axis (for tests, bench and
 * the autotuner): partitions, rows per partition, coupling window, and the
 * largest dropped reduced-system coefficient (0 when untruncated). */
TPF_API tpf_status tpf_ccd_axis_info(tpf_ccd_plan* plan, int axis, int* partitions,
                             int* max_rows, int* window, double* trunc_coeff);

/* Kernel a single-axis call on `axis` runs (no counterpart in the
 * reference): 1 lane per line with block recurrences, 2 wavefront per
 * line, 3 lane per line with the explicit partition inverse (the values of
 * tpf_ccd_plan_set_launch's mode). */
TPF_API tpf_status tpf_ccd_axis_kernel(tpf_ccd_plan* plan, int axis, int* kernel);

/* Launch shape of one axis (no counterpart in the reference: the knobs the
 * autotuner in twophaseflow_amd/autotune.py sets).  mode 0: automatic (a
 * wavefront per line for < 4096 lines of > 256 unknowns, and for >= 4096
 * contiguous lines of 33-256 unknowns, several lines per wavefront; else a
 * lane per line, with the explicit partition inverse on strided FP64 axes
 * of <= 16 x 16 rows); 1: a lane per line, block recurrences (<= 16
 * partitions, rows cap 8/16/32/64); 2: a wavefront per line (<= 64
 * partitions, rows cap 8/16); 3: a lane per line, explicit partition
 * inverse (strided axes, <= 16 partitions, rows cap 8/16).  An
 * explicit-inverse axis fused in one tpf_ccd_apply launch with a
 * recurrence axis runs the recurrence kernel.  rows_cap 0 picks the default
 * for the mode.  Validated on the host; the axis' tables are rebuilt on
 * next use. */
TPF_API tpf_status tpf_ccd_plan_set_launch(tpf_ccd_plan* plan, int axis, int mode,
                                   int rows_cap);

/* Store chunk of the explicit-inverse kernel on a strided axis (no
 * counterpart in the reference; a launch knob like tpf_ccd_plan_set_launch).
 * align 8 or 16: each workgroup writes whole `align`-element aligned chunks
 * of the outputs, solving `align` lines of the neighbouring group again
 * (rows of 2^k + 1 node axes start at 8 B offsets, and chunks written by
 * two workgroups are slow); 0: every workgroup writes its own 64 lines;
 * -1: the default (8).  Results do not depend on it.  A fused launch takes
 * the first axis' value. */
TPF_API tpf_status tpf_ccd_plan_set_store_align(tpf_ccd_plan* plan, int axis, int align);

/* Grid metric of a non-uniform axis: replaces apply_ccd_metric
 * (ccd/ccd_solver_helpers.py:509-519) as CCDSolver applies it on grids that
 * are not uniform (ccd_solver.py:126-130, 169-170).  J[j] = dxi/dx and
 * dJ[j] = dJ/dxi at the axis' n_nodes[axis] nodes (host arrays, copied);
 * the plan's h[axis] must then be the computational spacing 1/N.  Every
 * later call of the plan returns d1 = J d1_xi and d2 = J^2 d2_xi +
 * J dJ d1_xi on that axis (a d2-only call solves d1_xi into a plan-owned
 * scratch array, allocated on first use; such calls of one plan on two
 * streams race).  J = dJ = NULL removes the metric.  Batches reject plans
 * with a metric. */
TPF_API tpf_status tpf_ccd_plan_set_metric(tpf_ccd_plan* plan, int axis, const double* J,
                                           const double* dJ);

/* Batches of independent CCD problems (no counterpart in the reference; the
 * batching extension).  A batch fixes `count` plans and the axes to
 * differentiate per plan; every (problem, axis) pair with the same launch
 * shape (rows cap, partitions) goes into one grid whose workgroups look up
 * their problem in a device table.  apply() takes per-problem device arrays
 * f[j], d1[j], d2[j] (outputs stacked per selected axis as tpf_ccd_apply);
 * asynchronous on `stream`; calls on one batch must share one stream. */
typedef struct tpf_ccd_batch tpf_ccd_batch;
TPF_API tpf_status tpf_ccd_batch_create(tpf_ccd_batch** batch, tpf_ccd_plan* const* plans,
                                const uint32_t* axis_masks, int count);
TPF_API tpf_status tpf_ccd_batch_apply(tpf_ccd_batch* batch, const double* const* f,
                               double* const* d1, double* const* d2, void* stream);
TPF_API tpf_status tpf_ccd_batch_destroy(tpf_ccd_batch* batch);

/* ---------------------------------------------------------------- PPE --- */
typedef struct tpf_ppe tpf_ppe;

typedef struct {
  double tol;          /* ||r||/||b|| target (SolverConfig.pseudo_tol)     */
  int64_t maxiter;     /* iteration cap (SolverConfig.pseudo_maxiter)      */
  int precond;         /* 0 none, 1 Jacobi (fvm_matrixfree.py:106-114)     */
  int check_every;     /* host convergence check period in iterations      */
  int use_x0;          /* 1: p holds the initial guess (p_init)            */
  int mixed;           /* 1: FP32 matvec/inner solve + FP64 refinement     */
  double inner_tol;    /* mixed: relative reduction per FP32 inner solve (0 = 1e-4) */
} tpf_solve_opts;

typedef struct {
  int64_t iterations;  /* Krylov iterations performed                        */
  double relres;       /* recursively updated ||r||/||b|| at exit           */
  double true_relres;  /* FP64 ||b - A p||/||b|| recomputed on the CG system */
  int converged;       /* 1 when true_relres <= tol                         */
  int refinements;     /* outer FP64 refinement sweeps (mixed mode)         */
} tpf_solve_report;

/* Replaces PPESolverFVMMatrixFree.__init__ (ppe/fvm_matrixfree.py:66-138)
 * for wall BC: L[a] domain length, pin_dof = centre node
 * (core/boundary.py:181-191). Allocates the solver workspace. */
TPF_API tpf_status tpf_ppe_create(tpf_ppe** ppe, int ndim, const int64_t* n_nodes,
                          const double* L, int64_t pin_dof, void* stream);
TPF_API tpf_status tpf_ppe_destroy(tpf_ppe* ppe);

/* Replaces prepare_operator + _prepare_preconditioner_state
 * (fvm_matrixfree.py:262-279, 352-363): keeps rho, builds the Jacobi
 * diagonal of the CG system. rho must stay valid until the next call. */
TPF_API tpf_status tpf_ppe_set_density(tpf_ppe* ppe, const double* rho);

/* Replaces PPESolverFVMMatrixFree.apply (fvm_matrixfree.py:281-300):
 * out = L(rho) p with out[pin] = p[pin]. */
TPF_API tpf_status tpf_ppe_apply(tpf_ppe* ppe, const double* p, double* out);

/* Replaces PPESolverFVMMatrixFree.solve in CG form
 * (fvm_matrixfree.py:150-260): solves (-W L) p = -W rhs with the pin
 * row/column = identity; on exit p[pin] = 0.  Synchronous.  Returns
 * TPF_ERR_RUNTIME when maxiter is reached above tol (report still filled). */
TPF_API tpf_status tpf_ppe_solve(tpf_ppe* ppe, const double* rhs, double* p,
                         const tpf_solve_opts* opts, tpf_solve_report* rep);

typedef struct {
  double tol;          /* rtol, atol = 0 (gmres_helpers.py:16-42)                  */
  int64_t maxiter;     /* restart cycles, SciPy semantics (pseudo_maxiter)         */
  int restart;         /* Krylov dimension per cycle (ppe_restart; host default
                          min(80, max(20, maxiter)), fvm_matrixfree.py:104-105)   */
  int precond;         /* 0 none, 1 Jacobi, 2 shifted line (line_pcr)              */
  int pcr_stages;      /* line solves: >0 PCR stages capped at ceil(log2 n)
                          (GPU backend, ppe_pcr_stages), 0 exact Thomas (CPU)     */
  int use_x0;          /* 1: p holds the initial guess (p_init)                    */
  double c_tau;        /* pseudo_c_tau of the line shift 2/(c_tau rho h_min^2)    */
} tpf_gmres_opts;

/* Replaces PPESolverFVMMatrixFree.solve in its default GMRES form
 * (fvm_matrixfree.py:150-260, 302-323; fvm_matrixfree_helpers.py:76-101):
 * left-preconditioned restarted GMRES on L p = rhs with rhs[pin] = 0 and the
 * pin row = identity, SciPy 1.15 stopping rules (true ||rhs - L p|| <= tol
 * ||rhs||).  On exit p[pin] = 0.  Synchronous.  rep->iterations counts
 * Arnoldi steps, rep->refinements restart cycles.  Returns TPF_ERR_RUNTIME
 * when the cycle budget ends above tol (SciPy info = maxiter; report filled). */
TPF_API tpf_status tpf_ppe_solve_gmres(tpf_ppe* ppe, const double* rhs, double* p,
                               const tpf_gmres_opts* opts, tpf_solve_report* rep);

/* z = M^-1 r of the GMRES preconditioner (precond 1 Jacobi: r / diag(L),
 * 2: per-axis shifted line solves averaged over axes; z[pin] = 0;
 * fvm_matrixfree.py:302-330).  Ordered on the handle's stream; synchronous. */
TPF_API tpf_status tpf_ppe_gmres_precond(tpf_ppe* ppe, const double* r, double* z,
                                 const tpf_gmres_opts* opts);

typedef struct {
  double tol;          /* ||r||/||b|| target on the CG system                       */
  int64_t maxiter;     /* FCG iteration cap                                         */
  int use_x0;          /* 1: p holds the initial guess                              */
  int nu;              /* red-black smoothing sweeps before and after each coarse step */
  int gamma;           /* 1: V-cycle, 2: W-cycle                                    */
  int coarse_sweeps;   /* forward+reverse sweeps on the coarsest level (<= 1024 nodes) */
  int max_levels;      /* hierarchy depth cap (0: as deep as the grid allows)       */
  int w_levels;        /* gamma repeats only on the top w_levels levels (0: all)    */
} tpf_mg_opts;

/* Solves the reference CG system (-W L) p = -W rhs (fvm_matrixfree.py:163-221)
 * with flexible CG (FCG(1)) preconditioned by one FP32 multigrid cycle of
 * the tpf_mg_* hierarchy (no counterpart in the reference; selected by
 * ppe_preconditioner="mg").  FP64 vectors, matvec and reductions; the
 * FP64 true residual judges convergence (residual replacement otherwise).
 * On exit p[pin] = 0.  Synchronous.  TPF_ERR_RUNTIME when maxiter is reached
 * above tol (report filled). */
TPF_API tpf_status tpf_ppe_solve_mg(tpf_ppe* ppe, const double* rhs, double* p,
                            const tpf_mg_opts* opts, tpf_solve_report* rep);

/* tpf_ppe_solve_mg with a learned coarse correction after every cycle (no
 * counterpart in the reference; the learned preconditioner's device path):
 * z = V(r); z += P W E^-1 W^T R (r - A z), R / P the restriction and
 * prolongation chains between level 0 and `level`, W [k][nodes(level)] FP32
 * device vectors (1 <= k <= 64), einv the k x k FP64 host matrix E^-1.
 * Synchronous; TPF_ERR_RUNTIME when maxiter is reached above tol. */
TPF_API tpf_status tpf_ppe_solve_mg_deflated(tpf_ppe* ppe, const double* rhs, double* p,
                                             const tpf_mg_opts* opts, const float* W, int k,
                                             int level, const double* einv,
                                             tpf_solve_report* rep);

/* tpf_ppe_solve_mg on `count` distinct handles at once (no counterpart in
 * the reference; replaces a loop of multigrid solves, one per simulation of
 * a batch, as in the per-simulation PPE calls of fvm_matrixfree.py:150-260).
 * Every round enqueues the next step of each unfinished problem on its own
 * handle's stream and then reads all their scalars, so the solves overlap
 * on the GPU instead of each waiting on the host once per iteration.  Each
 * problem runs the kernels of its single solve in the same order: p[i]
 * equals tpf_ppe_solve_mg(ppes[i], rhs[i], p[i], opts) bit for bit.
 * reports[count] filled.  Synchronous.  TPF_ERR_RUNTIME names the problems
 * that ended above tol (the others are solved). */
TPF_API tpf_status tpf_ppe_solve_mg_many(tpf_ppe* const* ppes, int count,
                                         const double* const* rhs, double* const* p,
                                         const tpf_mg_opts* opts, tpf_solve_report* reports);

/* FP64 ||b - A p||_2 / ||b||_2 on the CG system (the judge of every solve). */
TPF_API tpf_status tpf_ppe_true_relres(tpf_ppe* ppe, const double* rhs,
                               const double* p, double* relres);

/* Batched Jacobi-PCG over independent PPE handles (uneven grids allowed):
 * one launch per PCG step covers every problem, iteration chunks replay as
 * one captured graph, and every problem keeps its own stop test and restart
 * rule.  Each problem's arithmetic and reduction order are those of
 * tpf_ppe_solve, so p equals its single solve bit for bit.  Replaces a loop
 * of PPESolverFVMMatrixFree.solve calls in CG form (fvm_matrixfree.py:
 * 150-260), one per simulation of a batch; the reference steps such
 * simulations one after another.  The handles' workspaces are used, so a
 * handle must not be solved elsewhere during a batched solve; densities are
 * those of the last tpf_ppe_set_density.  Calls are ordered on `stream`. */
typedef struct tpf_ppe_batch tpf_ppe_batch;
TPF_API tpf_status tpf_ppe_batch_create(tpf_ppe_batch** batch, tpf_ppe* const* ppes, int count,
                                        void* stream);
/* rhs[i], p[i]: problem i's right-hand side and solution (initial guess
 * when opts->use_x0); both NULL skip problem i this call (its report is
 * zeroed).  opts: precond 1, mixed 0.  reports[count] filled.
 * Synchronous.  TPF_ERR_RUNTIME names the problems that ended above tol
 * (their reports are filled; the others are solved). */
TPF_API tpf_status tpf_ppe_batch_solve(tpf_ppe_batch* batch, const double* const* rhs,
                                       double* const* p, const tpf_solve_opts* opts,
                                       tpf_solve_report* reports);
TPF_API tpf_status tpf_ppe_batch_destroy(tpf_ppe_batch* batch);

/* ----------------------------------------------------------- FCCD PPE --- */
/* The FCCD matrix-free PPE, L p = D_f[alpha_f G_f p] with the FCCD face
 * gradient G_f p = (p_hi - p_lo)/H - (H/24)(q_hi - q_lo), q the nodal CCD
 * second derivative, full_face boundary space, legacy operator pairing,
 * uniform grids; the coefficient schemes and interface couplings below.
 * Calls are ordered on the stream given to tpf_fccd_ppe_create. */
typedef struct tpf_fccd_ppe tpf_fccd_ppe;

/* ppe_coefficient_scheme / ppe_interface_coupling_scheme
 * (ppe/fccd_matrixfree.py:141-186). */
enum { TPF_FCCD_PHASE_DENSITY = 0, TPF_FCCD_PHASE_SEPARATED = 1 };
enum { TPF_FCCD_COUPLING_NONE = 0, TPF_FCCD_JUMP_DECOMPOSITION = 1, TPF_FCCD_AFFINE_JUMP = 2 };

/* Replaces PPESolverFCCDMatrixFree.__init__ with its FCCDSolver / CCDSolver
 * (ppe/fccd_matrixfree.py:108-213, ccd/fccd.py:165-195): bc[a] in
 * {TPF_BC_WALL, TPF_BC_PERIODIC}, L[a] domain length (H = L/N), pin_dof =
 * the gauge node (centre node by default, fccd_matrixfree.py:187-192). */
TPF_API tpf_status tpf_fccd_ppe_create(tpf_fccd_ppe** ppe, int ndim, const int64_t* n_nodes,
                                       const double* L, const int* bc, int64_t pin_dof,
                                       void* stream);
TPF_API tpf_status tpf_fccd_ppe_destroy(tpf_fccd_ppe* ppe);

/* Replaces prepare_operator (ppe/fccd_matrixfree_lifecycle.py:44-141): keeps
 * rho and builds the Jacobi inverse (fccd_matrixfree_helpers.py:251-264; no
 * pinned row under the mean gauge).  phase_density: the face coefficients
 * 2/(rho_lo + rho_hi) are formed inside each matvec.  phase_separated: the
 * face coefficients are built here (helpers.py:207-248) -- zero across the
 * density threshold 0.5 (rho_min + rho_max), with the per-phase
 * control-volume mean gauge (helpers.py:56-160; a uniform density keeps the
 * pin and 2/(rho_lo + rho_hi)), or with affine_jump the cut-face resistance
 * 1/(theta rho_lo + (1 - theta) rho_hi) where psi crosses the phase threshold
 * (coupling/interface_stress_closure.py:225-277) and the pin; affine_jump
 * without interface data is TPF_ERR_RUNTIME.  rho (and the interface arrays)
 * must stay valid until the next call.  Synchronous for phase_separated. */
TPF_API tpf_status tpf_fccd_ppe_set_density(tpf_fccd_ppe* ppe, const double* rho);

/* Selects the coefficient scheme and interface coupling (TPF_FCCD_*;
 * validation of fccd_matrixfree.py:165-183: phase_density takes coupling
 * none).  Rebuilds the prepared state when a density is set. */
TPF_API tpf_status tpf_fccd_ppe_set_scheme(tpf_fccd_ppe* ppe, int coefficient_scheme,
                                           int interface_coupling);

/* Replaces set_interface_jump_context / clear_interface_jump_context
 * (fccd_matrixfree.py:230-304) for the nodal_cut_face route: psi (NULL
 * clears) with kappa = kappa_lg (Young-Laplace j = p_gas - p_liquid =
 * -sigma kappa: the affine drive takes kappa at the cut point, the jump
 * decomposition J = j (1 - psi) when sigma > 0), or, with kappa NULL, a nodal
 * jump j (face-averaged on cut faces; the jump decomposition stays off, as
 * the reference's sigma = 0 context).  phase_threshold: the psi level of the
 * interface (0.5).  Device arrays of the field shape, kept by pointer. */
TPF_API tpf_status tpf_fccd_ppe_set_interface(tpf_fccd_ppe* ppe, const double* psi,
                                              const double* kappa, const double* jump,
                                              double sigma, double phase_threshold);

/* The right-hand side the solve iterates on (fccd_matrixfree.py:506-517):
 * rhs - L J (jump decomposition), + D_f[alpha_f B_Gamma(j)] (affine_jump,
 * fccd_matrixfree.py:462-504, interface_stress_closure.py:285-401), then the
 * per-phase mean projection or b[pin] = 0 (helpers.py:339-463), zero image
 * rows.  rhs and b distinct. */
TPF_API tpf_status tpf_fccd_ppe_system_rhs(tpf_fccd_ppe* ppe, const double* rhs, double* b);

/* The face coefficients alpha_f of `axis` in use (PPESolverFCCDMatrixFree.
 * _coeff_face), stored at each face's lower node (the last node of a line
 * holds 0); out has the field shape. */
TPF_API tpf_status tpf_fccd_ppe_face_coeff(tpf_fccd_ppe* ppe, int axis, double* out);

/* out = p + J of the jump decomposition, or p (apply_interface_jump,
 * fccd_matrixfree.py:644-654). */
TPF_API tpf_status tpf_fccd_ppe_interface_jump(tpf_fccd_ppe* ppe, const double* p, double* out);

/* Whether the per-phase mean gauge is active, and its density threshold
 * (NaN without it; PPESolverFCCDMatrixFree._phase_threshold). */
TPF_API tpf_status tpf_fccd_ppe_gauge_info(tpf_fccd_ppe* ppe, int* mean_gauge, double* threshold);

/* Replaces PPESolverFCCDMatrixFree.apply (fccd_matrixfree.py:378-397):
 * out = L p, out[pin] = p[pin], or under the per-phase mean gauge
 * out = P L P p + (I - P) p; periodic image rows out = p_img - p_src.
 * Asynchronous; p and out distinct. */
TPF_API tpf_status tpf_fccd_ppe_apply(tpf_fccd_ppe* ppe, const double* p, double* out);

/* Replaces PPESolverFCCDMatrixFree.solve (fccd_matrixfree.py:505-611): SciPy
 * 1.15 GMRES (gmres_helpers.py:16-42) on the operator of tpf_fccd_ppe_apply
 * with b = tpf_fccd_ppe_system_rhs(rhs) (for phase_density: rhs with
 * b[pin] = 0 and zero image rows); under the mean gauge the Jacobi step
 * projects the phase means and the result is mean-projected, and the jump
 * decomposition returns p + J; opts->precond 0 none, 1 Jacobi, 2 multigrid (no
 * counterpart in the reference; wall grids: one FP32 W(2,2) cycle of the FVM
 * operator's hierarchy on -W r per step, the low-order operator preconditioning
 * the FCCD one); opts->restart as SciPy's
 * (the reference passes ppe_restart, None = 20); opts->use_x0: p holds p_init
 * (images synced, pin 0).  On exit p[pin] = 0 and image nodes equal their
 * sources; rep->true_relres = ||L p - b|| / ||b|| on that field.
 * Synchronous; TPF_ERR_RUNTIME when GMRES ends above tol (report filled). */
TPF_API tpf_status tpf_fccd_ppe_solve_gmres(tpf_fccd_ppe* ppe, const double* rhs, double* p,
                                            const tpf_gmres_opts* opts, tpf_solve_report* rep);

/* FCCD face primitives of FCCDSolver along `axis` (ccd/fccd.py:165-521;
 * uniform weights ccd/fccd_helpers.py:17-35, H = L/N).  Nodal arrays have the
 * field shape; face arrays the field shape with N along `axis` (face j between
 * nodes j and j+1; periodic: face N-1 pairs node N-1 with node 0).  q is the
 * nodal CCD second derivative along `axis`.  op:
 *   0 face_gradient          (u_hi - u_lo)/H - (H/24)(q_hi - q_lo)   :243-277
 *   1 face_value             (u_lo + u_hi)/2 - (H^2/16)(q_lo + q_hi) :279-299
 *   2 face_second_derivative (q_lo + q_hi)/2                          :301-317
 *   3 node_gradient          R4 Hermite, one-sided at walls (nodal)  :392-462
 *   4 face_divergence        u = face array -> nodal; walls 0         :464-521
 *   5 upwind_face_value      Taylor state from the upwind node with d
 *                            (nodal f') and vel (face velocity)       :333-390
 * Asynchronous on `stream`; d and vel only for op 5, q unused by op 4. */
TPF_API tpf_status tpf_fccd_face_op(int op, int ndim, const int64_t* n_nodes, const double* L,
                                    const int* bc, int axis, const double* u, const double* q,
                                    const double* d, const double* vel, double* out,
                                    void* stream);

/* Replaces RhieChowInterpolator.face_velocity_divergence
 * (spatial/rhie_chow.py:40-284): the divergence of the Rhie-Chow face
 * velocities u_f = (u_L+u_R)/2 - dt 2/(rho_L+rho_R) [(p_R-p_L)/h - (dp_L+dp_R)/2]
 * (wall nodes take one-sided dp; face 0 is the periodic wrap or a zero wall
 * face), nodal (F_{k+1}-F_k)/h with node N = (F_0-F_N)/h (periodic) or -F_N/h
 * (wall).  vel: ndim device pointers (host array); dp: the CCD first
 * derivative of p, stacked [ndim][n_total] as tpf_ccd_apply writes it.
 * kappa, psi, dpsi (dpsi stacked like dp) and we != 0 switch on the
 * balanced-force bracket (:152-204); all three NULL switch it off. */
TPF_API tpf_status tpf_rhie_chow_divergence(int ndim, const int64_t* n_nodes, const double* L,
                                            const int* bc, const double* const* vel,
                                            const double* p, const double* rho, const double* dp,
                                            const double* kappa, const double* psi,
                                            const double* dpsi, double we, double dt,
                                            double* out, void* stream);

/* Dissipative-CCD level-set advection (DissipativeCCDAdvection,
 * levelset/advection_dccd.py:20-152) in three calls per TVD-RK3 stage
 * (time_integration/tvd_rk3.py:26-48); `periodic[a]` = the level-set
 * boundary of axis a is periodic (else Neumann / wall: end nodes unfiltered).
 * tpf_dccd_flux: f[a] = q * vel[a], f stacked [ndim][n_total]; vel is a host
 * array of ndim device pointers. */
TPF_API tpf_status tpf_dccd_flux(int ndim, const int64_t* n_nodes, const double* q,
                                 const double* const* vel, double* f, void* stream);
/* out = clip(ca q0 + cb (q + dt L(q)), lo, hi) (clip = 0: no clip), with
 * L(q) = -sum_a [fp_a + eps_d (2q-1)^2 (fp_a(+1) - 2 fp_a + fp_a(-1))]
 * (levelset/advection_kernels.py:31-34, periodic ghosts core/boundary.py:
 * 273-276), fp = the CCD first derivatives of f stacked like f; periodic
 * image nodes take their roots' values (core/boundary.py:116-133).  out must
 * not alias q or q0. */
TPF_API tpf_status tpf_dccd_stage(int ndim, const int64_t* n_nodes, const int* periodic,
                                  const double* q0, const double* q, const double* fp,
                                  double eps_d, double dt, double ca, double cb, int clip,
                                  double lo, double hi, double* out, void* stream);
/* out = in with every periodic image node copied from its root (in == out
 * allowed; core/boundary.py:116-133). */
TPF_API tpf_status tpf_sync_periodic_images(int ndim, const int64_t* n_nodes,
                                            const int* periodic, const double* in, double* out,
                                            void* stream);
/* Interface-weighted mass correction to m_target (levelset/heaviside.py:
 * 118-149) with Grid.cell_volumes weights, then the image copies; synchronous
 * (two FP64 reductions).  mass_before (may be NULL) receives sum q dV. */
TPF_API tpf_status tpf_dccd_mass_correct(int ndim, const int64_t* n_nodes, const double* L,
                                         const int* periodic, const double* q, double m_target,
                                         double* out, double* mass_before, void* stream);
/* tpf_dccd_mass_correct without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (code under stream capture, tight stepping loops, harnesses that forbid
 * hipMalloc).  Its partial sums and correction factor live in a fixed
 * device array of 2 * 2048 + 8 doubles (33 KB, one per device, kept until
 * the process ends); the factor is formed on the device.  With mass_before
 * NULL the call is asynchronous on `stream` and synchronises nothing; with
 * mass_before set it copies sum q dV back and synchronises the stream.
 * Every call on one device shares the array, so calls that overlap on a
 * device (other streams, other threads) race; nothing in it needs clearing
 * between calls. */
TPF_API tpf_status tpf_dccd_mass_correct_noalloc(int ndim, const int64_t* n_nodes,
                                                 const double* L, const int* periodic,
                                                 const double* q, double m_target, double* out,
                                                 double* mass_before, void* stream);

/* ------------------------------------------------- legacy time step --- */
/* Material properties (levelset/heaviside.py:167-187): rho = rho_g +
 * (rho_l - rho_g) psi, mu likewise; phi (may be NULL) = the inverse
 * Heaviside eps ln(c / (1 - c)), c = clip(psi, 1e-6, 1 - 1e-6)
 * (heaviside.py:90-118), which update_legacy_curvature stores as sim.phi
 * (simulation/legacy_flow_helpers.py:83-86).  Asynchronous. */
TPF_API tpf_status tpf_ns_properties(int64_t n, const double* psi, double rho_l, double rho_g,
                                     double mu_l, double mu_g, double eps, double* rho,
                                     double* mu, double* phi, void* stream);
/* UCCD6ConvectionTerm.compute (ns_terms/uccd6_convection.py:88-128) on the
 * plan's grid: out[j] = sum_k [-(u_k D1_k u_j + D1_k(u_k u_j))/2
 * - sigma h_k^7 max|u_k| D2_k^4 u_j], h_k = L[k]/(n_nodes[k]-1), the
 * maxima kept on the device.  vel, out: ndim device arrays; work: at least
 * tpf_ns_uccd6_work(ndim, n_nodes) doubles.  Asynchronous. */
TPF_API int64_t tpf_ns_uccd6_work(int ndim, const int64_t* n_nodes);
TPF_API tpf_status tpf_ns_uccd6_convection(tpf_ccd_plan* plan, int ndim, const int64_t* n_nodes,
                                           const double* L, const double* const* vel,
                                           double sigma, double* const* out, double* work,
                                           void* stream);
/* The viscous body of ViscousTerm with the default ccd_bulk scheme
 * (ns_terms/viscous_spatial.py:39-73, 93-130, 156-169): out[a] = sum_b
 * Dlow_b[mu (D1_b u_a + D1_a u_b)] / (Re rho), Dlow first-order one-sided at
 * both ends of each axis and three-point inside.  work: at least
 * tpf_ns_viscous_work(ndim, n_nodes) doubles.  Asynchronous. */
TPF_API int64_t tpf_ns_viscous_work(int ndim, const int64_t* n_nodes);
TPF_API tpf_status tpf_ns_viscous(tpf_ccd_plan* plan, int ndim, const int64_t* n_nodes,
                                  const double* L, const double* const* vel, const double* mu,
                                  const double* rho, double Re, double* const* out, double* work,
                                  void* stream);
/* Predictor.compute's explicit right-hand side (time_integration/
 * ab2_predictor.py:150-190): rhs[c] = rho ab2_c - [c = last] rho/Fr^2 +
 * kappa dpsi_c / We - dp_c, ab2 = conv (conv_prev NULL: Euler start-up) or
 * 3/2 conv - 1/2 conv_prev; dp_c (the CCD gradient of p, stacked by axis)
 * is taken as zero on the wall planes normal to c (coupling/
 * velocity_corrector.py:39-50).  Also writes the Euler point of the Picard
 * CN advance u_pred = u + dt (rhs / rho + visc_n) (time_integration/
 * cn_advance/picard_cn.py:66-84).  Asynchronous. */
TPF_API tpf_status tpf_ns_predictor_rhs(int ndim, const int64_t* n_nodes, const int* periodic,
                                        const double* const* vel, const double* const* conv,
                                        const double* const* conv_prev, const double* rho,
                                        const double* kappa, const double* dpsi, const double* dp,
                                        double Fr, double We, double dt,
                                        const double* const* visc_n, double* const* rhs,
                                        double* const* u_pred, void* stream);
/* The Picard CN trapezoid u* = u + dt (rhs / rho + visc_n / 2 + visc_pred / 2)
 * (picard_cn.py:86-96); periodic is accepted for symmetry with
 * tpf_ns_predictor_rhs and unused.  Asynchronous. */
TPF_API tpf_status tpf_ns_predictor_final(int ndim, const int64_t* n_nodes, const int* periodic,
                                          const double* const* vel, const double* const* rhs,
                                          const double* rho, const double* const* visc_n,
                                          const double* const* visc_pred, double dt,
                                          double* const* u_star, void* stream);
/* Unit normal nhat[a] = g_a / max(sqrt(max(|g|^2, 1e-28)), 1e-14) of a
 * gradient stacked by axis (ClosestPointExtender.compute_normal,
 * levelset/closest_point_extender.py:112-131).  Asynchronous. */
TPF_API tpf_status tpf_ns_normal(int ndim, int64_t n, const double* grad, double* const* nhat,
                                 void* stream);
/* BoundaryConditionHandler.apply (simulation/boundary_condition.py:36-87)
 * on ncomp velocity components in place: wall axes zero both planes,
 * periodic axes copy the root plane into the image plane.  Asynchronous. */
TPF_API tpf_status tpf_ns_wall_bc(int ndim, const int64_t* n_nodes, const int* periodic,
                                  int ncomp, double* const* comps, void* stream);
/* Host scalars for CFLCalculator.compute (time_integration/cfl.py:85-137):
 * max |a_k| of nk <= 3 arrays, or max (a / b); synchronous.  work: at least
 * 3 * 1024 + 8 doubles of device memory. */
TPF_API tpf_status tpf_ns_absmax(int nk, const double* const* arrays, int64_t n, double* work,
                                 double* out_host, void* stream);
TPF_API tpf_status tpf_ns_ratio_max(const double* a, const double* b, int64_t n, double* work,
                                    double* out_host, void* stream);
/* Number of NaN / Inf values over nk <= 64 arrays of n doubles, for the run
 * loop's rollback test (simulation/legacy_run_loop.py:70-82); synchronous.
 * work: at least 8 + nk doubles of device memory. */
TPF_API tpf_status tpf_ns_nonfinite(int nk, const double* const* arrays, int64_t n, double* work,
                                    int64_t* count_host, void* stream);
/* out = x / s (the PPE right-hand side div_rc / dt, simulation/
 * legacy_flow_helpers.py:52) and out = x + y (p^{n+1} = p^n + dp, :57);
 * out may alias x.  Asynchronous. */
TPF_API tpf_status tpf_ns_div_scalar(int64_t n, const double* x, double s, double* out,
                                     void* stream);
TPF_API tpf_status tpf_ns_add(int64_t n, const double* x, const double* y, double* out,
                              void* stream);
/* ClosestPointExtender.extend in 2-D (levelset/closest_point_extender.py:
 * 136-228): nodes with phi >= 0 take the tensor-product quintic Hermite
 * interpolant of q at the closest point (x - phi nx, y - phi ny), clipped to
 * the domain; its cell is found by binary search on the node coordinates x
 * (n_nodes[0]) and y (n_nodes[1]) and shifted at most three times per axis
 * until its far node has phi < 0.  The derivatives q_x ... q_xxyy come from
 * the plan (four CCD solves).  work: at least 8 n doubles.  Asynchronous. */
TPF_API tpf_status tpf_ns_extend_2d(tpf_ccd_plan* plan, const int64_t* n_nodes, const double* x,
                                    const double* y, const double* q, const double* phi,
                                    const double* nx, const double* ny, double* out,
                                    double* work, void* stream);

/* Curvature straight from psi (CurvatureCalculatorPsi,
 * levelset/curvature_psi.py:45-196).  tpf_dccd_filter_nd: the DCCD filter of
 * levelset/curvature.py:60-128 on f (successive 3-point passes with
 * coefficient eps (2 psi - 1)^2 along every axis, periodic axes over the
 * unique nodes, then the wall boundary planes restored from f); out and
 * scratch are distinct n_total arrays. */
TPF_API tpf_status tpf_dccd_filter_nd(int ndim, const int64_t* n_nodes, const int* periodic,
                                      const double* psi, double eps, const double* f,
                                      double* out, double* scratch, void* stream);

/* tpf_dccd_filter_nd over nf <= 8 fields that share psi (the derivative
 * loop of CurvatureCalculatorPsi, levelset/curvature_psi.py:89-104, filters
 * every first and second derivative with the same psi): one launch per pass
 * for all fields, each field bit-identical to its own tpf_dccd_filter_nd.
 * f, out, scratch: arrays of nf device pointers, all 3 nf distinct. */
TPF_API tpf_status tpf_dccd_filter_nd_multi(int ndim, const int64_t* n_nodes,
                                            const int* periodic, const double* psi, double eps,
                                            int nf, const double* const* f, double* const* out,
                                            double* const* scratch, void* stream);
/* kappa from the stacked CCD derivatives d1, d2 [ndim][n_total] and the mixed
 * derivatives (2-D: [psi_xy]; 3-D: [xy, xz, yz]); 0 where psi <= psi_min or
 * psi >= 1 - psi_min.  2-D or 3-D. */
TPF_API tpf_status tpf_curvature_psi(int ndim, const int64_t* n_nodes, const double* psi,
                                     const double* d1, const double* d2, const double* mixed,
                                     double psi_min, double* out, void* stream);
/* 3x3 Gaussian [1 2 1; 2 4 2; 1 2 1]/16 with edge replication on a 2-D field
 * (curvature_psi.py:171-196); in and out distinct. */
TPF_API tpf_status tpf_gaussian_3x3(const int64_t* n_nodes, const double* in, double* out,
                                    void* stream);

/* Operator-split CLS reinitialisation (SplitReinitializer,
 * levelset/reinit_split.py:22-102, levelset/reinit_ops.py:36-204): n_steps
 * pseudo-time steps of DCCD-filtered compression (eps_d_comp) and CN-ADI
 * diffusion, each clipped to [0, 1], then (mass_correction != 0) the
 * interface-weighted mass correction.  ccd: a plan of this grid (its own
 * wall / periodic closures give the derivatives); bc[a]: the level-set
 * boundary of axis a, TPF_REINIT_BC_* (ghosts of core/boundary.py:249-293;
 * periodic axes also synchronise their image nodes).  psi and out distinct;
 * ordered on `stream` (two short host waits for the mass sums when
 * mass_correction is set). */
enum { TPF_REINIT_BC_ZERO = 0, TPF_REINIT_BC_PERIODIC = 1, TPF_REINIT_BC_NEUMANN = 2,
       TPF_REINIT_BC_OUTFLOW = 3 };
TPF_API tpf_status tpf_reinit_split(tpf_ccd_plan* ccd, int ndim, const int64_t* n_nodes,
                                    const double* L, const int* bc, double eps, int n_steps,
                                    double eps_d_comp, int mass_correction, const double* psi,
                                    double* out, void* stream);

/* ------------------------------------------------------------- fields --- */
/* Replaces InitialConditionBuilder.build (gas spheres in liquid) +
 * update_properties (simulation/initial_conditions/builder.py:100-172,
 * levelset/heaviside.py:167-187): psi = 1/(1+exp(clip(-phi_gas/eps))),
 * rho = rho_g + (rho_l-rho_g) psi.  centres: n_spheres*ndim host doubles,
 * radii: n_spheres host doubles.  Either output may be NULL. */
TPF_API tpf_status tpf_fields_bubbles(int ndim, const int64_t* n_nodes,
                              const double* L, int n_spheres,
                              const double* centres, const double* radii,
                              double eps, double rho_l, double rho_g,
                              double* psi, double* rho, void* stream);

/* Mass integral sum(q * dV) in FP64 with Grid.cell_volumes weights
 * (core/grid.py:1123-1170), synchronous (heaviside.py:118-149). */
TPF_API tpf_status tpf_fields_mass(int ndim, const int64_t* n_nodes, const double* L,
                           const int* bc, const double* q, double* mass,
                           void* stream);

/* u_a -= dt/rho * g with g zeroed on the wall planes normal to `axis`
 * when bc[axis] is wall (coupling/velocity_corrector.py:39-89). */
TPF_API tpf_status tpf_fields_correct_velocity(int ndim, const int64_t* n_nodes,
                                       const int* bc, int axis, double dt,
                                       const double* rho, const double* grad,
                                       double* u, void* stream);

/* --------------------------------------------- learned preconditioner --- */
/* No counterpart in the reference: BF16 MFMA building blocks of the U-Net
 * preconditioner used inside flexible CG.  Activations are channels-last
 * [D][H][W][C] BF16 (raw 16-bit), C a multiple of 16; weights Wt[27][Cout][Cin]
 * BF16 with tap = (dz+1)*9 + (dy+1)*3 + (dx+1); FP32 accumulation.
 * Forward: Y = act(bias + conv3x3x3(X)), zero padding, same size; Y is BF16
 * or FP32 (out_f32).  Cout <= 128. */
TPF_API tpf_status tpf_conv3d_fwd(const void* X, const void* Wt, const float* bias, void* Y,
                                  int D, int H, int W, int Cin, int Cout, int relu,
                                  int out_f32, void* stream);
/* dW[tap][co][ci] += sum_v dY[v][co] X[v+off(tap)][ci] (FP32, accumulates).
 * With Cin, Cout in {16, 32, 64} and Cin * Cout >= 1024 it runs on the
 * 16x16x32 BF16 MFMA with workgroup partials reduced in a fixed order
 * (deterministic); the partials live in one grow-only buffer per device
 * (<= 113 MB, kept until the process ends), so calls that overlap on
 * different streams of one device race. */
TPF_API tpf_status tpf_conv3d_wgrad(const void* X, const void* dY, float* dW, int D, int H,
                                    int W, int Cin, int Cout, void* stream);
/* b[i] = bf16(a[i]), round to nearest even. */
TPF_API tpf_status tpf_f32_to_bf16(const float* a, void* b, int64_t n, void* stream);

/* CG-system matvec out = A p (FP64) with the handle's current density, the
 * same operator tpf_ppe_solve iterates on (used by flexible CG).  Both
 * system_apply calls are asynchronous on the stream given to tpf_ppe_create. */
TPF_API tpf_status tpf_ppe_system_apply(tpf_ppe* ppe, const double* p, double* out);
/* Builds the handle's FP32 copies of rho and of the Jacobi diagonal (once
 * per density; synchronous). */
TPF_API tpf_status tpf_ppe_prepare_f32(tpf_ppe* ppe);
/* out = D^-1, the Jacobi diagonal inverse of the CG system (FP64). */
TPF_API tpf_status tpf_ppe_jacobi_inv(tpf_ppe* ppe, double* out);
/* b = -W rhs with b[pin] = 0: the right-hand side of the CG system. */
TPF_API tpf_status tpf_ppe_system_rhs(tpf_ppe* ppe, const double* rhs, double* b);
/* FP32 CG-system matvec out = A p on the handle's FP32 density copy
 * (tpf_ppe_prepare_f32 first). */
TPF_API tpf_status tpf_ppe_system_apply_f32(tpf_ppe* ppe, const float* p, float* out);

/* Geometric multigrid on the CG system (FP32), the backbone of the learned
 * preconditioner; no counterpart in the reference.  Level 0 is the fine
 * grid; coarse levels halve every axis (node-centred) with homogenised face
 * conductances.  All calls below except tpf_mg_setup enqueue on the stream
 * given to tpf_ppe_create and return without synchronising. */
/* Builds (or refreshes) the hierarchy for the current density: at most
 * max_levels levels, stopping at <= 1024 nodes.  Synchronous. */
TPF_API tpf_status tpf_mg_setup(tpf_ppe* ppe, int max_levels, int* levels);
/* Node counts (3 slots, leading 1s for 2-D) and pin index of a level. */
TPF_API tpf_status tpf_mg_level_info(tpf_ppe* ppe, int level, int64_t* n3, int64_t* pin);
/* z = M(r): one multigrid cycle (gamma 1: V, 2: W; nu_pre / nu_post
 * red-black sweeps, reverse colour order after the coarse correction) down
 * to stop_level (-1: the coarsest), where a level of <= 1024 nodes gets
 * coarse_sweeps forward+reverse sweeps in LDS and a larger stop level
 * coarse_sweeps forward+reverse grid sweeps.  z[pin] = 0. */
TPF_API tpf_status tpf_mg_vcycle(tpf_ppe* ppe, const float* r, float* z, int nu_pre,
                         int nu_post, int coarse_sweeps, int stop_level, int gamma);
/* Descending half of the V-cycle from r to `level`: bc = that level's
 * restricted residual right-hand side. */
TPF_API tpf_status tpf_mg_down(tpf_ppe* ppe, const float* r, int level, int nu_pre,
                       float* bc);
/* Ascending half: xc is the correction on `level`; z = fine result. */
TPF_API tpf_status tpf_mg_up(tpf_ppe* ppe, int level, const float* xc, int nu_post, float* z);
/* z = P x_c: trilinear interpolation from `level` to the fine grid (the
 * fine and intermediate pins stay 0). */
TPF_API tpf_status tpf_mg_prolong(tpf_ppe* ppe, int level, const float* xc, float* z);
/* bc = R r: full-weighting restriction from the fine grid to `level`
 * (pins of the coarser levels set to 0). */
TPF_API tpf_status tpf_mg_restrict(tpf_ppe* ppe, const float* r, int level, float* bc);
/* y = A_level x (y[pin] = 0). */
TPF_API tpf_status tpf_mg_apply(tpf_ppe* ppe, int level, const float* x, float* y);
/* `sweeps` forward then `sweeps` reverse red-black sweeps on A_level x = b. */
TPF_API tpf_status tpf_mg_smooth(tpf_ppe* ppe, int level, const float* b, float* x,
                         int sweeps);
/* out[0..3n) = face conductances of the 3 slots, out[3n..4n) = 1/diag. */
TPF_API tpf_status tpf_mg_coeffs(tpf_ppe* ppe, int level, float* out);

#ifdef __cplusplus
}
#endif
#endif /* TPF_H */
