/*
 * regot_hip.h -- C-ABI of the MI355X (gfx950) EROT hot path.
 *
 * One shared library, libregot_hip.so, holds the HIP kernels and the host
 * C++ solver drivers.  Everything crosses this boundary as plain pointers,
 * sizes and int status codes (0 = ok, < 0 = error, message in
 * regot_last_error()).  No torch or Eigen type appears here.
 *
 * Reference: yixuan/regot-python v0.0.3.  Each entry point names the
 * reference interface it replaces (file:line).  Vectors passed as
 * `const double*` / `double*` are HOST pointers unless the name says d_.
 *
 * Threading: a problem handle is not thread-safe; distinct handles may be
 * used from different threads.  Calls block until their results are ready.
 */
#ifndef REGOT_HIP_H
#define REGOT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define REGOT_F32 0
#define REGOT_F64 1

/* Solver ids for regot_solve(). */
#define REGOT_SOLVER_BCD 0
#define REGOT_SOLVER_SSNS 1
#define REGOT_SOLVER_SPARSE_NEWTON 2
#define REGOT_SOLVER_SPLR 3
#define REGOT_SOLVER_NEWTON 4
#define REGOT_SOLVER_LBFGS_DUAL 5
#define REGOT_SOLVER_APDAGD 6

typedef struct regot_problem regot_problem;
typedef struct regot_result regot_result;

/* Thread-local message of the last failed call ("" if none). */
const char* regot_last_error(void);

/* Library version, e.g. 100 for 0.1.0. */
int regot_version(void);

/* Number of visible GPUs (0 when the HIP runtime finds none). */
int regot_device_count(void);

/* y[i] = 2^x[i] by the device FP64 exponential of the one-pass objective /
 * gradient and APDAGD kernels (host arrays; no reference counterpart: a
 * self-test of that routine against a correctly rounded exp2). */
int regot_check_exp2(const double* x, double* y, long long count, int device);

/* y[i] = ln x[i] by the device FP64 logarithm of the APDAGD plan pass (host
 * arrays; no reference counterpart: a self-test of that routine against a
 * correctly rounded log). */
int regot_check_log(const double* x, double* y, long long count, int device);

/*
 * Problem handle: M (n x m, row-major, row pitch ldm elements) resident on
 * one GPU in the bulk precision, plus a, b, log a, log b and reg.
 * Replaces the Sinkhorn::Problem constructor, src/sinkhorn_problem.h:36-44
 * (and the pybind11 numpy -> Eigen conversion, src/solvers.cpp:359-373).
 *
 *   M_dtype   REGOT_F32 or REGOT_F64: element type of the caller's M.
 *   bulk      REGOT_F32: M held in FP32, exp/sum in FP32, potentials FP64.
 *             REGOT_F64: M held in FP64, everything in FP64.
 *   M_device  0: M is host memory (copied once); 1: M is device memory on
 *             `device` and is borrowed (must outlive the handle) when its
 *             dtype equals bulk and rows are 16-byte aligned, else copied.
 *   ldm       row pitch in elements (>= m); 0 means m.
 *   a, b      host FP64 marginals (copied).
 */
int regot_problem_create(regot_problem** out, int n, int m, const void* M, int M_dtype,
                         int bulk, int M_device, long long ldm, const double* a,
                         const double* b, double reg, int device);
int regot_problem_destroy(regot_problem* p);
int regot_problem_dims(const regot_problem* p, int* n, int* m, int* bulk);

/* alpha_i = reg (log a_i - LSE_j((beta_j - M_ij)/reg)).
 * Replaces Problem::optimal_alpha, src/sinkhorn_problem_bcd.cpp:177-201. */
int regot_optimal_alpha(regot_problem* p, const double* beta, double* alpha);

/* beta_j = reg (log b_j - LSE_i((alpha_i - M_ij)/reg)).
 * Replaces Problem::optimal_beta, src/sinkhorn_problem_bcd.cpp:149-174. */
int regot_optimal_beta(regot_problem* p, const double* alpha, double* beta);

/* f(gamma) = reg sum(T) - alpha'a - betat'bt, gamma = [alpha; beta_0..m-2].
 * Replaces Problem::dual_obj, src/sinkhorn_problem_obj.cpp:324-342. */
int regot_dual_obj(regot_problem* p, const double* gamma, double* obj);

/* f and grad (n+m-1) at gamma.
 * Replaces Problem::dual_obj_grad, src/sinkhorn_problem_obj.cpp:363-393. */
int regot_dual_obj_grad(regot_problem* p, const double* gamma, double* obj, double* grad);

/* The same f and gradient on DEVICE buffers, allocating nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops) and the variant of
 * regot_dual_obj_grad with the same kernels and results.  M (n x m, row pitch
 * ldm elements, 0 = m; pointer and pitch 16-byte aligned) is FP32 (bulk =
 * REGOT_F32: the FP32 row and column passes) or FP64 (REGOT_F64: the one-pass
 * FP64 kernel); d_a (n), d_b (m), d_gamma (n+m-1) are FP64.  Outputs:
 * d_out[0] = f, d_out[1] = ||grad||^2, d_grad (n+m-1).  Work is enqueued on
 * `stream` (a hipStream_t; NULL = the null stream) and not waited for.  The
 * kernels' scratch is a fixed 32 MiB device array per GPU (held until the
 * process ends): sizes needing more (beyond about n = m = 16384 in FP64 bulk)
 * are rejected.  Calls on one GPU share that array and must not overlap. */
int regot_dual_obj_grad_noalloc(const void* M, int bulk, long long ldm, int n, int m,
                                const double* d_a, const double* d_b, double reg,
                                const double* d_gamma, double* d_out, double* d_grad,
                                void* stream, int device);

/* plan = exp((alpha (+) beta - M)/reg), FP64 row-major n x m, beta_{m-1} = 0.
 * Replaces SinkhornResult::get_plan, src/sinkhorn_result.h:32-50. */
int regot_get_plan(regot_problem* p, const double* gamma, double* plan);

/* Feasible rounding of plan(gamma) onto the transport polytope of (a, b),
 * FP64 row-major n x m into host memory.  Replaces approx_proj,
 * src/approx_proj.h:6-26, applied to T(gamma) as Problem::primal_val does. */
int regot_round_plan(regot_problem* p, const double* gamma, double* plan);

/* Primal objective <P, M> - reg sum P (1 - log P) of the rounded plan
 * P = approx_proj(T(gamma)).  Replaces Problem::primal_val,
 * src/sinkhorn_problem_obj.cpp:300-319. */
int regot_primal_val(regot_problem* p, const double* gamma, double* value);

/*
 * Sparsified Hessian Hs = [diag(h1) sigma; sigma' diag(h2)] at gamma (host),
 * held on the device in CSR and CSC.  The gradient at gamma is computed here
 * (h1, h2 are the exact row / column sums / reg).
 *
 * Threshold rule.  Replaces Problem::dual_sparsified_hess,
 * src/sinkhorn_problem_obj.cpp:432-443 -> Hessian::compute_hess,
 * src/sinkhorn_hess.cpp:15-33 -> sparsify_mat3, src/sinkhorn_sparsify3.cpp:276-405.
 * *nnz counts kept entries as the reference does (exact FP64 zeros included). */
int regot_sparsified_hess(regot_problem* p, const double* gamma, double delta,
                          double density_hint, long long* nnz);
/* Top-k rule, row 0 and column 0 preserved.  Replaces
 * Problem::dual_sparsified_hess_with_density, src/sinkhorn_problem_obj.cpp:446-457 ->
 * sparsify_mat_with_density, src/sinkhorn_sparsify4.cpp:77-174. */
int regot_sparsified_hess_with_density(regot_problem* p, const double* gamma, double density,
                                       long long* nnz);
/* Stored entries, FP64-zero kept entries and Hessian::density(),
 * src/sinkhorn_hess.h:75-80. */
int regot_hess_info(regot_problem* p, long long* nnz_stored, long long* nnz_zero,
                    double* density);
/* How the handle holds the sparsified Hessian (no reference counterpart; the
 * reference always stores it, src/sinkhorn_hess.h:22-60):
 *   -1 (default)  the solver drivers keep a dense kept set (threshold rule,
 *                 or every entry kept) as a bitmask and recompute sigma from
 *                 M in every product (FP32 bulk, m - 1 <= 16384);
 *                 regot_sparsified_hess* store CSR + CSC;
 *    0            always CSR + CSC;
 *    1            matrix-free wherever the bulk and shape allow, including
 *                 regot_sparsified_hess* (regot_hess_export then fails).
 * Products, CG and the drivers' iterates agree with the stored form to
 * rounding (the sums run in another order). */
int regot_set_hess_storage(regot_problem* p, int mode);
/* Copy h1 (n), h2 (m-1), CSC (colptr m, rowind nnz, cval nnz) and CSR
 * (rowptr n+1, colind nnz, rval nnz) to host; NULL skips an array.
 * (Hessian::to_spmat, src/sinkhorn_hess.cpp:57-96, builds the same matrix.) */
int regot_hess_export(regot_problem* p, double* h1, double* h2, int* colptr, int* rowind,
                      double* cval, int* rowptr, int* colind, double* rval);
/* res = Delta x, x and res of length m-1.  Replaces Hessian::apply_Deltax,
 * src/sinkhorn_hess.cpp:136-148. */
int regot_hess_apply_delta(regot_problem* p, const double* x, double shift, double* res);
/* res = Hs x, length n+m-1.  Replaces Hessian::apply_Hsx, src/sinkhorn_hess.cpp:151-164. */
int regot_hess_apply_hs(regot_problem* p, const double* x, double* res);
/* res = (Hs + shift I)^{-1} rhs by block elimination and CG on Delta (FP64
 * dots), relative residual tol, max_iter <= 0 -> 2m.  Replaces hess_cg,
 * src/sinkhorn_cg.cpp:9-50. */
int regot_hess_cg(regot_problem* p, const double* rhs, double shift, double tol, int max_iter,
                  double* res, int* iters, double* tol_error);
/* Two right-hand sides of the same (Hs + shift I) system, their CGs run in
 * the same kernel launches (the reference's SPLR solves its two systems with
 * one factorisation, src/sinkhorn_linsolve.cpp:160-225); each result, count
 * and residual equals regot_hess_cg's for that right-hand side bit for bit.
 * iters[2] and tol_error[2] (nullable). */
int regot_hess_cg2(regot_problem* p, const double* rhs0, const double* rhs1, double shift,
                   double tol, int max_iter, double* res0, double* res1, int* iters,
                   double* tol_error);

/*
 * Solver options; zero-initialise, then call regot_opts_default().
 * Mirrors SinkhornSolverOpts, src/sinkhorn_solvers.h:14-38, and the kwargs
 * parsed by parse_sinkhorn_opts, src/solvers.cpp:314-341.
 */
typedef struct {
    int method;          /* linear solver; this build always runs CG on the GPU */
    double mu0;          /* SSNS */
    double shift;        /* sparse Newton / SPLR */
    double density;      /* sparse Newton / SPLR */
    double cg_tol;       /* <= 0: driver default */
    int cg_max_iter;     /* <= 0: driver default */
    const double* x0;    /* optional warm start (n+m-1), host; NULL = reference init */
    int bcd_batch;       /* BCD iterations enqueued per batch (two in flight); <= 0: 8 */
    /* BCD iteration form.  0: two passes over M per iteration (row pass,
     * column pass); nonzero (default 1): one pass per iteration when M is
     * FP32, 1024 < m <= 16384, at most 1024 rows fall to each compute unit
     * and min a min b >= 2^-60 sum a sum b; the two-pass form otherwise. */
    int bcd_fused;
    double tau;          /* QROT gradient descent / L-BFGS: extra term tau/2 (1'alpha - 1'beta)^2
                            (QROTSolverOpts::tau, src/qrot_solvers.h:15-30); default 0 */
    int ls_max_iter;     /* strong-Wolfe trials of sparse Newton / SPLR (the max_iter argument
                            of Problem::line_search_wolfe, src/sinkhorn_problem.h:136-144);
                            <= 0: the reference's 20 */
} regot_opts;

void regot_opts_default(regot_opts* o, int solver);

/*
 * Run a solver on a problem handle.
 * Replaces sinkhorn_{bcd,ssns,sparse_newton,splr,newton,lbfgs_dual,apdagd}_internal,
 * src/sinkhorn_solvers.h:40-94 (drivers src/sinkhorn_bcd.cpp:14-124,
 * src/sinkhorn_ssns.cpp:16-234, src/sinkhorn_sparse_newton.cpp:24-170,
 * src/sinkhorn_splr.cpp:24-199, src/sinkhorn_newton.cpp:17-129,
 * src/sinkhorn_lbfgs_dual.cpp:84-125, src/sinkhorn_apdagd.cpp:14-183).
 * The plan is not formed here, except APDAGD's primal plan, which the result
 * keeps on the device; see regot_result_plan().  verbose >= 1 prints the
 * reference's progress lines to stdout.  A failed L-BFGS line search returns
 * -6 with LBFGSpp's message.
 */
int regot_solve(regot_problem* p, int solver, double tol, int max_iter, int verbose,
                const regot_opts* opts, regot_result** out);

/* Result accessors (SinkhornResult, src/sinkhorn_result.h:11-51). */
int regot_result_niter(const regot_result* r);
int regot_result_hist_len(const regot_result* r);       /* niter + 1 */
int regot_result_dens_len(const regot_result* r);       /* 0 for BCD */
/* Copies out: dual (n+m-1), obj_vals, mar_errs, run_times (hist_len),
 * densities (dens_len).  Any pointer may be NULL. */
int regot_result_get(const regot_result* r, double* dual, double* obj_vals, double* mar_errs,
                     double* run_times, double* densities);
/* FP64 plan, row-major n x m, into host memory: APDAGD's primal plan, else
 * the plan of the result's dual (SinkhornResult::get_plan). */
int regot_result_plan(regot_problem* p, const regot_result* r, double* plan);
/* Marginal error ||grad|| re-evaluated at the result's dual in FP64 arithmetic
 * (exp in FP64 regardless of bulk), plus the FP64 objective. */
int regot_result_verify(regot_problem* p, const regot_result* r, double* mar_err_fp64,
                        double* obj_fp64);
/* Per-phase statistics of the Newton-type and first-order drivers (no
 * reference counterpart; zeros for BCD): out[8] = {objective/gradient ms,
 * count, sparsification ms, count, CG ms, CG solves, CG iterations,
 * algorithmic HBM bytes}.  Host wall time per phase; the bytes count one
 * read of M per objective or gradient, one read of M plus the CSR and CSC
 * writes per Hessian build and the CSR + CSC reads per Hessian product. */
int regot_result_stats(const regot_result* r, double* out);
/* BCD in FP32 bulk (no reference counterpart; the reference runs FP64
 * throughout): the iterate from which the loop ran with FP64 exp and sums
 * over the FP32 M -- 0 when tol lies below the FP32 resolution of the
 * marginals from the start, k when ||g|| stalled above tol at iterate k --
 * or -1 when every iteration ran in FP32. */
int regot_result_precise_from(const regot_result* r);
void regot_result_free(regot_result* r);

/*
 * Batched BCD: many independent problems solved together (no reference
 * counterpart; each problem follows sinkhorn_bcd_internal,
 * src/sinkhorn_bcd.cpp:14-124, exactly as regot_solve with REGOT_SOLVER_BCD
 * does, but only the final state is kept -- no per-iteration histories).
 *
 * regot_batch_create copies every problem into one packed device buffer:
 *   n[p], m[p]      sizes of problem p;
 *   M[p]            row-major n[p] x m[p] cost matrix (host memory, or device
 *                   memory on `device` when M_device = 1), element type M_dtype;
 *   bulk            REGOT_F32 or REGOT_F64, as for regot_problem_create;
 *   a[p], b[p]      host FP64 marginals; reg[p] > 0.
 */
typedef struct regot_batch regot_batch;
int regot_batch_create(regot_batch** out, int nprob, const int* n, const int* m,
                       const void* const* M, int M_dtype, int bulk, int M_device,
                       const double* const* a, const double* const* b, const double* reg,
                       int device);
int regot_batch_destroy(regot_batch* batch);
/* Options of the batched solver (no reference counterpart):
 * REGOT_BATCH_GRAPH     1: every group of 8 rounds is one hipGraph launch
 *                       (captured once per tile-table configuration and
 *                       replayed); 0 (default): the rounds' kernels are
 *                       launched one by one, which measured slightly faster
 *                       (the GPU, not the launch path, bounds the rounds, and
 *                       shrinking tables force re-captures).  Results do not
 *                       depend on this option.
 * REGOT_BATCH_AUTOTUNE  1: the next solve first runs one whole solve per
 *                       coarse-tile candidate (rows per wave {2, 4} FP32 /
 *                       {1, 2} FP64 x column-block rows {256, 512, 1024}) and
 *                       keeps the fastest for the handle; 0 (default): fixed
 *                       rules (4 / 2 rows per wave, 512 rows).  Results depend
 *                       on the tiles at rounding level.
 * REGOT_BATCH_FUSED     1 (default): FP32 problems with 1024 < m <= 4096 and
 *                       min a min b >= 2^-60 sum a sum b run single-read
 *                       rounds (one read of M per BCD iteration, the
 *                       iteration of the single-problem solver's single-read
 *                       pass); 0: every problem runs two-pass rounds (row
 *                       tiles, then column tiles).  Results agree to rounding
 *                       (the sums run in another order). */
#define REGOT_BATCH_GRAPH 0
#define REGOT_BATCH_AUTOTUNE 1
#define REGOT_BATCH_FUSED 2
int regot_batch_set_option(regot_batch* batch, int option, long long value);
/* out[12]: rows per wave, column-block rows, tuned (0/1), graph
 * instantiations, graph launches, the six candidates' solve times (ms; 0 for
 * candidates not run), then the number of problems on single-read rounds. */
int regot_batch_stats(const regot_batch* batch, double* out);
/* kind[p] = 1 if problem p runs single-read rounds (one read of its cost
 * matrix per iteration), 0 if two-pass rounds (two reads), under the current
 * options; kind has nprob entries. */
int regot_batch_round_kinds(const regot_batch* batch, int* kind);
/* Number of problems, total n*m elements, and the packed dual length
 * sum_p (n[p] + m[p] - 1). */
int regot_batch_info(const regot_batch* batch, int* nprob, long long* total_elems,
                     long long* dual_len);
/* Solve every problem from the reference's initial point to ||g|| < tol or
 * max_iter rounds.  Per problem: niter, the final objective f and ||g||
 * (that of iterate niter); dual (packed, problem p at sum_{q<p} (n[q]+m[q]-1))
 * in the gamma layout.  Any output may be NULL.  *rounds receives the number
 * of BCD rounds launched. */
int regot_batch_solve_bcd(regot_batch* batch, double tol, int max_iter, int* niter,
                          double* obj, double* mar_err, double* dual, long long* rounds);

/*
 * Squared-Euclidean cost from BF16 point clouds on the matrix cores (no
 * reference counterpart; the cost of the differentiable loss):
 * M[i][j] = |x_i|^2 + |y_j|^2 - 2 x_i.y_j, X (n x d, row pitch ldx) and Y
 * (m x d, pitch ldy) BF16 device arrays, M FP32 device array (pitch ldm).
 * The contraction accumulates in FP32; norms and the sum are FP64.
 */
int regot_sqeuclid_bf16(const void* X, long long ldx, const void* Y, long long ldy, int n, int m,
                        int d, float* M, long long ldm, int device);

/*
 * Gradient of a loss of the optimal plan with respect to M, by the implicit
 * function theorem on the dual optimality conditions (no reference
 * counterpart).  gamma (host, n+m-1) is the dual optimum of the handle's
 * problem; loss 0 = entropic loss <T,M> - reg h(T) (= -f(gamma), the entropy
 * of SinkhornProblem::primal_val, src/sinkhorn_problem_obj.cpp:300-319),
 * 1 = sharp loss <T,M>.  Solves H w = dL/dgamma by the device CG of the
 * sparsified Hessian with every entry kept (cg_tol <= 0: 1e-10), then writes
 * G = dL/dM (n x m, row pitch ldg, REGOT_F32 or REGOT_F64 device array).
 * Optional outputs: the loss value, w (host, n+m-1), CG iterations and the
 * final relative residual.
 */
int regot_implicit_grad(regot_problem* p, const double* gamma, int loss, double cg_tol,
                        int cg_max_iter, void* G, int G_dtype, long long ldg, double* loss_value,
                        double* w_out, int* cg_iters, double* cg_err);

/*
 * Quadratically regularised OT (QROT) on the same problem handle: the dual
 * variable is gamma = [alpha (n); beta (m)] and
 *   f = 0.5 ||(alpha (+) beta - M)_+||^2 - reg (a'alpha + b'beta) + tau/2 (1'alpha - 1'beta)^2.
 * regot_qrot_solve replaces qrot_bcd_internal, qrot_gd_internal,
 * qrot_lbfgs_dual_internal, qrot_lbfgs_semi_dual_internal and
 * qrot_apdagd_internal (src/qrot_apdagd.cpp:14-181), qrot_grssn_internal
 * (src/qrot_grssn.cpp:16-133), qrot_assn_internal (src/qrot_assn.cpp:16-170)
 * qrot_s5n_internal (src/qrot_s5n.cpp:16-196) and qrot_pdaam_internal
 * (src/qrot_pdaam.cpp:14-188)
 * (src/qrot_bcd.cpp:15-101, src/qrot_gd.cpp:15-121,
 * src/qrot_lbfgs_dual.cpp:85-126, src/qrot_lbfgs_semi_dual.cpp:78-118;
 * bindings src/solvers.cpp:53-164).  Its result
 * carries prim_vals (regot_result_prim_vals, one per history entry) and a
 * dual of n + m entries (regot_result_dual_len); mar_errs are ||grad|| / reg;
 * regot_result_plan gives (alpha (+) beta - M)_+ / reg.  opts may be NULL;
 * opts->x0 (n + m; the semi-dual uses its first n) and opts->tau are used
 * (BCD and the semi-dual force tau = 0).
 */
#define REGOT_QROT_BCD 0
#define REGOT_QROT_GD 1
#define REGOT_QROT_LBFGS_DUAL 2
#define REGOT_QROT_LBFGS_SEMI_DUAL 3   /* src/qrot_lbfgs_semi_dual.cpp:78-118 */
#define REGOT_QROT_APDAGD 4            /* src/qrot_apdagd.cpp:14-181; the result keeps its
                                          averaged primal plan (regot_result_plan) */
#define REGOT_QROT_GRSSN 5             /* src/qrot_grssn.cpp:16-133; reads opts->shift
                                          (default 1e-3), opts->method (0, the default
                                          with NULL opts: CG to 1e-8; other values name
                                          the reference's direct factorisations, run as
                                          CG to 1e-12), opts->cg_tol / cg_max_iter */
#define REGOT_QROT_ASSN 6              /* src/qrot_assn.cpp:16-170; reads opts->method
                                          and cg_tol / cg_max_iter as GRSSN (its shift
                                          is adaptive) */
#define REGOT_QROT_S5N 7               /* src/qrot_s5n.cpp:16-196; reads opts->method,
                                          opts->mu0 (default 1; with NULL opts 1) and
                                          cg_tol / cg_max_iter */
#define REGOT_QROT_PDAAM 8             /* src/qrot_pdaam.cpp:14-188; the result keeps its
                                          averaged primal plan (regot_result_plan) */
int regot_qrot_solve(regot_problem* p, int solver, double tol, int max_iter, int verbose,
                     const regot_opts* opts, regot_result** out);
int regot_result_prim_vals(const regot_result* r, double* out);
/* QROT generalised Hessian solve at host gamma (n + m): out = (H + shift I +
 * tau vv')^-1 rhs, v = [1 (n); -1 (m)], H = [diag(sigma 1) sigma; sigma'
 * diag(sigma' 1)], sigma_ij = [alpha_i + beta_j - M_ij > 0]: the operator
 * behind QROT::Hessian::compute_D_hess and hess_cg (src/qrot_hess.cpp:42-96,
 * src/qrot_cg.cpp:9-51; CG to tol with at most max_iter steps, <= 0: 2m);
 * *cg_iters (may be NULL) receives the CG steps. */
int regot_qrot_hess_solve(regot_problem* p, const double* gamma, const double* rhs,
                          double shift, double tau, double tol, int max_iter, double* out,
                          int* cg_iters);
int regot_result_dual_len(const regot_result* r);
/* QROT Problem::dual_obj_grad (src/qrot_problem.cpp:226-258): f and grad (n + m,
 * may be NULL) at host gamma (n + m). */
int regot_qrot_obj_grad(regot_problem* p, const double* gamma, double tau, double* f,
                        double* grad);
/* QROT Problem::optimal_beta / optimal_alpha (src/qrot_problem.cpp:397-418). */
int regot_qrot_optimal_beta(regot_problem* p, const double* alpha, double* beta);
int regot_qrot_optimal_alpha(regot_problem* p, const double* beta, double* alpha);
/* QROT Problem::primal_val (src/qrot_problem.cpp:10-26): <P, M> + reg/2 ||P||^2 of
 * P = approx_proj((alpha (+) beta - M)_+ / reg) at host gamma (n + m). */
int regot_qrot_primal_val(regot_problem* p, const double* gamma, double* value);

/* Device-side timing of the hot kernels (HIP events on the handle's stream).
 * kind: 0 = row pass, 1 = column pass (partial + combine), 2 = single-read
 * BCD pass, 3 = its column-sum combine.  Enabled by regot_profile_enable(p, 1);
 * totals accumulate until reset. */
int regot_profile_enable(regot_problem* p, int on);
int regot_profile_get(regot_problem* p, int kind, long long* launches, double* total_ms);
int regot_profile_reset(regot_problem* p);

#ifdef __cplusplus
}
#endif
#endif
