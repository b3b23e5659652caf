/*
 * erot_oracle.h -- FP64 CPU restatement of the EROT path of
 * yixuan/regot-python v0.0.3 (src/sinkhorn_*).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity checker: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 * The product (libregot_hip.so and the regot package) never links or calls it.
 *
 * Storage: the cost matrix M and every dense n x m array here are row-major
 * (C order), the layout the GPU build keeps in HBM.  The reference stores them
 * column-major (Eigen); only the summation order differs.
 *
 * Every function names the reference file:line it restates.
 */
#ifndef EROT_ORACLE_H
#define EROT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Objectives for the generic line searches and L-BFGS: f(x), and f(x) with
 * its gradient written to grad. */
typedef double (*orc_objgrad_fn)(void* ctx, const double* x, double* grad);
typedef double (*orc_obj_fn)(void* ctx, const double* x);

/* ---- Problem dense passes (src/sinkhorn_problem_bcd.cpp, _obj.cpp) ---- */
void orc_optimal_beta(int n, int m, const double* M, const double* b, double reg,
                      const double* alpha, double* beta);
void orc_optimal_alpha(int n, int m, const double* M, const double* a, double reg,
                       const double* beta, double* alpha);
double orc_compute_T(int n, int m, const double* M, double reg,
                     const double* gamma, double* T);
void orc_compute_sums(int n, int m, const double* T, double* Tsums);
double orc_dual_obj(int n, int m, const double* M, const double* a, const double* b,
                    double reg, const double* gamma);
double orc_dual_obj_grad(int n, int m, const double* M, const double* a, const double* b,
                         double reg, const double* gamma, double* grad,
                         double* T, int computeT);
void orc_get_plan(int n, int m, const double* M, double reg, const double* gamma,
                  double* plan);

/* ---- Sparsifiers (src/sinkhorn_sparsify3.cpp, src/sinkhorn_sparsify4.cpp) ----
 * Output: CSC of the n x (m-1) kept block of T.  colptr has m entries;
 * *rowind and *val are malloc'ed here and released with orc_free(). */
long long orc_sparsify_mat3(int n, int m, const double* T, double delta, double density_hint,
                            int* colptr, int** rowind, double** val);
long long orc_sparsify_density(int n, int m, const double* T, double density, int preserve,
                               int* colptr, int** rowind, double** val);
void orc_free(void* p);

/* ---- Hessian operators (src/sinkhorn_hess.cpp:99-164) on an explicit CSC ---- */
void orc_apply_sigmadx(int n, int m, const int* colptr, const int* rowind, const double* val,
                       const double* x, double* res);
void orc_apply_sigmadtx(int n, int m, const int* colptr, const int* rowind, const double* val,
                        const double* x, double* res);
void orc_apply_deltax(int n, int m, const double* h1, const double* h2,
                      const int* colptr, const int* rowind, const double* val,
                      const double* x, double shift, double* res);
void orc_apply_hsx(int n, int m, const double* h1, const double* h2,
                   const int* colptr, const int* rowind, const double* val,
                   const double* x, double* res);
int orc_hess_cg(int n, int m, const double* h1, const double* h2,
                const int* colptr, const int* rowind, const double* val,
                const double* rhs, double shift, double tol, int max_iter,
                double* res, double* tol_error);
int orc_hess_direct(int n, int m, const double* h1, const double* h2,
                    const int* colptr, const int* rowind, const double* val,
                    const double* rhs, double shift, double* res);

/* SPLR's inverse-BFGS combination given A^-1 r and A^-1 y
 * (src/sinkhorn_linsolve.cpp:207-213). */
void orc_low_rank_combine(long N, const double* iAr, const double* iAy, const double* r,
                          const double* y, const double* s, double* res);

/* ---- Line-search helpers (src/sinkhorn_problem_ls.cpp:72-223) ---- */
double orc_cubic_minimizer(double a, double b, double fa, double fb, double ga, double gb,
                           int* exists);
double orc_step_selection(double al, double au, double at, double fl, double fu, double ft,
                          double gl, double gu, double gt);

/* ---- Solver drivers (src/sinkhorn_{bcd,ssns,sparse_newton,splr}.cpp) ----
 * History arrays need max_iter + 1 slots; dual needs n + m - 1.
 * Returns niter; *hist_len receives the number of history entries.
 * lin_method: 0 = CG, otherwise a dense Cholesky solve of (Hs + shift I)
 * (the mathematical equivalent of the reference's SimplicialLDLT; small sizes only).
 * cg_tol <= 0 keeps the driver's hard-coded 1e-8. */
typedef struct {
    int lin_method;
    double mu0;
    double shift;
    double density;
    double cg_tol;
    int cg_max_iter;      /* <= 0: Eigen's default 2 * HessianCG::cols() = 2m */
    int lowrank_cg;       /* SPLR: 1 = CG for the two low-rank solves instead of direct */
    int dense_T;          /* 1: keep the n x m T like the reference (needed by Newton drivers) */
    int ls_max_iter;      /* Wolfe trials of sparse Newton / SPLR; <= 0: the reference's 20 */
} orc_opts;

/* ---- Line searches (src/sinkhorn_problem_ls.cpp) over generic objectives ---- */
double orc_line_selection(long N, orc_obj_fn fobj, void* ctx, const double* cand, int nc,
                          const double* gamma, const double* direc, double curobj,
                          double* objval, double* work, double* last_cand);
double orc_line_search_wolfe(long N, orc_objgrad_fn fg, void* ctx, double init_step,
                             const double* gamma, const double* direc, double curobj,
                             const double* curgrad, int max_iter, int* recompute_T, double* x,
                             double* grad, double* last_step, int* ntrials);
double orc_erot_line_search_wolfe(int n, int m, const double* M, const double* a,
                                  const double* b, double reg, double init_step,
                                  const double* gamma, const double* direc, double curobj,
                                  const double* curgrad, int max_iter, int* recompute_T,
                                  double* x, double* grad, double* last_step, int* ntrials);

/* ---- Driver trace: one record per SSNS / sparse Newton / SPLR iteration ----
 * ORC_TRACE_HDR scalars, then gamma (before the step), the point whose T is
 * in use, the direction, y and s (SPLR's curvature pair, zeros otherwise),
 * each n + m - 1 long.
 * SSNS scalars:   kind 1, iter, f, ||g||, shift, mu, delta, step, f at the
 *                 step, rho, its numerator, its denominator, last candidate
 *                 evaluated, H density, mu after the update, 0.
 * sparse Newton (kind 2) / SPLR (kind 3): kind, iter, f, ||g||, shift,
 *                 density of H, initial step, step, recompute_T, low-rank
 *                 flag, Wolfe trials, last trial step, 0, 0, 0, 0. */
#define ORC_TRACE_HDR 16
void orc_trace_set(double* buf, long cap);
long orc_trace_used(void);

int orc_sinkhorn_bcd(int n, int m, const double* M, const double* a, const double* b,
                     double reg, double tol, int max_iter, int verbose,
                     double* dual, double* obj_vals, double* mar_errs, double* run_times,
                     int* hist_len);
int orc_sinkhorn_ssns(int n, int m, const double* M, const double* a, const double* b,
                      double reg, double tol, int max_iter, int verbose, const orc_opts* opts,
                      double* dual, double* obj_vals, double* mar_errs, double* run_times,
                      double* densities, int* hist_len);
int orc_sinkhorn_sparse_newton(int n, int m, const double* M, const double* a, const double* b,
                               double reg, double tol, int max_iter, int verbose,
                               const orc_opts* opts,
                               double* dual, double* obj_vals, double* mar_errs,
                               double* run_times, double* densities, int* hist_len);
int orc_sinkhorn_splr(int n, int m, const double* M, const double* a, const double* b,
                      double reg, double tol, int max_iter, int verbose, const orc_opts* opts,
                      double* dual, double* obj_vals, double* mar_errs, double* run_times,
                      double* densities, int* hist_len);

/* ---- Dense Newton, L-BFGS dual, APDAGD (src/sinkhorn_{newton,lbfgs_dual,apdagd}.cpp) ---- */
/* Newton direction -H^{-1} g with the dense Hessian and a Cholesky solve
 * (src/sinkhorn_problem_obj.cpp:396-418, src/sinkhorn_newton.cpp:87-88). */
int orc_newton_direction(int n, int m, const double* M, const double* a, const double* b,
                         double reg, const double* gamma, double* direc);
int orc_sinkhorn_newton(int n, int m, const double* M, const double* a, const double* b,
                        double reg, double tol, int max_iter, int verbose,
                        double* dual, double* obj_vals, double* mar_errs, double* run_times,
                        int* hist_len);
/* Two-loop L-BFGS product H v after adding npairs pairs (S[k], Y[k]) in order
 * (src/LBFGSpp/LBFGSpp/BFGSMat.h:86-97,276-302). */
void orc_lbfgs_apply(long N, int mm, int npairs, const double* S, const double* Y,
                     const double* v, double* res);
/* One strong-Wolfe line search on the dual (LineSearchNocedalWright). */
int orc_nw_line_search(int n, int m, const double* M, const double* a, const double* b,
                       double reg, const double* gamma, const double* direc, double init_step,
                       double* step, double* f, double* grad, double* last_f);
int orc_sinkhorn_lbfgs_dual(int n, int m, const double* M, const double* a, const double* b,
                            double reg, double tol, int max_iter, int verbose, int hist_cap,
                            double* dual, double* obj_vals, double* mar_errs, double* run_times,
                            int* hist_len);
/* plan (n x m) receives the primal plan. */
int orc_sinkhorn_apdagd(int n, int m, const double* M, const double* a, const double* b,
                        double reg, double tol, int max_iter, int verbose,
                        double* dual, double* plan, double* obj_vals, double* mar_errs,
                        double* run_times, int* hist_len);

/* ---- Feasible rounding and primal value (src/approx_proj.h, sinkhorn_problem_obj.cpp:300-319) ---- */
void orc_approx_proj(int n, int m, const double* plan, const double* a, const double* b,
                     double* out);
double orc_primal_val(int n, int m, const double* M, const double* a, const double* b,
                      double reg, const double* gamma);

/* ---- Generic L-BFGS (LBFGSSolver::minimize with the Nocedal-Wright search) ---- */
typedef void (*orc_iter_fn)(void* ctx, const double* x, double last_f, double gnorm);
int orc_lbfgs_minimize(long N, orc_objgrad_fn fn, void* fctx, double eps_abs, double eps_rel,
                       int max_iter, orc_iter_fn it, void* ictx, double* x);

/* ---- QROT (qrot_oracle.c): gamma = [alpha (n); beta (m)] ---- */
double orc_qrot_dual_obj_grad(int n, int m, const double* M, const double* a, const double* b,
                              double reg, double tau, const double* gamma, double* grad);
double orc_qrot_pl_root(int len, const double* y, double c);
void orc_qrot_optimal_beta(int n, int m, const double* M, const double* b, double reg,
                           const double* alpha, double* beta);
void orc_qrot_optimal_alpha(int n, int m, const double* M, const double* a, double reg,
                            const double* beta, double* alpha);
void orc_qrot_get_plan(int n, int m, const double* M, double reg, const double* gamma,
                       double* plan);
double orc_qrot_primal_val(int n, int m, const double* M, const double* a, const double* b,
                           double reg, const double* gamma);
/* solver: 0 = BCD, 1 = gradient descent, 2 = L-BFGS on the dual, 3 = L-BFGS on
 * the semi-dual, 4 = APDAGD (plan, n x m, receives its averaged primal plan;
 * may be NULL for the others) */
int orc_qrot_solve(int solver, int n, int m, const double* M, const double* a, const double* b,
                   double reg, double tau, double tol, int max_iter, int verbose, int hist_cap,
                   const double* x0, double* x, double* obj_vals, double* prim_vals,
                   double* mar_errs, double* run_times, int* hist_len, double* plan);
/* (H + shift I + tau vv')^-1 rhs with the generalised Hessian at gamma;
 * returns the CG steps. */
int orc_qrot_hess_solve(int n, int m, const double* M, const double* gamma, const double* rhs,
                        double shift, double tau, double tol, int max_iter, double* out);
int orc_qrot_grssn(int n, int m, const double* M, const double* a, const double* b, double reg,
                   double tau, double shift, int method, double tol, int max_iter, int verbose,
                   int hist_cap, const double* x0, double* x, double* obj_vals,
                   double* prim_vals, double* mar_errs, double* run_times, int* hist_len,
                   int* cg_iters);
int orc_qrot_assn(int n, int m, const double* M, const double* a, const double* b, double reg,
                  double tau, int method, double tol, int max_iter, int verbose, int hist_cap,
                  const double* x0, double* x, double* obj_vals, double* prim_vals,
                  double* mar_errs, double* run_times, int* hist_len, int* cg_iters);
int orc_qrot_s5n(int n, int m, const double* M, const double* a, const double* b, double reg,
                 double tau, int method, double mu0, double tol, int max_iter, int verbose,
                 int hist_cap, const double* x0, double* x, double* obj_vals, double* prim_vals,
                 double* mar_errs, double* run_times, int* hist_len, int* cg_iters);

/* BCD iterations only (no convergence test), for timing a bounded CPU sample. */
double orc_bcd_iterations(int n, int m, const double* M, const double* a, const double* b,
                          double reg, int iters, double* gamma_out);

#ifdef __cplusplus
}
#endif
#endif
