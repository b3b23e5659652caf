/*
 * qrot_oracle.c -- FP64 CPU restatement of the QROT (quadratically
 * regularised OT) dual passes and drivers of yixuan/regot-python v0.0.3.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP build.  Only
 * tests/ load it.
 *
 * Storage is row-major (C order): M[i*m + j].  The dual variable is
 * gamma = [alpha (n); beta (m)] (no elimination of beta_m, unlike EROT).
 */
#include "erot_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double q_now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static double q_norm2(const double* x, long n)
{
    double s = 0.0;
    for (long i = 0; i < n; i++) s += x[i] * x[i];
    return sqrt(s);
}

/* f = 0.5 ||D||_F^2 - reg (a'alpha + b'beta) + tau/2 (1'alpha - 1'beta)^2,
 * D = (alpha_i + beta_j - M_ij)_+; grad = [D 1 - reg a + tau d; D'1 - reg b -
 * tau d], d = 1'alpha - 1'beta (grad may be NULL).  Restates
 * compute_D_stats and Problem::dual_obj_grad, src/qrot_problem.cpp:37-96,
 * 226-258. */
double orc_qrot_dual_obj_grad(int n, int m, const double* M, const double* a, const double* b,
                              double reg, double tau, const double* gamma, double* grad)
{
    const double* alpha = gamma;
    const double* beta = gamma + n;
    double sq = 0.0;
    if (grad) memset(grad, 0, sizeof(double) * (size_t)(n + m));
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < m; j++) {
            const double c = alpha[i] + beta[j] - M[(long)i * m + j];
            if (c > 0.0) {
                sq += c * c;
                if (grad) {
                    grad[i] += c;
                    grad[n + j] += c;
                }
            }
        }
    }
    double aa = 0.0, bb = 0.0, sa = 0.0, sb = 0.0;
    for (int i = 0; i < n; i++) {
        aa += alpha[i] * a[i];
        sa += alpha[i];
    }
    for (int j = 0; j < m; j++) {
        bb += beta[j] * b[j];
        sb += beta[j];
    }
    double f = 0.5 * sq - reg * (aa + bb);
    const double d = sa - sb;
    if (tau > 0.0) f += 0.5 * tau * d * d;
    if (grad) {
        for (int i = 0; i < n; i++) grad[i] -= reg * a[i] - (tau > 0.0 ? tau * d : 0.0);
        for (int j = 0; j < m; j++) grad[n + j] -= reg * b[j] + (tau > 0.0 ? tau * d : 0.0);
    }
    return f;
}

static int q_cmp(const void* x, const void* y)
{
    const double u = *(const double*)x, v = *(const double*)y;
    return (u < v) ? -1 : (u > v) ? 1 : 0;
}

/* The root x of sum_k (x - y_k)_+ = c (c > 0): sort y ascending and find the
 * segment [y_(i), y_(i+1)) where the left side crosses c; then x = (c +
 * sum_{k<=i} y_(k)) / (i + 1).  Restates piecewise_linear_equation,
 * src/qrot_problem.cpp:365-394. */
double orc_qrot_pl_root(int len, const double* y, double c)
{
    double* ys = (double*)malloc(sizeof(double) * (size_t)len);
    memcpy(ys, y, sizeof(double) * (size_t)len);
    qsort(ys, (size_t)len, sizeof(double), q_cmp);
    double cum = 0.0, f_lo = 0.0, x = 0.0;
    int found = 0;
    for (int i = 0; i < len - 1; i++) {
        cum += ys[i];
        const double f_hi = (i + 1) * ys[i + 1] - cum;
        if (c >= f_lo && c < f_hi) {
            x = (c + cum) / (i + 1);
            found = 1;
            break;
        }
        f_lo = f_hi;
    }
    if (!found) {
        cum += ys[len - 1];
        x = (c + cum) / len;
    }
    free(ys);
    return x;
}

/* beta_j = root of sum_i (x - (M_ij - alpha_i))_+ = reg b_j.
 * Problem::optimal_beta, src/qrot_problem.cpp:397-406. */
void orc_qrot_optimal_beta(int n, int m, const double* M, const double* b, double reg,
                           const double* alpha, double* beta)
{
    double* y = (double*)malloc(sizeof(double) * (size_t)n);
    for (int j = 0; j < m; j++) {
        for (int i = 0; i < n; i++) y[i] = M[(long)i * m + j] - alpha[i];
        beta[j] = orc_qrot_pl_root(n, y, reg * b[j]);
    }
    free(y);
}

/* alpha_i = root of sum_j (x - (M_ij - beta_j))_+ = reg a_i.
 * Problem::optimal_alpha, src/qrot_problem.cpp:409-418. */
void orc_qrot_optimal_alpha(int n, int m, const double* M, const double* a, double reg,
                            const double* beta, double* alpha)
{
    double* y = (double*)malloc(sizeof(double) * (size_t)m);
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < m; j++) y[j] = M[(long)i * m + j] - beta[j];
        alpha[i] = orc_qrot_pl_root(m, y, reg * a[i]);
    }
    free(y);
}

/* plan_ij = (alpha_i + beta_j - M_ij)_+ / reg.  QROTResult::get_plan,
 * src/qrot_result.h:35-49. */
void orc_qrot_get_plan(int n, int m, const double* M, double reg, const double* gamma,
                       double* plan)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++) {
            const double c = gamma[i] + gamma[n + j] - M[(long)i * m + j];
            plan[(long)i * m + j] = (c > 0.0 ? c : 0.0) / reg;
        }
}

/* <P, M> + reg/2 ||P||_F^2 of P = approx_proj(plan(gamma)).
 * Problem::primal_val, src/qrot_problem.cpp:10-26. */
double orc_qrot_primal_val(int n, int m, const double* M, const double* a, const double* b,
                           double reg, const double* gamma)
{
    double* T = (double*)malloc(sizeof(double) * (size_t)n * m);
    double* P = (double*)malloc(sizeof(double) * (size_t)n * m);
    orc_qrot_get_plan(n, m, M, reg, gamma, T);
    orc_approx_proj(n, m, T, a, b, P);
    double mp = 0.0, sq = 0.0;
    for (long q = 0; q < (long)n * m; q++) {
        mp += P[q] * M[q];
        sq += P[q] * P[q];
    }
    free(T);
    free(P);
    return mp + 0.5 * reg * sq;
}

typedef struct {
    int n, m;
    const double *M, *a, *b;
    double reg, tau;
    int h, hist_cap, verbose;
    double t0;
    double *obj_vals, *prim_vals, *mar_errs, *run_times;
} qrot_ctx;

static double qrot_objgrad(void* ctx, const double* x, double* grad)
{
    const qrot_ctx* Q = (const qrot_ctx*)ctx;
    return orc_qrot_dual_obj_grad(Q->n, Q->m, Q->M, Q->a, Q->b, Q->reg, Q->tau, x, grad);
}

/* One history entry: f, the primal value, ||grad|| / reg and the time since
 * the previous entry (src/qrot_bcd.cpp:58-64, src/qrot_lbfgs_dual.cpp:56-73). */
static void qrot_record(qrot_ctx* Q, const double* x, double f, double gnorm)
{
    const double t1 = q_now_ms();
    if (Q->h < Q->hist_cap) {
        Q->obj_vals[Q->h] = f;
        Q->prim_vals[Q->h] = orc_qrot_primal_val(Q->n, Q->m, Q->M, Q->a, Q->b, Q->reg, x);
        Q->mar_errs[Q->h] = gnorm / Q->reg;
        Q->run_times[Q->h] = (Q->h > 0 ? Q->run_times[Q->h - 1] : 0.0) + (t1 - Q->t0);
        Q->h++;
    }
    Q->t0 = q_now_ms();
}

static void qrot_lbfgs_iter(void* ctx, const double* x, double last_f, double gnorm)
{
    qrot_ctx* Q = (qrot_ctx*)ctx;
    qrot_record(Q, x, last_f, gnorm);
    if (Q->verbose >= 1)
        printf("iter = %d, objval = %g, ||grad|| = %g\n", Q->h - 1, last_f, gnorm);
}

/* Semi-dual objective at alpha (n): beta = optimal_beta(alpha), f = the dual
 * objective at [alpha; beta] (tau = 0), grad = D 1 - reg a.  Restates
 * Problem::semi_dual_obj_grad, src/qrot_problem.cpp:421-441. */
static double qrot_semi_objgrad(void* ctx, const double* x, double* grad)
{
    const qrot_ctx* Q = (const qrot_ctx*)ctx;
    const int n = Q->n, m = Q->m;
    double* g = (double*)malloc(sizeof(double) * (size_t)(n + m));
    double* gm = (double*)malloc(sizeof(double) * (size_t)(n + m));
    memcpy(g, x, sizeof(double) * n);
    orc_qrot_optimal_beta(n, m, Q->M, Q->b, Q->reg, x, g + n);
    const double f = orc_qrot_dual_obj_grad(n, m, Q->M, Q->a, Q->b, Q->reg, 0.0, g,
                                            grad ? gm : NULL);
    if (grad) memcpy(grad, gm, sizeof(double) * n);
    free(g);
    free(gm);
    return f;
}

/* iterate() of src/qrot_lbfgs_semi_dual.cpp:51-75: the primal value at
 * [alpha; optimal_beta(alpha)]. */
static void qrot_semi_iter(void* ctx, const double* x, double last_f, double gnorm)
{
    qrot_ctx* Q = (qrot_ctx*)ctx;
    double* g = (double*)malloc(sizeof(double) * (size_t)(Q->n + Q->m));
    memcpy(g, x, sizeof(double) * Q->n);
    orc_qrot_optimal_beta(Q->n, Q->m, Q->M, Q->b, Q->reg, x, g + Q->n);
    qrot_record(Q, g, last_f, gnorm);
    free(g);
    if (Q->verbose >= 1)
        printf("iter = %d, objval = %g, ||grad|| = %g\n", Q->h - 1, last_f, gnorm);
}

/* APDAGD on the QROT dual (src/qrot_apdagd.cpp:14-181): dual variables
 * lambda = -gamma, L halved per iteration and doubled per failed test (at
 * most 20 inner trials, the last one accepted), the primal plan the running
 * combination (1 - tau) P + tau plan(-lambda~) starting from plan(gamma_0);
 * histories: f of the last objective evaluated, the primal value of the
 * rounded plan at -lambda~, the full marginal error of P (all m columns);
 * stop when <M, P> + reg/2 ||P||^2 - phi <= tol and the marginal error <=
 * tol.  x receives -lambda~ of the last iteration, plan the averaged P. */
static int qrot_apdagd(qrot_ctx* Q, double tol, int max_iter, double* x, double* plan)
{
    const int n = Q->n, m = Q->m;
    const long N = (long)n + m;
    const double reg = Q->reg;
    double* grad = (double*)malloc(sizeof(double) * N);
    double* dual = (double*)malloc(sizeof(double) * N);
    double* dualp = (double*)malloc(sizeof(double) * N);
    double* dualt = (double*)malloc(sizeof(double) * N);
    double* dual_new = (double*)malloc(sizeof(double) * N);
    double* dualp_new = (double*)malloc(sizeof(double) * N);
    double* T = (double*)malloc(sizeof(double) * (size_t)n * m);
    for (long q = 0; q < N; q++) dual[q] = dualp[q] = dualt[q] = -x[q];
    orc_qrot_get_plan(n, m, Q->M, reg, x, plan);
    double f = qrot_objgrad(Q, x, grad);
    qrot_record(Q, x, f, q_norm2(grad, N));
    double L = 1.0, pbeta = 0.0;
    int it;
    for (it = 0; it < max_iter; it++) {
        L *= 0.5;
        double ptau = 0.0, phi_new = 0.0;
        for (int j = 0; j < 20; j++) {
            const double palpha = (1.0 + sqrt(4.0 * L * pbeta + 1.0)) / (2.0 * L);
            const double pbeta_new = pbeta + palpha;
            ptau = palpha / pbeta_new;
            for (long q = 0; q < N; q++) {
                dualt[q] = (1.0 - ptau) * dual[q] + ptau * dualp[q];
                x[q] = -dualt[q];
            }
            f = qrot_objgrad(Q, x, grad);
            const double phit = -f / reg;
            double gd = 0.0, dd = 0.0;
            for (long q = 0; q < N; q++) {
                dualp_new[q] = dualp[q] + (palpha / reg) * grad[q];
                dual_new[q] = (1.0 - ptau) * dual[q] + ptau * dualp_new[q];
                x[q] = -dual_new[q];
                const double df = dual_new[q] - dualt[q];
                gd += grad[q] * df;
                dd += df * df;
            }
            f = qrot_objgrad(Q, x, NULL);
            phi_new = -f / reg;
            const double rhs = phit + gd / reg - 0.5 * L * dd;
            if (phi_new >= rhs || j >= 19) {
                pbeta = pbeta_new;
                memcpy(dualp, dualp_new, sizeof(double) * N);
                memcpy(dual, dual_new, sizeof(double) * N);
                break;
            }
            L *= 2.0;
        }
        for (long q = 0; q < N; q++) x[q] = -dualt[q];
        orc_qrot_get_plan(n, m, Q->M, reg, x, T);
        double mp = 0.0, sq = 0.0;
        for (long q = 0; q < (long)n * m; q++) {
            plan[q] = (1.0 - ptau) * plan[q] + ptau * T[q];
            mp += Q->M[q] * plan[q];
            sq += plan[q] * plan[q];
        }
        const double primal_obj = mp + 0.5 * reg * sq;
        double ra = 0.0, rb = 0.0;
        for (int i = 0; i < n; i++) {
            double r = 0.0;
            for (int jj = 0; jj < m; jj++) r += plan[(long)i * m + jj];
            ra += (r - Q->a[i]) * (r - Q->a[i]);
        }
        for (int jj = 0; jj < m; jj++) {
            double c = 0.0;
            for (int i = 0; i < n; i++) c += plan[(long)i * m + jj];
            rb += (c - Q->b[jj]) * (c - Q->b[jj]);
        }
        const double mar_err = sqrt(ra + rb);
        qrot_record(Q, x, f, mar_err * reg);   /* qrot_record divides by reg */
        if (Q->verbose >= 1)
            printf("iter = %d, primal_obj = %g, phi_new = %g, mar_err = %g\n", it, primal_obj,
                   phi_new, mar_err);
        if (primal_obj - phi_new <= tol && mar_err <= tol) break;
    }
    free(grad);
    free(dual);
    free(dualp);
    free(dualt);
    free(dual_new);
    free(dualp_new);
    free(T);
    return it;
}

/* QROT drivers (solver 0 = BCD, 1 = gradient descent, 2 = L-BFGS on the
 * dual, 3 = L-BFGS on the semi-dual, 4 = APDAGD; plan (n x m, solver 4 only)
 * receives the averaged primal plan).  Initial point alpha = 0, beta = optimal_beta(alpha), or x0 (n+m)
 * when non-NULL.  x receives the final dual; histories hold at most
 * hist_cap entries (niter + 1 for BCD and GD).  Returns niter, or a negative
 * line-search status (L-BFGS).
 *   BCD: src/qrot_bcd.cpp:15-101 (tau forced to 0; test ||g|| < tol, then
 *        alpha = optimal_alpha(beta), beta = optimal_beta(alpha)).
 *   GD:  src/qrot_gd.cpp:15-121 (step 1 halved up to 20 times until
 *        f(gamma - step g) <= f - step 0.5 ||g||^2; the last trial is taken).
 *   L-BFGS: src/qrot_lbfgs_dual.cpp:85-126 (epsilon = epsilon_rel = tol).
 *   Semi-dual L-BFGS: src/qrot_lbfgs_semi_dual.cpp:78-118 (over alpha, from
 *        alpha = 0 or x0's head; tau = 0; the result is [alpha;
 *        optimal_beta(alpha)]). */
int orc_qrot_solve(int solver, int n, int m, const double* M, const double* a, const double* b,
                   double reg, double tau, double tol, int max_iter, int verbose, int hist_cap,
                   const double* x0, double* x, double* obj_vals, double* prim_vals,
                   double* mar_errs, double* run_times, int* hist_len, double* plan)
{
    const long N = (long)n + m;
    qrot_ctx Q = {n, m, M, a, b, reg, (solver == 0 || solver == 3) ? 0.0 : tau, 0, hist_cap,
                  verbose, 0.0, obj_vals, prim_vals, mar_errs, run_times};
    if (solver == 3) {
        if (x0)
            memcpy(x, x0, sizeof(double) * n);
        else
            memset(x, 0, sizeof(double) * n);
        Q.t0 = q_now_ms();
        const int rc = orc_lbfgs_minimize(n, qrot_semi_objgrad, &Q, tol, tol, max_iter,
                                          qrot_semi_iter, &Q, x);
        orc_qrot_optimal_beta(n, m, M, b, reg, x, x + n);
        *hist_len = Q.h;
        return rc;
    }
    if (x0) {
        memcpy(x, x0, sizeof(double) * N);
    } else {
        memset(x, 0, sizeof(double) * n);
        orc_qrot_optimal_beta(n, m, M, b, reg, x, x + n);
    }
    Q.t0 = q_now_ms();
    if (solver == 4) {
        Q.tau = 0.0;
        const int it = qrot_apdagd(&Q, tol, max_iter, x, plan);
        *hist_len = Q.h;
        return it;
    }
    if (solver == 2) {
        const int rc = orc_lbfgs_minimize(N, qrot_objgrad, &Q, tol, tol, max_iter,
                                          qrot_lbfgs_iter, &Q, x);
        *hist_len = Q.h;
        return rc;
    }
    double* grad = (double*)malloc(sizeof(double) * N);
    double* xn = (double*)malloc(sizeof(double) * N);
    double f = qrot_objgrad(&Q, x, grad);
    double gnorm = q_norm2(grad, N);
    qrot_record(&Q, x, f, gnorm);
    int it;
    for (it = 0; it < max_iter; it++) {
        if (verbose >= 1) printf("iter = %d, objval = %g, ||grad|| = %g\n", it, f, gnorm);
        if (gnorm < tol) break;
        if (solver == 0) {
            orc_qrot_optimal_alpha(n, m, M, a, reg, x + n, x);
            orc_qrot_optimal_beta(n, m, M, b, reg, x, x + n);
        } else {
            double step = 1.0;
            const double thresh = 0.5 * gnorm * gnorm;
            for (int k = 0; k < 20; k++) {
                for (long q = 0; q < N; q++) xn[q] = x[q] - step * grad[q];
                const double fn = qrot_objgrad(&Q, xn, NULL);
                if (fn <= f - step * thresh) break;
                step *= 0.5;
            }
            memcpy(x, xn, sizeof(double) * N);
        }
        f = qrot_objgrad(&Q, x, grad);
        gnorm = q_norm2(grad, N);
        qrot_record(&Q, x, f, gnorm);
    }
    free(grad);
    free(xn);
    *hist_len = Q.h;
    return it;
}
