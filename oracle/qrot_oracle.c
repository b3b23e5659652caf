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

/* PDAAM on the QROT dual (src/qrot_pdaam.cpp:14-188): lambda = -gamma; per
 * iteration L_new = L / 2, doubled per failed test (at most 20 trials, the
 * last one accepted); alpha_new = (1 + sqrt(4 alpha^2 L L_new + 1)) / (2
 * L_new), tau = 1 / (alpha_new L_new), lambda~ = (1 - tau) lambda + tau
 * lambda'; at gamma = -lambda~ the block with the larger gradient norm is
 * minimised exactly (optimal_alpha or optimal_beta), accepted when phi_new >=
 * phi~ + ||g||^2 / (2 L_new reg^2); then lambda' += (alpha / reg) g and the
 * plan becomes (alpha_old / alpha)^2 (L_old / L) P + tau plan(-lambda~).
 * Histories as APDAGD: f of the last objective, the primal value at
 * -lambda~, the full marginal error of P; stop when <M, P> + reg/2 ||P||^2 -
 * phi_new <= tol and the marginal error <= tol.  x receives -lambda~, plan
 * the averaged P. */
static int qrot_pdaam(qrot_ctx* Q, double tol, int max_iter, double* x, double* plan)
{
    const int n = Q->n, m = Q->m;
    const long N = (long)n + m;
    const double reg = Q->reg;
    double* grad = (double*)malloc(sizeof(double) * N);
    double* dual = (double*)malloc(sizeof(double) * N);
    double* dualp = (double*)malloc(sizeof(double) * N);
    double* dualt = (double*)malloc(sizeof(double) * N);
    double* T = (double*)malloc(sizeof(double) * (size_t)n * m);
    for (long q = 0; q < N; q++) dual[q] = dualp[q] = dualt[q] = -x[q];
    orc_qrot_get_plan(n, m, Q->M, reg, x, plan);
    double f = qrot_objgrad(Q, x, grad);
    qrot_record(Q, x, f, q_norm2(grad, N));
    double L = 1.0, palpha = 0.0;
    int it;
    for (it = 0; it < max_iter; it++) {
        double L_new = 0.5 * L, phi_new = 0.0, c_old = 0.0, c_new = 0.0;
        for (int j = 0; j < 20; j++) {
            const double palpha_new =
                (1.0 + sqrt(4.0 * palpha * palpha * L * L_new + 1.0)) / (2.0 * L_new);
            const double ptau = 1.0 / (palpha_new * L_new);
            for (long q = 0; q < N; q++) {
                dualt[q] = (1.0 - ptau) * dual[q] + ptau * dualp[q];
                x[q] = -dualt[q];
            }
            f = qrot_objgrad(Q, x, grad);
            const double phit = -f / reg;
            double lr = 0.0, mr = 0.0;
            for (int i = 0; i < n; i++) lr += grad[i] * grad[i];
            for (int jj = 0; jj < m; jj++) mr += grad[n + jj] * grad[n + jj];
            if (lr > mr)
                orc_qrot_optimal_alpha(n, m, Q->M, Q->a, reg, x + n, x);
            else
                orc_qrot_optimal_beta(n, m, Q->M, Q->b, reg, x, x + n);
            f = qrot_objgrad(Q, x, NULL);
            phi_new = -f / reg;
            const double rhs = phit + (lr + mr) / (2.0 * L_new * reg * reg);
            if (phi_new >= rhs || j >= 19) {
                const double ar = palpha / palpha_new, lrat = L / L_new;
                c_old = ar * ar * lrat;
                c_new = ptau;
                palpha = palpha_new;
                L = L_new;
                for (long q = 0; q < N; q++) dual[q] = -x[q];
                break;
            }
            L_new *= 2.0;
        }
        for (long q = 0; q < N; q++) {
            dualp[q] += (palpha / reg) * grad[q];
            x[q] = -dualt[q];
        }
        orc_qrot_get_plan(n, m, Q->M, reg, x, T);
        double mp = 0.0, sq = 0.0;
        for (long q = 0; q < (long)n * m; q++) {
            plan[q] = c_old * plan[q] + c_new * T[q];
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
    if (solver == 8) {
        Q.tau = 0.0;
        const int it = qrot_pdaam(&Q, tol, max_iter, x, plan);
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

/* ---- Generalised Hessian and GRSSN ----
 * H = [diag(h1) sigma; sigma' diag(h2)], sigma_ij = [alpha_i + beta_j - M_ij
 * > 0], h1 = sigma 1, h2 = sigma' 1 (src/qrot_hess.h:11-36,
 * src/qrot_hess.cpp:42-96).  (H + shift I + tau vv') x = rhs, v = [1; -1], is
 * solved by block elimination with CG on the Schur complement
 * Delta = diag(h2) + shift I - C A^-1 B over beta (src/qrot_hess.cpp:130-187,
 * src/qrot_cg.cpp:9-51). */
typedef struct {
    int n, m;
    unsigned char* sig; /* n x m, row-major */
    double *h1, *h2, *tn, *tm, *tn2;
    double shift, tau;
} qrot_hess;

static void qh_build(qrot_hess* H, int n, int m, const double* M, const double* gamma)
{
    H->n = n;
    H->m = m;
    H->sig = (unsigned char*)malloc((size_t)n * m);
    H->h1 = (double*)calloc((size_t)n, sizeof(double));
    H->h2 = (double*)calloc((size_t)m, sizeof(double));
    H->tn = (double*)malloc(sizeof(double) * (size_t)n);
    H->tn2 = (double*)malloc(sizeof(double) * (size_t)n);
    H->tm = (double*)malloc(sizeof(double) * (size_t)m);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++) {
            const int on = gamma[i] + gamma[n + j] - M[(long)i * m + j] > 0.0;
            H->sig[(long)i * m + j] = (unsigned char)on;
            H->h1[i] += on;
            H->h2[j] += on;
        }
}

static void qh_free(qrot_hess* H)
{
    free(H->sig);
    free(H->h1);
    free(H->h2);
    free(H->tn);
    free(H->tn2);
    free(H->tm);
}

static double q_sum(const double* x, long n)
{
    double s = 0.0;
    for (long i = 0; i < n; i++) s += x[i];
    return s;
}

/* B x = sigma x - tau (1'x) 1 (n); x has m entries */
static void qh_Bx(const qrot_hess* H, const double* x, double* res)
{
    const double c = H->tau > 0.0 ? H->tau * q_sum(x, H->m) : 0.0;
    for (int i = 0; i < H->n; i++) {
        double s = 0.0;
        for (int j = 0; j < H->m; j++)
            if (H->sig[(long)i * H->m + j]) s += x[j];
        res[i] = s - c;
    }
}

/* C x = sigma' x - tau (1'x) 1 (m); x has n entries */
static void qh_Cx(const qrot_hess* H, const double* x, double* res)
{
    const double c = H->tau > 0.0 ? H->tau * q_sum(x, H->n) : 0.0;
    for (int j = 0; j < H->m; j++) res[j] = 0.0;
    for (int i = 0; i < H->n; i++)
        for (int j = 0; j < H->m; j++)
            if (H->sig[(long)i * H->m + j]) res[j] += x[i];
    for (int j = 0; j < H->m; j++) res[j] -= c;
}

/* A^-1 x with A = diag(h1 + shift) + tau 11' (Sherman-Morrison when tau > 0) */
static void qh_solve_A(const qrot_hess* H, const double* x, double* res)
{
    const int n = H->n;
    if (H->tau > 0.0) {
        double c1 = 1.0 / H->tau, c2 = 0.0;
        for (int i = 0; i < n; i++) {
            const double h = H->h1[i] + H->shift;
            res[i] = x[i] / h;
            c1 += 1.0 / h;
            c2 += res[i];
        }
        for (int i = 0; i < n; i++) res[i] -= (c2 / c1) / (H->h1[i] + H->shift);
    } else {
        for (int i = 0; i < n; i++) res[i] = x[i] / (H->h1[i] + H->shift);
    }
}

/* Delta x = h2 .* x + shift x - C A^-1 B x + tau (1'x) 1 */
static void qh_delta(const qrot_hess* H, const double* x, double* res)
{
    qh_Bx(H, x, H->tn);
    qh_solve_A(H, H->tn, H->tn2);
    qh_Cx(H, H->tn2, H->tm);
    const double c = H->tau > 0.0 ? H->tau * q_sum(x, H->m) : 0.0;
    for (int j = 0; j < H->m; j++) res[j] = H->h2[j] * x[j] + H->shift * x[j] - H->tm[j] + c;
}

/* Conjugate gradient with an identity preconditioner from x = guess (or 0),
 * stopping at ||r||^2 < max(tol^2 ||rhs||^2, DBL_MIN) or after max_iter steps
 * (the rules of Eigen::ConjugateGradient that src/qrot_cg.cpp:31-38 calls;
 * max_iter defaults to 2m there; a zero rhs gives x = 0).  Returns the number
 * of steps. */
static int qh_cg(qrot_hess* H, const double* rhs, double tol, int max_iter, const double* guess,
                 double* x)
{
    const int m = H->m;
    double* r = (double*)malloc(sizeof(double) * (size_t)m);
    double* p = (double*)malloc(sizeof(double) * (size_t)m);
    double* q = (double*)malloc(sizeof(double) * (size_t)m);
    double rhs2 = 0.0;
    for (int j = 0; j < m; j++) rhs2 += rhs[j] * rhs[j];
    if (guess) {
        qh_delta(H, guess, q);
        for (int j = 0; j < m; j++) {
            x[j] = guess[j];
            r[j] = rhs[j] - q[j];
        }
    } else {
        for (int j = 0; j < m; j++) {
            x[j] = 0.0;
            r[j] = rhs[j];
        }
    }
    int it = 0;
    if (rhs2 == 0.0) {
        for (int j = 0; j < m; j++) x[j] = 0.0;
    } else {
        double thr = tol * tol * rhs2;
        if (thr < 2.2250738585072014e-308) thr = 2.2250738585072014e-308;
        double rr = 0.0;
        for (int j = 0; j < m; j++) {
            rr += r[j] * r[j];
            p[j] = r[j];
        }
        if (rr >= thr) {
            double abs_new = rr;
            while (it < max_iter) {
                qh_delta(H, p, q);
                double pq = 0.0;
                for (int j = 0; j < m; j++) pq += p[j] * q[j];
                const double al = abs_new / pq;
                rr = 0.0;
                for (int j = 0; j < m; j++) {
                    x[j] += al * p[j];
                    r[j] -= al * q[j];
                    rr += r[j] * r[j];
                }
                if (rr < thr) break;
                const double bt = rr / abs_new;
                abs_new = rr;
                for (int j = 0; j < m; j++) p[j] = r[j] + bt * p[j];
                it++;
            }
        }
    }
    free(r);
    free(p);
    free(q);
    return it;
}

/* (H + shift I + tau vv')^-1 rhs by block elimination (src/qrot_cg.cpp:9-51):
 * v = A^-1 w, r2 = Delta^-1 (z - C v) by CG, r1 = A^-1 (w - B r2). */
static int qh_solve(qrot_hess* H, const double* rhs, double tol, int max_iter,
                    const double* guess, double* out)
{
    const int n = H->n, m = H->m;
    double* v = (double*)malloc(sizeof(double) * (size_t)n);
    double* zy = (double*)malloc(sizeof(double) * (size_t)m);
    double* w = (double*)malloc(sizeof(double) * (size_t)n);
    qh_solve_A(H, rhs, v);
    qh_Cx(H, v, zy);
    for (int j = 0; j < m; j++) zy[j] = rhs[n + j] - zy[j];
    const int it = qh_cg(H, zy, tol, max_iter, guess, out + n);
    qh_Bx(H, out + n, w);
    for (int i = 0; i < n; i++) w[i] = rhs[i] - w[i];
    qh_solve_A(H, w, out);
    free(v);
    free(zy);
    free(w);
    return it;
}

int orc_qrot_hess_solve(int n, int m, const double* M, const double* gamma, const double* rhs,
                        double shift, double tau, double tol, int max_iter, double* out)
{
    qrot_hess H;
    qh_build(&H, n, m, M, gamma);
    H.shift = shift;
    H.tau = tau;
    const int it = qh_solve(&H, rhs, tol, max_iter, NULL, out);
    qh_free(&H);
    return it;
}

/* GRSSN (src/qrot_grssn.cpp:16-133): direction (H + shift I + tau vv')^-1
 * (-g) by the CG above (tol 1e-8, at most 2m steps; method != 0 asks the
 * reference for a direct factorisation, an exact solve, here CG to 1e-12 with
 * at most 10m + 1000 steps), Armijo backtracking from step 1 (theta = 0.5,
 * kappa = 0.5, 20 trials, the last trial taken), histories niter + 1 long;
 * cg_iters[k] receives the CG steps of iteration k. */
int orc_qrot_grssn(int n, int m, const double* M, const double* a, const double* b, double reg,
                   double tau, double shift, int method, double tol, int max_iter, int verbose,
                   int hist_cap, const double* x0, double* x, double* obj_vals,
                   double* prim_vals, double* mar_errs, double* run_times, int* hist_len,
                   int* cg_iters)
{
    const long N = (long)n + m;
    qrot_ctx Q = {n, m, M, a, b, reg, tau, 0, hist_cap, verbose, 0.0,
                  obj_vals, prim_vals, mar_errs, run_times};
    if (x0) {
        memcpy(x, x0, sizeof(double) * N);
    } else {
        memset(x, 0, sizeof(double) * n);
        orc_qrot_optimal_beta(n, m, M, b, reg, x, x + n);
    }
    const double cg_tol = method == 0 ? 1e-8 : 1e-12;
    const int cg_max = method == 0 ? 2 * m : 10 * m + 1000;
    double* grad = (double*)malloc(sizeof(double) * N);
    double* rhs = (double*)malloc(sizeof(double) * N);
    double* dir = (double*)malloc(sizeof(double) * N);
    double* xn = (double*)malloc(sizeof(double) * N);
    Q.t0 = q_now_ms();
    double f = qrot_objgrad(&Q, x, grad);
    double gnorm = q_norm2(grad, N);
    qrot_record(&Q, x, f, gnorm);
    int it;
    for (it = 0; it < max_iter; it++) {
        if (verbose >= 1) printf("iter = %d, objval = %g, ||grad|| = %g\n", it, f, gnorm);
        if (gnorm < tol) break;
        qrot_hess H;
        qh_build(&H, n, m, M, x);
        H.shift = shift;
        H.tau = tau;
        for (long q = 0; q < N; q++) rhs[q] = -grad[q];
        const int cgi = qh_solve(&H, rhs, cg_tol, cg_max, NULL, dir);
        qh_free(&H);
        if (cg_iters && it < hist_cap) cg_iters[it] = cgi;
        double gd = 0.0;
        for (long q = 0; q < N; q++) gd += grad[q] * dir[q];
        const double thresh = 0.5 * gd;
        double step = 1.0;
        for (int k = 0; k < 20; k++) {
            for (long q = 0; q < N; q++) xn[q] = x[q] + step * dir[q];
            const double fn = qrot_objgrad(&Q, xn, NULL);
            if (fn <= f + step * thresh) break;
            step *= 0.5;
        }
        memcpy(x, xn, sizeof(double) * N);
        f = qrot_objgrad(&Q, x, grad);
        gnorm = q_norm2(grad, N);
        qrot_record(&Q, x, f, gnorm);
    }
    free(grad);
    free(rhs);
    free(dir);
    free(xn);
    *hist_len = Q.h;
    return it;
}

/* ASSN (src/qrot_assn.cpp:16-170): Newton direction with shift mu = lam ||g||
 * (CG as in GRSSN); u = gamma + d is taken when ||g(u)|| <= nu xi (xi the
 * last accepted gradient norm, ||g_0|| at the start); otherwise rho =
 * -g(u)'d / ||d||^2 picks the projection point v = gamma + (g(u)'d /
 * ||g(u)||^2) g(u) (rho >= eta1 and ||g(v)|| <= ||g||) or a gradient step
 * gamma - 0.5 g, and lam moves to (lam_min + lam) / 2, 2 lam or 4 lam for rho
 * >= eta2, >= eta1 or below.  nu = 0.99, eta1 = 0.25, eta2 = 0.75,
 * lam_min = 0.01, lam_0 = 1. */
int orc_qrot_assn(int n, int m, const double* M, const double* a, const double* b, double reg,
                  double tau, int method, double tol, int max_iter, int verbose, int hist_cap,
                  const double* x0, double* x, double* obj_vals, double* prim_vals,
                  double* mar_errs, double* run_times, int* hist_len, int* cg_iters)
{
    const long N = (long)n + m;
    qrot_ctx Q = {n, m, M, a, b, reg, tau, 0, hist_cap, verbose, 0.0,
                  obj_vals, prim_vals, mar_errs, run_times};
    if (x0) {
        memcpy(x, x0, sizeof(double) * N);
    } else {
        memset(x, 0, sizeof(double) * n);
        orc_qrot_optimal_beta(n, m, M, b, reg, x, x + n);
    }
    const double nu = 0.99, eta1 = 0.25, eta2 = 0.75, lam_min = 0.01;
    const double cg_tol = method == 0 ? 1e-8 : 1e-12;
    const int cg_max = method == 0 ? 2 * m : 10 * m + 1000;
    double xi = 0.0, lam = 1.0;
    double* g = (double*)malloc(sizeof(double) * N);
    double* rhs = (double*)malloc(sizeof(double) * N);
    double* d = (double*)malloc(sizeof(double) * N);
    double* u = (double*)malloc(sizeof(double) * N);
    double* gu = (double*)malloc(sizeof(double) * N);
    double* v = (double*)malloc(sizeof(double) * N);
    double* gv = (double*)malloc(sizeof(double) * N);
    Q.t0 = q_now_ms();
    double f = qrot_objgrad(&Q, x, g);
    double gnorm = q_norm2(g, N);
    qrot_record(&Q, x, f, gnorm);
    int it;
    for (it = 0; it < max_iter; it++) {
        if (verbose >= 1) printf("iter = %d, objval = %g, ||grad|| = %g\n", it, f, gnorm);
        if (gnorm < tol) break;
        if (it == 0) xi = gnorm;
        qrot_hess H;
        qh_build(&H, n, m, M, x);
        H.shift = lam * gnorm;
        H.tau = tau;
        for (long q = 0; q < N; q++) rhs[q] = -g[q];
        const int cgi = qh_solve(&H, rhs, cg_tol, cg_max, NULL, d);
        qh_free(&H);
        if (cg_iters && it < hist_cap) cg_iters[it] = cgi;
        for (long q = 0; q < N; q++) u[q] = x[q] + d[q];
        qrot_objgrad(&Q, u, gu);
        const double gunorm = q_norm2(gu, N);
        if (gunorm <= nu * xi) {
            memcpy(x, u, sizeof(double) * N);
            xi = gunorm;
        } else {
            double numer = 0.0, denom = 0.0;
            for (long q = 0; q < N; q++) {
                numer -= gu[q] * d[q];
                denom += d[q] * d[q];
            }
            const double rho = numer / denom;
            const double c = (-numer) / (gunorm * gunorm);
            for (long q = 0; q < N; q++) v[q] = x[q] + c * gu[q];
            qrot_objgrad(&Q, v, gv);
            const double gvnorm = q_norm2(gv, N);
            if (rho >= eta1 && gvnorm <= gnorm) {
                memcpy(x, v, sizeof(double) * N);
            } else {
                for (long q = 0; q < N; q++) x[q] -= 0.5 * g[q];
            }
            if (rho >= eta2)
                lam = 0.5 * (lam_min + lam);
            else if (rho >= eta1)
                lam *= 2.0;
            else
                lam *= 4.0;
        }
        f = qrot_objgrad(&Q, x, g);
        gnorm = q_norm2(g, N);
        qrot_record(&Q, x, f, gnorm);
    }
    free(g);
    free(rhs);
    free(d);
    free(u);
    free(gu);
    free(v);
    free(gv);
    *hist_len = Q.h;
    return it;
}

/* (H + shift I + tau vv') x: [h1 .* w + sigma z; h2 .* z + sigma' w] + shift x
 * + tau (1'w - 1'z) v (src/qrot_hess.cpp:193-214). */
static void qh_apply_H(const qrot_hess* H, const double* x, double shift, double tau, double* res)
{
    const int n = H->n, m = H->m;
    for (int i = 0; i < n; i++) {
        double sz = 0.0;
        for (int j = 0; j < m; j++)
            if (H->sig[(long)i * m + j]) sz += x[n + j];
        res[i] = sz;
    }
    for (int j = 0; j < m; j++) res[n + j] = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++)
            if (H->sig[(long)i * m + j]) res[n + j] += x[i];
    for (int i = 0; i < n; i++) res[i] += H->h1[i] * x[i];
    for (int j = 0; j < m; j++) res[n + j] += H->h2[j] * x[n + j];
    if (shift != 0.0)
        for (long q = 0; q < (long)n + m; q++) res[q] += shift * x[q];
    if (tau > 0.0) {
        const double c = q_sum(x, n) - q_sum(x + n, m);
        for (int i = 0; i < n; i++) res[i] += tau * c;
        for (int j = 0; j < m; j++) res[n + j] -= tau * c;
    }
}

/* Problem::line_selection2 (src/qrot_problem.cpp:337-361): the first candidate
 * whose objective is below cur, else the best one; *objval its objective. */
static double q_line_selection2(qrot_ctx* Q, const double* cand, int nc, const double* x,
                                const double* d, double cur, double* objval, double* work)
{
    const long N = (long)Q->n + Q->m;
    double best = 1.0;
    *objval = 1.0 / 0.0;
    for (int k = 0; k < nc; k++) {
        for (long q = 0; q < N; q++) work[q] = x[q] + cand[k] * d[q];
        const double fk = qrot_objgrad(Q, work, NULL);
        if (fk < *objval) {
            best = cand[k];
            *objval = fk;
        }
        if (*objval < cur) return best;
    }
    return best;
}

/* S5N (src/qrot_s5n.cpp:16-196): shift = mu (primal value + f / reg)^1 (the
 * duality gap of the rounded plan), direction by the CG solve (warm-started
 * from the last direction's beta block after a rejected step), step from
 * line_selection2 over {1, 0.5, 0.1}, then {0.05, 0.01, 1e-3, 1e-4} on
 * alternate failures; rho = (f - f_new) / (-g's - s'Hs / 2) accepts the step
 * when > 0 and moves mu by x4 (rho < 0.25) or x0.5 down to 1e-3 (rho > 0.75);
 * H, f and g are recomputed only after an accepted step; every iteration
 * records f, the primal value, ||g|| / reg. */
int orc_qrot_s5n(int n, int m, const double* M, const double* a, const double* b, double reg,
                 double tau, int method, double mu0, double tol, int max_iter, int verbose,
                 int hist_cap, const double* x0, double* x, double* obj_vals, double* prim_vals,
                 double* mar_errs, double* run_times, int* hist_len, int* cg_iters)
{
    const long N = (long)n + m;
    qrot_ctx Q = {n, m, M, a, b, reg, tau, 0, hist_cap, verbose, 0.0,
                  obj_vals, prim_vals, mar_errs, run_times};
    if (x0) {
        memcpy(x, x0, sizeof(double) * N);
    } else {
        memset(x, 0, sizeof(double) * n);
        orc_qrot_optimal_beta(n, m, M, b, reg, x, x + n);
    }
    const double cg_tol = method == 0 ? 1e-8 : 1e-12;
    const int cg_max = method == 0 ? 2 * m : 10 * m + 1000;
    const double alphas[3] = {1.0, 0.5, 0.1};
    const double alphas_small[4] = {0.05, 0.01, 0.001, 1e-4};
    double mu = mu0, rho = 1.0;
    int use_small = 1;
    double* g = (double*)malloc(sizeof(double) * N);
    double* rhs = (double*)malloc(sizeof(double) * N);
    double* d = (double*)calloc((size_t)N, sizeof(double));
    double* st = (double*)malloc(sizeof(double) * N);
    double* hs = (double*)malloc(sizeof(double) * N);
    double* work = (double*)malloc(sizeof(double) * N);
    Q.t0 = q_now_ms();
    double f = qrot_objgrad(&Q, x, g);
    double gnorm = q_norm2(g, N);
    qrot_hess H;
    qh_build(&H, n, m, M, x);
    H.tau = tau;
    qrot_record(&Q, x, f, gnorm);
    double prim = Q.prim_vals[Q.h - 1];
    int it;
    for (it = 0; it < max_iter; it++) {
        const double dgap = prim + f / reg;
        if (verbose >= 1)
            printf("iter = %d, objval = %g, ||grad|| = %g, ldgap = %g\n", it, f, gnorm,
                   log10(dgap));
        if (gnorm < tol) break;
        H.shift = mu * pow(dgap, 1.0);
        for (long q = 0; q < N; q++) rhs[q] = -g[q];
        const int cgi = qh_solve(&H, rhs, cg_tol, cg_max, rho <= 0.0 ? d + n : NULL, d);
        if (cg_iters && it < hist_cap) cg_iters[it] = cgi;
        double newf;
        double alpha = q_line_selection2(&Q, alphas, 3, x, d, f, &newf, work);
        if (newf >= f) {
            if (use_small) alpha = q_line_selection2(&Q, alphas_small, 4, x, d, f, &newf, work);
            use_small = !use_small;
        } else {
            use_small = 1;
        }
        for (long q = 0; q < N; q++) st[q] = alpha * d[q];
        qh_apply_H(&H, st, 0.0, tau, hs);
        double gs = 0.0, shs = 0.0;
        for (long q = 0; q < N; q++) {
            gs += g[q] * st[q];
            shs += st[q] * hs[q];
        }
        rho = (f - newf) / (-gs - 0.5 * shs);
        if (rho > 0.0)
            for (long q = 0; q < N; q++) x[q] += st[q];
        if (rho < 0.25)
            mu *= 4.0;
        else if (rho > 0.75)
            mu = fmax(0.5 * mu, 0.001);
        if (rho > 0.0) {
            f = qrot_objgrad(&Q, x, g);
            gnorm = q_norm2(g, N);
            qh_free(&H);
            qh_build(&H, n, m, M, x);
            H.tau = tau;
        }
        qrot_record(&Q, x, f, gnorm);
        prim = Q.prim_vals[Q.h - 1 < hist_cap ? Q.h - 1 : hist_cap - 1];
    }
    qh_free(&H);
    free(g);
    free(rhs);
    free(d);
    free(st);
    free(hs);
    free(work);
    *hist_len = Q.h;
    return it;
}
