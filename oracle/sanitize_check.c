/* sanitize_check.c -- runs the oracle's entry points on small problems in a
 * build with AddressSanitizer and UndefinedBehaviorSanitizer (test
 * infrastructure only: `make -C oracle sanitize`, run by the CPU tests).
 * A memory error or undefined behaviour aborts with the sanitizer's report;
 * a non-finite result is counted and fails the run. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "erot_oracle.h"

static int fails = 0;

static void expect_finite(const char* what, const double* x, long len)
{
    for (long i = 0; i < len; i++)
        if (!isfinite(x[i])) {
            printf("non-finite %s[%ld]\n", what, i);
            fails++;
            return;
        }
}

/* The reference README's recipe: 1-D grids, squared distance, an exponential
 * and a Gaussian mixture as marginals (normalised). */
static void problem(int n, int m, double* M, double* a, double* b)
{
    double sa = 0.0, sb = 0.0;
    for (int i = 0; i < n; i++) {
        const double x = n > 1 ? 5.0 * i / (n - 1) : 0.0;
        a[i] = exp(-x);
        sa += a[i];
        for (int j = 0; j < m; j++) {
            const double y = m > 1 ? 5.0 * j / (m - 1) : 0.0;
            M[(long)i * m + j] = (x - y) * (x - y);
        }
    }
    for (int j = 0; j < m; j++) {
        const double y = m > 1 ? 5.0 * j / (m - 1) : 0.0;
        b[j] = 0.2 * exp(-0.5 * (y - 1.0) * (y - 1.0) / 0.04) +
               0.8 * exp(-0.5 * (y - 3.0) * (y - 3.0) / 0.25);
        sb += b[j];
    }
    for (int i = 0; i < n; i++) a[i] /= sa;
    for (int j = 0; j < m; j++) b[j] /= sb;
}

static void run_shape(int n, int m)
{
    const int it = 60, N = n + m - 1;
    const double reg = 0.1, tol = 1e-9;
    double* M = malloc(sizeof(double) * (size_t)n * m);
    double* T = malloc(sizeof(double) * (size_t)n * m);
    double* plan = malloc(sizeof(double) * (size_t)n * m);
    double* a = malloc(sizeof(double) * n);
    double* b = malloc(sizeof(double) * m);
    double* dual = malloc(sizeof(double) * (N + m + 1));
    double* grad = malloc(sizeof(double) * (N + m + 1));
    double* h[4];
    for (int q = 0; q < 4; q++) h[q] = malloc(sizeof(double) * (it + 1));
    int hl = 0;
    problem(n, m, M, a, b);

    orc_opts o;
    memset(&o, 0, sizeof o);
    o.lin_method = 1;
    o.mu0 = 1.0;
    o.shift = 1e-3;
    o.density = 0.1;
    o.dense_T = 1;

    static double trace[1 << 16];
    orc_trace_set(trace, sizeof trace / sizeof trace[0]);

    orc_sinkhorn_bcd(n, m, M, a, b, reg, tol, it, 0, dual, h[0], h[1], h[2], &hl);
    expect_finite("bcd dual", dual, N);
    orc_sinkhorn_ssns(n, m, M, a, b, reg, tol, it, 0, &o, dual, h[0], h[1], h[2], h[3], &hl);
    expect_finite("ssns dual", dual, N);
    o.lin_method = 0;
    orc_sinkhorn_ssns(n, m, M, a, b, reg, tol, it, 0, &o, dual, h[0], h[1], h[2], h[3], &hl);
    expect_finite("ssns (CG) dual", dual, N);
    orc_sinkhorn_sparse_newton(n, m, M, a, b, reg, tol, it, 0, &o, dual, h[0], h[1], h[2], h[3],
                               &hl);
    expect_finite("sparse Newton dual", dual, N);
    orc_sinkhorn_splr(n, m, M, a, b, reg, tol, it, 0, &o, dual, h[0], h[1], h[2], h[3], &hl);
    expect_finite("SPLR dual", dual, N);
    orc_sinkhorn_newton(n, m, M, a, b, reg, tol, it, 0, dual, h[0], h[1], h[2], &hl);
    expect_finite("Newton dual", dual, N);
    orc_sinkhorn_lbfgs_dual(n, m, M, a, b, reg, tol, it, 0, it + 1, dual, h[0], h[1], h[2], &hl);
    expect_finite("L-BFGS dual", dual, N);
    orc_sinkhorn_apdagd(n, m, M, a, b, reg, tol, it, 0, dual, plan, h[0], h[1], h[2], &hl);
    expect_finite("APDAGD plan", plan, (long)n * m);
    orc_trace_set(NULL, 0);

    /* passes, sparsifiers and Hessian operators at the BCD point */
    orc_sinkhorn_bcd(n, m, M, a, b, reg, 1e-6, it, 0, dual, h[0], h[1], h[2], &hl);
    double f = orc_dual_obj_grad(n, m, M, a, b, reg, dual, grad, T, 1);
    expect_finite("objective", &f, 1);
    expect_finite("gradient", grad, N);
    orc_get_plan(n, m, M, reg, dual, plan);
    orc_approx_proj(n, m, plan, a, b, T);
    expect_finite("rounded plan", T, (long)n * m);
    /* the entropy term is P (1 - log P) summed as in the reference: NaN
     * exactly when the rounded plan has an entry <= 0 (0 * inf, or the log of
     * an entry that the rank-one correction's rounding left a few ulps below
     * zero) */
    f = orc_primal_val(n, m, M, a, b, reg, dual);
    int nonpos = 0;
    for (long q = 0; q < (long)n * m; q++) nonpos |= T[q] <= 0.0;
    if (nonpos ? !isnan(f) : !isfinite(f)) {
        printf("primal value %g with%s entries <= 0 in the rounded plan\n", f,
               nonpos ? "" : "out");
        fails++;
    }
    orc_compute_T(n, m, M, reg, dual, T);
    int* colptr = malloc(sizeof(int) * m);
    for (int rule = 0; rule < 2; rule++) {
        int* rowind = NULL;
        double* val = NULL;
        if (rule == 0)
            orc_sparsify_mat3(n, m, T, 1e-3, 0.5, colptr, &rowind, &val);
        else
            orc_sparsify_density(n, m, T, 0.1, 1, colptr, &rowind, &val);
        double* h1 = malloc(sizeof(double) * n);
        double* h2 = malloc(sizeof(double) * m);
        for (int i = 0; i < n; i++) {
            h1[i] = 0.0;
            for (int j = 0; j < m; j++) h1[i] += T[(long)i * m + j];
        }
        for (int j = 0; j < m - 1; j++) {
            h2[j] = 0.0;
            for (int i = 0; i < n; i++) h2[j] += T[(long)i * m + j];
        }
        double terr = 0.0;
        orc_apply_hsx(n, m, h1, h2, colptr, rowind, val, grad, dual);
        expect_finite("Hs x", dual, N);
        orc_hess_cg(n, m, h1, h2, colptr, rowind, val, grad, 1e-3, 1e-10, 1000, dual, &terr);
        expect_finite("CG solve", dual, N);
        orc_hess_direct(n, m, h1, h2, colptr, rowind, val, grad, 1e-3, dual);
        expect_finite("direct solve", dual, N);
        free(h1);
        free(h2);
        orc_free(rowind);
        orc_free(val);
    }
    free(colptr);

    /* QROT: gamma = [alpha; beta] */
    double* x = malloc(sizeof(double) * (n + m));
    int* cg = malloc(sizeof(int) * (it + 1));   /* CG steps per Newton iteration */
    for (int solver = 0; solver <= 4; solver++) {
        orc_qrot_solve(solver, n, m, M, a, b, reg, 1e-3, 1e-8, it, 0, it + 1, NULL, x, h[0],
                       h[1], h[2], h[3], &hl, plan);
        expect_finite("QROT solve", x, n + m);
    }
    orc_qrot_grssn(n, m, M, a, b, reg, 1e-3, 1e-6, 0, 1e-8, it, 0, it + 1, NULL, x, h[0], h[1],
                   h[2], h[3], &hl, cg);
    expect_finite("QROT GRSSN", x, n + m);
    orc_qrot_assn(n, m, M, a, b, reg, 1e-3, 0, 1e-8, it, 0, it + 1, NULL, x, h[0], h[1], h[2],
                  h[3], &hl, cg);
    expect_finite("QROT ASSN", x, n + m);
    orc_qrot_s5n(n, m, M, a, b, reg, 1e-3, 0, 1.0, 1e-8, it, 0, it + 1, NULL, x, h[0], h[1], h[2],
                 h[3], &hl, cg);
    expect_finite("QROT S5N", x, n + m);
    free(x);
    free(cg);

    free(M);
    free(T);
    free(plan);
    free(a);
    free(b);
    free(dual);
    free(grad);
    for (int q = 0; q < 4; q++) free(h[q]);
}

int main(void)
{
    const int shapes[][2] = {{23, 17}, {12, 40}, {2, 3}};
    for (unsigned s = 0; s < sizeof shapes / sizeof shapes[0]; s++) {
        printf("shape %d x %d\n", shapes[s][0], shapes[s][1]);
        fflush(stdout);
        run_shape(shapes[s][0], shapes[s][1]);
    }
    printf("%s (%d non-finite results)\n", fails ? "FAILED" : "ok", fails);
    return fails ? 1 : 0;
}
