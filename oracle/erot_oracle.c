/*
 * erot_oracle.c -- FP64 CPU restatement of the EROT path of
 * yixuan/regot-python v0.0.3.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP build.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Storage is row-major (C order): M[i*m + j].  Each function cites the
 * reference file:line whose semantics it restates.  The code is written from
 * the reference's documented behaviour, not transcribed: the reference uses
 * Eigen packets and x86-simd-sort; this file uses plain loops and qsort.
 */
#include "erot_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

static double dot(const double* x, const double* y, long n)
{
    double s = 0.0;
    for (long i = 0; i < n; i++) s += x[i] * y[i];
    return s;
}

static double norm2(const double* x, long n) { return sqrt(dot(x, x, n)); }

void orc_free(void* p) { free(p); }

/* ------------------------------------------------------------------------ */
/* Dense passes                                                              */
/* ------------------------------------------------------------------------ */

/* beta_j = reg * (log b_j - LSE_i((alpha_i - M_ij)/reg)), max-shifted LSE.
 * Restates Problem::optimal_beta, src/sinkhorn_problem_bcd.cpp:149-174, and
 * log_sum_exp_colwise, :111-146 (column max, then sum of exp(D - max)). */
void orc_optimal_beta(int n, int m, const double* M, const double* b, double reg,
                      const double* alpha, double* beta)
{
    double* cmax = (double*)malloc(sizeof(double) * m);
    double* csum = (double*)calloc(m, sizeof(double));
    for (int j = 0; j < m; j++) cmax[j] = -INFINITY;
    for (int i = 0; i < n; i++) {
        const double* Mi = M + (long)i * m;
        for (int j = 0; j < m; j++) {
            double d = (alpha[i] - Mi[j]) / reg;
            if (d > cmax[j]) cmax[j] = d;
        }
    }
    for (int i = 0; i < n; i++) {
        const double* Mi = M + (long)i * m;
        for (int j = 0; j < m; j++) csum[j] += exp((alpha[i] - Mi[j]) / reg - cmax[j]);
    }
    for (int j = 0; j < m; j++) beta[j] = reg * (log(b[j]) - (cmax[j] + log(csum[j])));
    free(cmax);
    free(csum);
}

/* alpha_i = reg * (log a_i - LSE_j((beta_j - M_ij)/reg)).
 * Restates Problem::optimal_alpha, src/sinkhorn_problem_bcd.cpp:177-201, and
 * log_sum_exp_rowwise, :82-108. */
void orc_optimal_alpha(int n, int m, const double* M, const double* a, double reg,
                       const double* beta, double* alpha)
{
    for (int i = 0; i < n; i++) {
        const double* Mi = M + (long)i * m;
        double c = -INFINITY;
        for (int j = 0; j < m; j++) {
            double d = (beta[j] - Mi[j]) / reg;
            if (d > c) c = d;
        }
        double s = 0.0;
        for (int j = 0; j < m; j++) s += exp((beta[j] - Mi[j]) / reg - c);
        alpha[i] = reg * (log(a[i]) - (c + log(s)));
    }
}

/* T_ij = exp((alpha_i + beta_j - M_ij)/reg) with gamma = [alpha; beta_0..m-2]
 * and beta_{m-1} = 0; returns sum(T).  T may be NULL (sum only).
 * Restates Problem::compute_T, src/sinkhorn_problem_obj.cpp:139-188. */
double orc_compute_T(int n, int m, const double* M, double reg, const double* gamma, double* T)
{
    const double* alpha = gamma;
    const double* beta = gamma + n;
    double total = 0.0;
    for (int i = 0; i < n; i++) {
        const double* Mi = M + (long)i * m;
        double* Ti = T ? T + (long)i * m : NULL;
        for (int j = 0; j < m; j++) {
            double bj = (j < m - 1) ? beta[j] : 0.0;
            double t = exp((alpha[i] + bj - Mi[j]) / reg);
            if (Ti) Ti[j] = t;
            total += t;
        }
    }
    return total;
}

/* Tsums[0:n] = row sums, Tsums[n:n+m-1] = column sums without the last column.
 * Restates Problem::compute_sums, src/sinkhorn_problem_obj.cpp:192-297. */
void orc_compute_sums(int n, int m, const double* T, double* Tsums)
{
    double* cs = Tsums + n;
    for (int j = 0; j < m - 1; j++) cs[j] = 0.0;
    for (int i = 0; i < n; i++) {
        const double* Ti = T + (long)i * m;
        double r = 0.0;
        for (int j = 0; j < m; j++) {
            r += Ti[j];
            if (j < m - 1) cs[j] += Ti[j];
        }
        Tsums[i] = r;
    }
}

/* f = reg * sum(T) - alpha'a - betat'bt.
 * Restates Problem::dual_obj, src/sinkhorn_problem_obj.cpp:324-342. */
double orc_dual_obj(int n, int m, const double* M, const double* a, const double* b,
                    double reg, const double* gamma)
{
    double Tsum = orc_compute_T(n, m, M, reg, gamma, NULL);
    return reg * Tsum - dot(gamma, a, n) - dot(gamma + n, b, m - 1);
}

/* f and g = [T 1_m - a ; (T' 1_n)[0:m-1] - b[0:m-1]], sum(T) taken as the sum of
 * the row sums.  T (n x m) is filled when computeT, else read.  If T is NULL
 * the pass streams without storing T.
 * Restates Problem::dual_obj_grad, src/sinkhorn_problem_obj.cpp:363-393. */
double orc_dual_obj_grad(int n, int m, const double* M, const double* a, const double* b,
                         double reg, const double* gamma, double* grad, double* T, int computeT)
{
    if (T) {
        if (computeT) orc_compute_T(n, m, M, reg, gamma, T);
        orc_compute_sums(n, m, T, grad);
    } else {
        const double* alpha = gamma;
        const double* beta = gamma + n;
        double* cs = grad + n;
        for (int j = 0; j < m - 1; j++) cs[j] = 0.0;
        for (int i = 0; i < n; i++) {
            const double* Mi = M + (long)i * m;
            double r = 0.0;
            for (int j = 0; j < m; j++) {
                double bj = (j < m - 1) ? beta[j] : 0.0;
                double t = exp((alpha[i] + bj - Mi[j]) / reg);
                r += t;
                if (j < m - 1) cs[j] += t;
            }
            grad[i] = r;
        }
    }
    double Tsum = 0.0;
    for (int i = 0; i < n; i++) Tsum += grad[i];
    for (int i = 0; i < n; i++) grad[i] -= a[i];
    for (int j = 0; j < m - 1; j++) grad[n + j] -= b[j];
    return reg * Tsum - dot(gamma, a, n) - dot(gamma + n, b, m - 1);
}

/* plan = exp((alpha (+) beta - M)/reg), beta_{m-1} = 0, row-major.
 * Restates SinkhornResult::get_plan, src/sinkhorn_result.h:32-50. */
void orc_get_plan(int n, int m, const double* M, double reg, const double* gamma, double* plan)
{
    orc_compute_T(n, m, M, reg, gamma, plan);
}

/* ------------------------------------------------------------------------ */
/* Sparsifiers                                                               */
/* ------------------------------------------------------------------------ */

static int cmp_double(const void* x, const void* y)
{
    double a = *(const double*)x, b = *(const double*)y;
    return (a < b) ? -1 : (a > b) ? 1 : 0;
}

/* Position L in the ascending order x_(0) <= ... <= x_(k-1) where the
 * cumulative mass first exceeds delta: x_(0)+...+x_(L) > delta; k if none.
 * Semantics of find_small, src/sinkhorn_sparsify3.cpp:104-148. */
static int small_mass_position(const double* sorted, int k, double delta)
{
    double s = 0.0;
    for (int l = 0; l < k; l++) {
        s += sorted[l];
        if (s > delta) return l;
    }
    return k;
}

/* Largest L with x_(L)+...+x_(k-1) >= target, scanning from the largest
 * value down; 0 if the suffix sums never reach target.
 * Semantics of find_large / select_and_find_reverse,
 * src/sinkhorn_sparsify3.cpp:66-102,150-196. */
static int large_mass_position(const double* sorted, int k, double target)
{
    double s = 0.0;
    for (int l = k - 1; l >= 0; l--) {
        s += sorted[l];
        if (s >= target) return l;
    }
    return 0;
}

/* SSNS threshold rule.  Per column j < m-1: threshold = the value at the
 * position where the ascending cumulative mass passes delta (find_small if
 * density_hint > 0.2, else find_large on colsum - delta); entries >= threshold
 * are kept.  Per row: if the dropped mass exceeds 2 delta, the largest dropped
 * entries are re-kept until the dropped mass falls to 2 delta (find_large on
 * rowsum - 2 delta).  An entry is in the output iff its Delta slot is exactly
 * zero, so exact zeros of T are stored too.
 * Restates sparsify_mat3, src/sinkhorn_sparsify3.cpp:276-405. */
long long orc_sparsify_mat3(int n, int m, const double* T, double delta, double density_hint,
                            int* colptr, int** rowind_out, double** val_out)
{
    const int m1 = m - 1;
    double* D = (double*)malloc(sizeof(double) * (size_t)n * m);
    memcpy(D, T, sizeof(double) * (size_t)n * m);
    for (int i = 0; i < n; i++) D[(long)i * m + m - 1] = 0.0;

    double* buf = (double*)malloc(sizeof(double) * (size_t)(n > m ? n : m));
    /* column pass */
    for (int j = 0; j < m1; j++) {
        double colsum = 0.0;
        for (int i = 0; i < n; i++) {
            buf[i] = T[(long)i * m + j];
            colsum += buf[i];
        }
        qsort(buf, n, sizeof(double), cmp_double);
        int L;
        if (density_hint > 0.2)
            L = small_mass_position(buf, n, delta);
        else
            L = large_mass_position(buf, n, colsum - delta);
        double thresh = (L >= n) ? INFINITY : buf[L];
        for (int i = 0; i < n; i++) {
            double t = T[(long)i * m + j];
            D[(long)i * m + j] = (t >= thresh) ? 0.0 : t;
        }
    }
    /* row pass */
    for (int i = 0; i < n; i++) {
        double* Di = D + (long)i * m;
        double rowsum = 0.0;
        for (int j = 0; j < m; j++) rowsum += Di[j];
        if (rowsum > 2.0 * delta) {
            memcpy(buf, Di, sizeof(double) * m1);
            qsort(buf, m1, sizeof(double), cmp_double);
            int L = large_mass_position(buf, m1, rowsum - 2.0 * delta);
            double thresh = (L >= m1) ? INFINITY : buf[L];
            for (int j = 0; j < m1; j++)
                if (Di[j] >= thresh) Di[j] = 0.0;
        }
    }
    free(buf);
    /* CSC build: kept <=> Delta == 0 */
    long long nnz = 0;
    for (int j = 0; j < m1; j++)
        for (int i = 0; i < n; i++)
            if (D[(long)i * m + j] == 0.0) nnz++;
    int* rowind = (int*)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    double* val = (double*)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
    long long k = 0;
    colptr[0] = 0;
    for (int j = 0; j < m1; j++) {
        for (int i = 0; i < n; i++) {
            if (D[(long)i * m + j] == 0.0) {
                rowind[k] = i;
                val[k] = T[(long)i * m + j];
                k++;
            }
        }
        colptr[j + 1] = (int)k;
    }
    free(D);
    *rowind_out = rowind;
    *val_out = val;
    return nnz;
}

typedef struct {
    double v;
    long long l;
} vl_pair;

static int cmp_topk(const void* x, const void* y)
{
    const vl_pair* a = (const vl_pair*)x;
    const vl_pair* b = (const vl_pair*)y;
    if (a->v > b->v) return -1;
    if (a->v < b->v) return 1;
    return (a->l < b->l) ? -1 : (a->l > b->l) ? 1 : 0;
}

/* Global top-k of T[:, 0:m-1], k = floor(n(m-1) * clamp(density, 0, 1)).
 * With preserve, column 0 and row 0 are always kept and top-k entries on
 * them are skipped so nothing is stored twice.  Ties at the k-th value are
 * broken towards the smaller column-major index i + j*n (the reference's
 * avx2_argselect leaves this order unspecified).
 * Restates sparsify_mat_with_density, src/sinkhorn_sparsify4.cpp:77-174. */
long long orc_sparsify_density(int n, int m, const double* T, double density, int preserve,
                               int* colptr, int** rowind_out, double** val_out)
{
    const int m1 = m - 1;
    const long long len = (long long)n * m1;
    if (density > 1.0) density = 1.0;
    if (density < 0.0) density = 0.0;
    long long k = (long long)(len * density);
    unsigned char* keep = (unsigned char*)calloc((size_t)(len > 0 ? len : 1), 1);
    if (k > 0) {
        vl_pair* p = (vl_pair*)malloc(sizeof(vl_pair) * len);
        for (int j = 0; j < m1; j++)
            for (int i = 0; i < n; i++) {
                long long l = i + (long long)j * n;
                p[l].v = T[(long)i * m + j];
                p[l].l = l;
            }
        qsort(p, len, sizeof(vl_pair), cmp_topk);
        for (long long q = 0; q < k; q++) keep[p[q].l] = 1;
        free(p);
    }
    if (preserve && m1 > 0) {
        for (int i = 0; i < n; i++) keep[i] = 1;                       /* column 0 */
        for (int j = 1; j < m1; j++) keep[(long long)j * n] = 1;       /* row 0 */
    }
    long long nnz = 0;
    for (long long l = 0; l < len; l++) nnz += keep[l];
    int* rowind = (int*)malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    double* val = (double*)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
    long long q = 0;
    colptr[0] = 0;
    for (int j = 0; j < m1; j++) {
        for (int i = 0; i < n; i++) {
            if (keep[i + (long long)j * n]) {
                rowind[q] = i;
                val[q] = T[(long)i * m + j];
                q++;
            }
        }
        colptr[j + 1] = (int)q;
    }
    free(keep);
    *rowind_out = rowind;
    *val_out = val;
    return nnz;
}

/* ------------------------------------------------------------------------ */
/* Hessian operators on CSC sigma (n x (m-1)), src/sinkhorn_hess.cpp:99-164  */
/* ------------------------------------------------------------------------ */

/* res = sigma x, x (m-1) -> res (n); src/sinkhorn_hess.cpp:99-104 */
void orc_apply_sigmadx(int n, int m, const int* colptr, const int* rowind, const double* val,
                       const double* x, double* res)
{
    for (int i = 0; i < n; i++) res[i] = 0.0;
    for (int j = 0; j < m - 1; j++)
        for (int k = colptr[j]; k < colptr[j + 1]; k++) res[rowind[k]] += val[k] * x[j];
}

/* res = sigma' x, x (n) -> res (m-1); src/sinkhorn_hess.cpp:107-112 */
void orc_apply_sigmadtx(int n, int m, const int* colptr, const int* rowind, const double* val,
                        const double* x, double* res)
{
    (void)n;
    for (int j = 0; j < m - 1; j++) {
        double s = 0.0;
        for (int k = colptr[j]; k < colptr[j + 1]; k++) s += val[k] * x[rowind[k]];
        res[j] = s;
    }
}

/* Delta x = h2 .* x + shift x - sigma' ((sigma x) ./ (h1 + shift));
 * src/sinkhorn_hess.cpp:136-148 */
void orc_apply_deltax(int n, int m, const double* h1, const double* h2,
                      const int* colptr, const int* rowind, const double* val,
                      const double* x, double shift, double* res)
{
    double* t = (double*)malloc(sizeof(double) * n);
    double* u = (double*)malloc(sizeof(double) * (m > 1 ? m - 1 : 1));
    orc_apply_sigmadx(n, m, colptr, rowind, val, x, t);
    for (int i = 0; i < n; i++) t[i] /= (h1[i] + shift);
    orc_apply_sigmadtx(n, m, colptr, rowind, val, t, u);
    for (int j = 0; j < m - 1; j++) res[j] = h2[j] * x[j] + shift * x[j] - u[j];
    free(t);
    free(u);
}

/* Hs x = [h1 .* w + sigma z ; h2 .* z + sigma' w], x = [w; z];
 * src/sinkhorn_hess.cpp:151-164 */
void orc_apply_hsx(int n, int m, const double* h1, const double* h2,
                   const int* colptr, const int* rowind, const double* val,
                   const double* x, double* res)
{
    orc_apply_sigmadx(n, m, colptr, rowind, val, x + n, res);
    orc_apply_sigmadtx(n, m, colptr, rowind, val, x, res + n);
    for (int i = 0; i < n; i++) res[i] += h1[i] * x[i];
    for (int j = 0; j < m - 1; j++) res[n + j] += h2[j] * x[n + j];
}

/* Unpreconditioned conjugate gradient on Delta (size m-1) started from zero,
 * stopping when ||r||^2 < max(tol^2 ||b||^2, DBL_MIN) or after max_iter
 * iterations; the textbook algorithm as Eigen's ConjugateGradient with an
 * IdentityPreconditioner runs it (the solver hess_cg calls).  Returns the
 * iteration count. */
static int cg_delta(int n, int m, const double* h1, const double* h2,
                    const int* colptr, const int* rowind, const double* val,
                    const double* rhs, double shift, double tol, int max_iter,
                    double* x, double* tol_error)
{
    const int d = m - 1;
    for (int j = 0; j < d; j++) x[j] = 0.0;
    double rhs_norm2 = dot(rhs, rhs, d);
    if (rhs_norm2 == 0.0) {
        if (tol_error) *tol_error = 0.0;
        return 0;
    }
    double threshold = tol * tol * rhs_norm2;
    if (threshold < 2.2250738585072014e-308) threshold = 2.2250738585072014e-308;
    double* r = (double*)malloc(sizeof(double) * d);
    double* p = (double*)malloc(sizeof(double) * d);
    double* q = (double*)malloc(sizeof(double) * d);
    memcpy(r, rhs, sizeof(double) * d);
    double rr = rhs_norm2;
    int it = 0;
    if (rr >= threshold) {
        memcpy(p, r, sizeof(double) * d);
        double abs_new = dot(r, p, d);
        while (it < max_iter) {
            orc_apply_deltax(n, m, h1, h2, colptr, rowind, val, p, shift, q);
            double alpha = abs_new / dot(p, q, d);
            for (int j = 0; j < d; j++) x[j] += alpha * p[j];
            for (int j = 0; j < d; j++) r[j] -= alpha * q[j];
            rr = dot(r, r, d);
            if (rr < threshold) break;
            double abs_old = abs_new;
            abs_new = dot(r, r, d);
            double beta = abs_new / abs_old;
            for (int j = 0; j < d; j++) p[j] = r[j] + beta * p[j];
            it++;
        }
    }
    if (tol_error) *tol_error = sqrt(rr / rhs_norm2);
    free(r);
    free(p);
    free(q);
    return it;
}

/* (Hs + shift I)^{-1} rhs by block elimination with CG on the Schur
 * complement Delta.  Restates hess_cg, src/sinkhorn_cg.cpp:9-50.
 * max_iter <= 0 uses 2m, the count Eigen derives from HessianCG::cols()
 * (src/sinkhorn_cg.h:53-54 returns m for the (m-1)-dimensional system). */
int orc_hess_cg(int n, int m, const double* h1, const double* h2,
                const int* colptr, const int* rowind, const double* val,
                const double* rhs, double shift, double tol, int max_iter,
                double* res, double* tol_error)
{
    const int d = m - 1;
    if (max_iter <= 0) max_iter = 2 * m;
    double* v = (double*)calloc(n > 0 ? n : 1, sizeof(double));
    double* y = (double*)malloc(sizeof(double) * (d > 0 ? d : 1));
    double* zy = (double*)malloc(sizeof(double) * (d > 0 ? d : 1));
    for (int i = 0; i < n; i++) v[i] = rhs[i] / (h1[i] + shift);
    orc_apply_sigmadtx(n, m, colptr, rowind, val, v, y);
    for (int j = 0; j < d; j++) zy[j] = rhs[n + j] - y[j];
    int it = cg_delta(n, m, h1, h2, colptr, rowind, val, zy, shift, tol, max_iter, res + n,
                      tol_error);
    orc_apply_sigmadx(n, m, colptr, rowind, val, res + n, v);
    for (int i = 0; i < n; i++) res[i] = (rhs[i] - v[i]) / (h1[i] + shift);
    free(v);
    free(y);
    free(zy);
    return it;
}

/* Dense Cholesky solve of (Hs + shift I) x = rhs, N = n + m - 1.  The
 * mathematical equivalent of the reference's default SimplicialLDLT path
 * (src/sinkhorn_linsolve.cpp:20-56,58-106); small sizes only. */
int orc_hess_direct(int n, int m, const double* h1, const double* h2,
                    const int* colptr, const int* rowind, const double* val,
                    const double* rhs, double shift, double* res)
{
    const long N = n + m - 1;
    double* A = (double*)calloc((size_t)N * N, sizeof(double));
    if (!A) return -1;
    for (int i = 0; i < n; i++) A[i * N + i] = h1[i] + shift;
    for (int j = 0; j < m - 1; j++) A[(n + j) * N + n + j] = h2[j] + shift;
    for (int j = 0; j < m - 1; j++)
        for (int k = colptr[j]; k < colptr[j + 1]; k++) {
            A[(long)rowind[k] * N + n + j] += val[k];
            A[(long)(n + j) * N + rowind[k]] += val[k];
        }
    /* in-place lower Cholesky */
    for (long j = 0; j < N; j++) {
        double s = A[j * N + j];
        for (long k = 0; k < j; k++) s -= A[j * N + k] * A[j * N + k];
        if (!(s > 0.0)) {
            free(A);
            return -2;
        }
        double d = sqrt(s);
        A[j * N + j] = d;
        for (long i = j + 1; i < N; i++) {
            double t = A[i * N + j];
            const double* Ai = A + i * N;
            const double* Aj = A + j * N;
            for (long k = 0; k < j; k++) t -= Ai[k] * Aj[k];
            A[i * N + j] = t / d;
        }
    }
    for (long i = 0; i < N; i++) {
        double t = rhs[i];
        for (long k = 0; k < i; k++) t -= A[i * N + k] * res[k];
        res[i] = t / A[i * N + i];
    }
    for (long i = N - 1; i >= 0; i--) {
        double t = res[i];
        for (long k = i + 1; k < N; k++) t -= A[k * N + i] * res[k];
        res[i] = t / A[i * N + i];
    }
    free(A);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Line searches, src/sinkhorn_problem_ls.cpp                               */
/* ------------------------------------------------------------------------ */

/* Minimiser of the quadratic through fa, ga (at a) and fb (at b);
 * src/sinkhorn_problem_ls.cpp:72-77 */
static double quad_min_fg(double a, double b, double fa, double ga, double fb)
{
    const double ba = b - a;
    const double w = 0.5 * ba * ga / (fa - fb + ba * ga);
    return a + w * ba;
}

/* Minimiser of the quadratic with derivatives ga (at a) and gb (at b);
 * src/sinkhorn_problem_ls.cpp:84-88 */
static double quad_min_gg(double a, double b, double ga, double gb)
{
    const double w = ga / (ga - gb);
    return a + w * (b - a);
}

/* Local minimiser of the cubic interpolating fa, ga, fb, gb;
 * src/sinkhorn_problem_ls.cpp:93-153 */
double orc_cubic_minimizer(double a, double b, double fa, double fb, double ga, double gb,
                           int* exists)
{
    const double apb = a + b, ba = b - a, ba2 = ba * ba;
    const double fba = fb - fa, gba = gb - ga;
    const double z3 = (ga + gb) * ba - 2.0 * fba;
    const double z2 = 0.5 * (gba * ba2 - 3.0 * apb * z3);
    const double z1 = fba * ba2 - apb * z2 - (a * apb + b * b) * z3;
    const double eps = 2.220446049250313e-16;
    if (fabs(z3) < eps * fabs(z2) || fabs(z3) < eps * fabs(z1)) {
        *exists = (z2 * ba > 0.0);
        return *exists ? (-0.5 * z1 / z2) : b;
    }
    const double u = z2 / (3.0 * z3), v = z1 / z2;
    const double vu = v / u;
    *exists = (vu <= 1.0);
    if (!*exists) return b;
    double r1, r2;
    if (fabs(u) >= fabs(v)) {
        const double w = 1.0 + sqrt(1.0 - vu);
        r1 = -u * w;
        r2 = -v / w;
    } else {
        const double sqrtd = sqrt(fabs(u)) * sqrt(fabs(v)) * sqrt(1 - u / v);
        r1 = -u - sqrtd;
        r2 = -u + sqrtd;
    }
    return (z3 * ba > 0.0) ? (r1 > r2 ? r1 : r2) : (r1 < r2 ? r1 : r2);
}

/* Next trial step from the bracket state; src/sinkhorn_problem_ls.cpp:157-223 */
double orc_step_selection(double al, double au, double at, double fl, double fu, double ft,
                          double gl, double gu, double gt)
{
    if (al == au) return al;
    if (!isfinite(ft) || !isfinite(gt)) return (al + at) / 2.0;
    int ac_exists;
    const double ac = orc_cubic_minimizer(al, at, fl, ft, gl, gt, &ac_exists);
    const double aq = quad_min_fg(al, at, fl, gl, ft);
    if (ft > fl) {
        if (!ac_exists) return aq;
        return (fabs(ac - al) < fabs(aq - al)) ? ac : ((aq + ac) / 2.0);
    }
    const double as = quad_min_gg(al, at, gl, gt);
    if (gt * gl < 0.0) return (fabs(ac - at) >= fabs(as - at)) ? ac : as;
    const double deltal = 1.1, deltau = 0.66;
    if (fabs(gt) < fabs(gl)) {
        const int choose_ac = ac_exists && ((ac - at) * (at - al) > 0.0) &&
                              (fabs(ac - at) < fabs(as - at));
        const double res = choose_ac ? ac : as;
        const double lim = at + deltau * (au - at);
        return (at > al) ? (lim < res ? lim : res) : (lim > res ? lim : res);
    }
    if (!isfinite(au) || !isfinite(fu) || !isfinite(gu)) return at + deltal * (at - al);
    int ae_exists;
    const double ae = orc_cubic_minimizer(at, au, ft, fu, gt, gu, &ae_exists);
    const double lim = at + deltau * (au - at);
    return (at > al) ? (lim < ae ? lim : ae) : (lim > ae ? lim : ae);
}

typedef struct {
    int n, m;
    const double* M;
    const double* a;
    const double* b;
    double reg;
    double* T; /* n x m */
} prob_t;

/* First candidate step that lowers f, else the best one (the first of equal
 * values, since only a strictly smaller value replaces it); *last_cand is the
 * last candidate evaluated, whose T the caller holds afterwards.
 * src/sinkhorn_problem_ls.cpp:13-41 */
double orc_line_selection(long N, orc_obj_fn fobj, void* ctx, const double* cand, int nc,
                          const double* gamma, const double* direc, double curobj,
                          double* objval, double* work, double* last_cand)
{
    double best = 1.0;
    *objval = INFINITY;
    for (int c = 0; c < nc; c++) {
        for (long i = 0; i < N; i++) work[i] = gamma[i] + cand[c] * direc[i];
        const double f = fobj(ctx, work);
        if (last_cand) *last_cand = cand[c];
        if (f < *objval) {
            best = cand[c];
            *objval = f;
        }
        if (*objval < curobj) return best;
    }
    return best;
}

/* More-Thuente-style strong-Wolfe search, c1 = 1e-4, c2 = 0.9, at most
 * max_iter trials (the reference's default is 20), steps in [1e-10, 2].
 * Restates Problem::line_search_wolfe, src/sinkhorn_problem_ls.cpp:226-470
 * (defaults at src/sinkhorn_problem.h:136-144) over a generic objective:
 * fg(ctx, x, grad) returns f(x) and writes its gradient.  On return x and
 * grad hold the last trial evaluated (*last_step; *ntrials trials), which is
 * the returned step unless *recompute_T is set or the trials ran out. */
double orc_line_search_wolfe(long N, orc_objgrad_fn fg, void* ctx, double init_step,
                             const double* gamma, const double* direc, double curobj,
                             const double* curgrad, int max_iter, int* recompute_T, double* x,
                             double* grad, double* last_step, int* ntrials)
{
    const double c1 = 1e-4, c2 = 0.9;
    if (max_iter <= 0) max_iter = 20;
    *recompute_T = 0;
    if (ntrials) *ntrials = 0;
    if (last_step) *last_step = 0.0;
    const double step_min = 1e-10, step_max = 2.0;
    if (init_step < step_min) init_step = step_min;
    if (init_step > step_max) init_step = step_max;
    double step = init_step;
    double fx = curobj;
    double dg = dot(curgrad, direc, N);
    const double fx_init = curobj, dg_init = dg;
    if (dg_init > 0.0) {
        *recompute_T = 1;
        return init_step;
    }
    const double test_decr = c1 * dg_init, test_curv = -c2 * dg_init;
    double I_lo = 0.0, I_hi = INFINITY;
    double fI_lo = 0.0, fI_hi = INFINITY;
    double gI_lo = (1.0 - c1) * dg_init, gI_hi = INFINITY;
    double psiI_lo = fI_lo;
    int bracketed = 0, use_step_min_safeguard = 1;
    double I_width = INFINITY, I_width_prev = INFINITY;
    int shrink_fail = 0;
    const double delta_max = 1.1, delta_min = 7.0 / 12, shrink = 0.66;
    for (int iter = 0; iter < max_iter; iter++) {
        for (long i = 0; i < N; i++) x[i] = gamma[i] + step * direc[i];
        fx = fg(ctx, x, grad);
        if (ntrials) *ntrials = iter + 1;
        if (last_step) *last_step = step;
        dg = dot(grad, direc, N);
        const double psit = fx - fx_init - step * test_decr;
        const double dpsit = dg - test_decr;
        if (psit <= 0.0 && fabs(dg) <= test_curv) return step;
        if (step <= step_min && (psit > 0.0 || dpsit >= 0.0)) return step;
        if (step >= step_max && (psit <= 0.0 && dpsit < 0.0)) return step;
        const double ft = psit, gt = dpsit; /* f is psi throughout (:558-571) */
        if (use_step_min_safeguard && (psit <= 0.0 && dpsit < 0.0)) use_step_min_safeguard = 0;
        double new_step;
        const int in_case_2 = (psit <= psiI_lo) && (dpsit * (I_lo - step) > 0.0);
        if (in_case_2) {
            new_step = step + delta_max * (step - I_lo);
            if (new_step > step_max) new_step = step_max;
        } else {
            new_step = orc_step_selection(I_lo, I_hi, step, fI_lo, fI_hi, ft, gI_lo, gI_hi, gt);
            if (new_step < step_min) new_step = step_min;
            if (new_step > step_max) new_step = step_max;
            if (use_step_min_safeguard) {
                const double upper = (step_min > delta_min * step) ? step_min : delta_min * step;
                if (new_step < step_min) new_step = step_min;
                if (new_step > upper) new_step = upper;
            }
        }
        if (psit > psiI_lo) {
            I_hi = step;
            fI_hi = ft;
            gI_hi = gt;
        } else if (in_case_2) {
            I_lo = step;
            fI_lo = ft;
            gI_lo = gt;
            psiI_lo = psit;
        } else {
            I_hi = I_lo;
            fI_hi = fI_lo;
            gI_hi = gI_lo;
            I_lo = step;
            fI_lo = ft;
            gI_lo = gt;
            psiI_lo = psit;
        }
        if (!bracketed && !in_case_2) {
            const double left = (I_lo < I_hi) ? I_lo : I_hi;
            const double right = (I_lo > I_hi) ? I_lo : I_hi;
            bracketed = (left >= step_min && right <= step_max);
        }
        if (bracketed) {
            I_width_prev = I_width;
            I_width = fabs(I_hi - I_lo);
            if (I_width_prev < INFINITY && I_width > shrink * I_width_prev)
                shrink_fail += 1;
            else
                shrink_fail = 0;
            if (shrink_fail >= 2) {
                new_step = (I_lo + I_hi) / 2.0;
                shrink_fail = 0;
            }
        }
        step = new_step;
    }
    /* Trials exhausted (:448-469): the step of the last update is kept when
     * psi formed from the last trial's f at that step is no worse than
     * psi(I_lo) -- T, f and the gradient are then the last trial's, as the
     * reference leaves recompute_T false -- else I_lo, or the initial step
     * when I_lo is not positive, both with recompute_T set. */
    const double psit = fx - fx_init - step * test_decr;
    if (psit <= psiI_lo) return step;
    if (I_lo <= 0.0) {
        *recompute_T = 1;
        return init_step;
    }
    *recompute_T = 1;
    return I_lo;
}

/* The EROT dual as a line-search objective: T is recomputed into P->T. */
static double erot_fg(void* ctx, const double* x, double* grad)
{
    const prob_t* P = (const prob_t*)ctx;
    return orc_dual_obj_grad(P->n, P->m, P->M, P->a, P->b, P->reg, x, grad, P->T, 1);
}

static double erot_f(void* ctx, const double* x)
{
    const prob_t* P = (const prob_t*)ctx;
    const double Tsum = orc_compute_T(P->n, P->m, P->M, P->reg, x, P->T);
    return P->reg * Tsum - dot(x, P->a, P->n) - dot(x + P->n, P->b, P->m - 1);
}

/* line_search_wolfe on the EROT dual of (M, a, b, reg); see above. */
double orc_erot_line_search_wolfe(int n, int m, const double* M, const double* a,
                                  const double* b, double reg, double init_step,
                                  const double* gamma, const double* direc, double curobj,
                                  const double* curgrad, int max_iter, int* recompute_T,
                                  double* x, double* grad, double* last_step, int* ntrials)
{
    prob_t P = {n, m, M, a, b, reg, (double*)malloc(sizeof(double) * (size_t)n * m)};
    const double st = orc_line_search_wolfe(n + m - 1, erot_fg, &P, init_step, gamma, direc,
                                            curobj, curgrad, max_iter, recompute_T, x, grad,
                                            last_step, ntrials);
    free(P.T);
    return st;
}

/* ---- Driver trace (test infrastructure) --------------------------------- *
 * When a buffer is set, the SSNS, sparse Newton and SPLR drivers append one
 * record per iteration: ORC_TRACE_HDR scalars, then five vectors of length
 * n + m - 1: gamma before the step, the point where the T (hence the
 * gradient and the sparsified Hessian) in use was formed, the direction, and
 * SPLR's y and s (zeros otherwise).  The scalars are listed in erot_oracle.h. */
static double* g_trace = NULL;
static long g_trace_cap = 0, g_trace_used = 0;

void orc_trace_set(double* buf, long cap)
{
    g_trace = buf;
    g_trace_cap = cap;
    g_trace_used = 0;
}

long orc_trace_used(void) { return g_trace_used; }

static void trace_put(const double* v, long k)
{
    if (!g_trace) return;
    for (long i = 0; i < k; i++) {
        if (g_trace_used < g_trace_cap) g_trace[g_trace_used] = v ? v[i] : 0.0;
        g_trace_used++;
    }
}

static void trace_record(const double* hdr, long N, const double* gamma, const double* tpt,
                         const double* direc, const double* y, const double* s)
{
    if (!g_trace) return;
    trace_put(hdr, ORC_TRACE_HDR);
    trace_put(gamma, N);
    trace_put(tpt, N);
    trace_put(direc, N);
    trace_put(y, N);
    trace_put(s, N);
}

/* ------------------------------------------------------------------------ */
/* Drivers                                                                   */
/* ------------------------------------------------------------------------ */

typedef struct {
    double* h1;
    double* h2;
    int* colptr;
    int* rowind;
    double* val;
    long long nnz;
} hess_t;

static void hess_free(hess_t* H)
{
    free(H->h1);
    free(H->h2);
    free(H->colptr);
    free(H->rowind);
    free(H->val);
    memset(H, 0, sizeof(*H));
}

/* h1 = r/reg, h2 = ct/reg (r, ct from the gradient), sigma = sparsified T/reg.
 * src/sinkhorn_problem_obj.cpp:432-457, src/sinkhorn_hess.cpp:15-55 */
static void hess_compute(const prob_t* P, const double* grad, int threshold_rule, double delta,
                         double hint_or_density, hess_t* H)
{
    const int n = P->n, m = P->m;
    hess_free(H);
    H->h1 = (double*)malloc(sizeof(double) * n);
    H->h2 = (double*)malloc(sizeof(double) * (m > 1 ? m - 1 : 1));
    H->colptr = (int*)malloc(sizeof(int) * m);
    for (int i = 0; i < n; i++) H->h1[i] = (grad[i] + P->a[i]) / P->reg;
    for (int j = 0; j < m - 1; j++) H->h2[j] = (grad[n + j] + P->b[j]) / P->reg;
    if (threshold_rule)
        H->nnz = orc_sparsify_mat3(n, m, P->T, delta, hint_or_density, H->colptr, &H->rowind,
                                   &H->val);
    else
        H->nnz = orc_sparsify_density(n, m, P->T, hint_or_density, 1, H->colptr, &H->rowind,
                                      &H->val);
    for (long long k = 0; k < H->nnz; k++) H->val[k] /= P->reg;
}

/* src/sinkhorn_hess.h:75-80 */
static double hess_density(const prob_t* P, const hess_t* H)
{
    const double N = (double)(P->n + P->m - 1);
    return (2.0 * (double)H->nnz + P->n + P->m - 1) / N / N;
}

/* src/sinkhorn_splr.cpp:17-22, src/sinkhorn_sparse_newton.cpp:17-22 */
static double density_T2H(double denT, int n, int m)
{
    const double nnz = 2.0 * denT * n * (m - 1) + n + m - 1;
    return nnz / (double)(n + m - 1) / (double)(n + m - 1);
}

void orc_low_rank_combine(long N, const double* iAr, const double* iAy, const double* r,
                          const double* y, const double* s, double* res);

static double opt_cg_tol(const orc_opts* o) { return (o && o->cg_tol > 0) ? o->cg_tol : 1e-8; }

/* SinkhornLinearSolver::solve, src/sinkhorn_linsolve.cpp:58-106 */
static void lin_solve(const prob_t* P, const hess_t* H, const orc_opts* o, const double* rhs,
                      double shift, double* res)
{
    if (o->lin_method == 0)
        orc_hess_cg(P->n, P->m, H->h1, H->h2, H->colptr, H->rowind, H->val, rhs, shift,
                    opt_cg_tol(o), o->cg_max_iter, res, NULL);
    else
        orc_hess_direct(P->n, P->m, H->h1, H->h2, H->colptr, H->rowind, H->val, rhs, shift, res);
}

/* Inverse-BFGS update of (Hs + shift I)^{-1} applied to rhs.
 * src/sinkhorn_linsolve.cpp:160-225 (two solves; the reference always uses
 * SimplicialLDLT here). */
static void lin_solve_low_rank(const prob_t* P, const hess_t* H, const orc_opts* o,
                               const double* rhs, double shift, const double* y,
                               const double* s, double* res)
{
    const long N = P->n + P->m - 1;
    double* iAr = (double*)malloc(sizeof(double) * N);
    double* iAy = (double*)malloc(sizeof(double) * N);
    if (o->lowrank_cg) {
        orc_hess_cg(P->n, P->m, H->h1, H->h2, H->colptr, H->rowind, H->val, rhs, shift,
                    opt_cg_tol(o), o->cg_max_iter, iAr, NULL);
        orc_hess_cg(P->n, P->m, H->h1, H->h2, H->colptr, H->rowind, H->val, y, shift,
                    opt_cg_tol(o), o->cg_max_iter, iAy, NULL);
    } else {
        orc_hess_direct(P->n, P->m, H->h1, H->h2, H->colptr, H->rowind, H->val, rhs, shift, iAr);
        orc_hess_direct(P->n, P->m, H->h1, H->h2, H->colptr, H->rowind, H->val, y, shift, iAy);
    }
    orc_low_rank_combine(N, iAr, iAy, rhs, y, s, res);
    free(iAr);
    free(iAy);
}

/* res = iAr + ((1 + y'iAy / y's) s'r / y's - y'iAr / y's) s - (s'r / y's) iAy, the
 * combination step of SinkhornLinearSolver::solve_low_rank,
 * src/sinkhorn_linsolve.cpp:207-213, given iAr = A^-1 r and iAy = A^-1 y. */
void orc_low_rank_combine(long N, const double* iAr, const double* iAy, const double* r,
                          const double* y, const double* s, double* res)
{
    const double rs = dot(s, r, N), ys = dot(y, s, N);
    const double yiAy = dot(y, iAy, N), yiAr = dot(y, iAr, N);
    const double rs_ys = rs / ys;
    const double cs = (1.0 + yiAy / ys) * rs_ys - yiAr / ys;
    for (long i = 0; i < N; i++) res[i] = iAr[i] + cs * s[i] - rs_ys * iAy[i];
}

/* gamma = [alpha + beta_{m-1}; beta_{0..m-2} - beta_{m-1}];
 * src/sinkhorn_bcd.cpp:50-51 */
static void normalise(int n, int m, const double* alpha, const double* beta, double* gamma)
{
    for (int i = 0; i < n; i++) gamma[i] = alpha[i] + beta[m - 1];
    for (int j = 0; j < m - 1; j++) gamma[n + j] = beta[j] - beta[m - 1];
}

/* Initial point shared by every driver: alpha = 0, beta = optimal_beta(0). */
static void initial_gamma(const prob_t* P, double* gamma)
{
    double* alpha = (double*)calloc(P->n, sizeof(double));
    double* beta = (double*)malloc(sizeof(double) * P->m);
    orc_optimal_beta(P->n, P->m, P->M, P->b, P->reg, alpha, beta);
    normalise(P->n, P->m, alpha, beta, gamma);
    free(alpha);
    free(beta);
}

/* Plain Sinkhorn / BCD.  Restates sinkhorn_bcd_internal, src/sinkhorn_bcd.cpp:14-124:
 * init alpha = 0, beta = optimal_beta(alpha); test ||g|| < tol before each
 * update; histories hold niter + 1 entries. */
int orc_sinkhorn_bcd(int n, int m, const double* M, const double* a, const double* b,
                     double reg, double tol, int max_iter, int verbose,
                     double* dual, double* obj_vals, double* mar_errs, double* run_times,
                     int* hist_len)
{
    const long N = n + m - 1;
    double* alpha = (double*)calloc(n, sizeof(double));
    double* beta = (double*)malloc(sizeof(double) * m);
    double* grad = (double*)malloc(sizeof(double) * N);
    orc_optimal_beta(n, m, M, b, reg, alpha, beta);
    double t0 = now_ms();
    normalise(n, m, alpha, beta, dual);
    double f = orc_dual_obj_grad(n, m, M, a, b, reg, dual, grad, NULL, 1);
    double gnorm = norm2(grad, N);
    double t1 = now_ms();
    obj_vals[0] = f;
    mar_errs[0] = gnorm;
    run_times[0] = t1 - t0;
    int h = 1, i;
    for (i = 0; i < max_iter; i++) {
        if (verbose >= 1) printf("iter = %d, objval = %g, ||grad|| = %g\n", i, f, gnorm);
        t0 = now_ms();
        if (gnorm < tol) break;
        orc_optimal_alpha(n, m, M, a, reg, beta, alpha);
        orc_optimal_beta(n, m, M, b, reg, alpha, beta);
        normalise(n, m, alpha, beta, dual);
        f = orc_dual_obj_grad(n, m, M, a, b, reg, dual, grad, NULL, 1);
        gnorm = norm2(grad, N);
        t1 = now_ms();
        obj_vals[h] = f;
        mar_errs[h] = gnorm;
        run_times[h] = run_times[h - 1] + (t1 - t0);
        h++;
    }
    free(alpha);
    free(beta);
    free(grad);
    *hist_len = h;
    return i;
}

double orc_bcd_iterations(int n, int m, const double* M, const double* a, const double* b,
                          double reg, int iters, double* gamma_out)
{
    const long N = n + m - 1;
    double* alpha = (double*)calloc(n, sizeof(double));
    double* beta = (double*)malloc(sizeof(double) * m);
    double* grad = (double*)malloc(sizeof(double) * N);
    orc_optimal_beta(n, m, M, b, reg, alpha, beta);
    double t0 = now_ms();
    for (int it = 0; it < iters; it++) {
        orc_optimal_alpha(n, m, M, a, reg, beta, alpha);
        orc_optimal_beta(n, m, M, b, reg, alpha, beta);
        normalise(n, m, alpha, beta, gamma_out);
        orc_dual_obj_grad(n, m, M, a, b, reg, gamma_out, grad, NULL, 1);
    }
    double t1 = now_ms();
    free(alpha);
    free(beta);
    free(grad);
    return (t1 - t0) * 1e-3;
}

/* Safe and sparse Newton (SSNS).  Restates sinkhorn_ssns_internal,
 * src/sinkhorn_ssns.cpp:16-234, including the reuse of T from the last
 * evaluated line-selection candidate (:185-187): after an accepted step the
 * gradient and the sparsified Hessian come from that candidate's T, and f
 * from that T with the accepted gamma. */
int orc_sinkhorn_ssns(int n, int m, const double* M, const double* a, const double* b,
                      double reg, double tol, int max_iter, int verbose, const orc_opts* opts,
                      double* gamma, double* obj_vals, double* mar_errs, double* run_times,
                      double* densities, int* hist_len)
{
    const long N = n + m - 1;
    const double rho_t1 = 0.25, rho_t2 = 0.75, mu_s1 = 4.0, mu_s2 = 0.5, kappa = 0.001,
                 nu0 = 0.01;
    const double cand[4] = {1.0, 0.5, 0.25, 0.1};
    double mu = opts->mu0, rho = 1.0;
    prob_t P = {n, m, M, a, b, reg, (double*)malloc(sizeof(double) * (size_t)n * m)};
    double* step = (double*)malloc(sizeof(double) * N);
    double* Hstep = (double*)malloc(sizeof(double) * N);
    double* direc = (double*)malloc(sizeof(double) * N);
    double* g = (double*)malloc(sizeof(double) * N);
    double* rhs = (double*)malloc(sizeof(double) * N);
    double* work = (double*)malloc(sizeof(double) * N);
    double* tpt = (double*)malloc(sizeof(double) * N);
    hess_t H;
    memset(&H, 0, sizeof(H));
    initial_gamma(&P, gamma);
    double t0 = now_ms();
    double f = orc_dual_obj_grad(n, m, M, a, b, reg, gamma, g, P.T, 1);
    memcpy(tpt, gamma, sizeof(double) * N);
    double gnorm = norm2(g, N);
    double delta = nu0 * gnorm;
    hess_compute(&P, g, 1, delta, 0.001, &H);
    double t1 = now_ms();
    obj_vals[0] = f;
    mar_errs[0] = gnorm;
    run_times[0] = t1 - t0;
    densities[0] = hess_density(&P, &H);
    int h = 1, i;
    for (i = 0; i < max_iter; i++) {
        if (verbose >= 1) printf("iter = %d, objval = %g, ||grad|| = %g\n", i, f, gnorm);
        t0 = now_ms();
        if (gnorm < tol) break;
        const double shift = mu * gnorm;
        for (long k = 0; k < N; k++) rhs[k] = -g[k];
        lin_solve(&P, &H, opts, rhs, shift, direc);
        double newf, c_last = 0.0;
        double alpha = orc_line_selection(N, erot_f, &P, cand, 4, gamma, direc, f, &newf, work,
                                          &c_last);
        for (long k = 0; k < N; k++) step[k] = alpha * direc[k];
        orc_apply_hsx(n, m, H.h1, H.h2, H.colptr, H.rowind, H.val, step, Hstep);
        const double numer = f - newf;
        const double denom = -dot(g, step, N) - 0.5 * dot(step, Hstep, N);
        rho = numer / denom;
        double hdr[ORC_TRACE_HDR] = {1, i, f, gnorm, shift, mu, delta, alpha, newf, rho, numer,
                                     denom, c_last, hess_density(&P, &H), 0, 0};
        if (g_trace) {
            memcpy(work, gamma, sizeof(double) * N);
        }
        if (rho > 0.0)
            for (long k = 0; k < N; k++) gamma[k] += step[k];
        if (rho < rho_t1)
            mu = mu_s1 * mu;
        else if (rho > rho_t2)
            mu = (mu_s2 * mu > kappa) ? mu_s2 * mu : kappa;
        hdr[14] = mu;
        trace_record(hdr, N, work, tpt, direc, NULL, NULL);
        if (rho > 0.0) {
            f = orc_dual_obj_grad(n, m, M, a, b, reg, gamma, g, P.T, 0);
            for (long k = 0; k < N; k++) tpt[k] = gamma[k] + (c_last - alpha) * direc[k];
            gnorm = norm2(g, N);
            delta = nu0 * gnorm;
            hess_compute(&P, g, 1, delta, hess_density(&P, &H), &H);
        }
        t1 = now_ms();
        obj_vals[h] = f;
        mar_errs[h] = gnorm;
        run_times[h] = run_times[h - 1] + (t1 - t0);
        densities[h] = hess_density(&P, &H);
        h++;
    }
    hess_free(&H);
    free(P.T);
    free(step);
    free(Hstep);
    free(direc);
    free(g);
    free(rhs);
    free(work);
    free(tpt);
    *hist_len = h;
    return i;
}

/* Shared body of the sparse Newton and SPLR drivers.
 * src/sinkhorn_sparse_newton.cpp:24-170, src/sinkhorn_splr.cpp:24-199 */
static int newton_family(int splr, int n, int m, const double* M, const double* a,
                         const double* b, double reg, double tol, int max_iter, int verbose,
                         const orc_opts* opts, double* gamma, double* obj_vals,
                         double* mar_errs, double* run_times, double* densities, int* hist_len)
{
    const long N = n + m - 1;
    const double density_max = opts->density, density_min = 0.01 * opts->density;
    double density = 0.1 * density_max;
    const double shift_max = opts->shift;
    const double eps = 1e-6;
    prob_t P = {n, m, M, a, b, reg, (double*)malloc(sizeof(double) * (size_t)n * m)};
    double* direc = (double*)malloc(sizeof(double) * N);
    double* g = (double*)malloc(sizeof(double) * N);
    double* g_pre = (double*)calloc(N, sizeof(double));
    double* gamma_pre = (double*)calloc(N, sizeof(double));
    double* y = (double*)malloc(sizeof(double) * N);
    double* s = (double*)malloc(sizeof(double) * N);
    double* rhs = (double*)malloc(sizeof(double) * N);
    double* x = (double*)malloc(sizeof(double) * N);
    double* xg = (double*)malloc(sizeof(double) * N);
    double* tpt = (double*)malloc(sizeof(double) * N);
    double* gamma_old = (double*)malloc(sizeof(double) * N);
    hess_t H;
    memset(&H, 0, sizeof(H));
    initial_gamma(&P, gamma);
    double t0 = now_ms();
    double f = orc_dual_obj_grad(n, m, M, a, b, reg, gamma, g, P.T, 1);
    memcpy(tpt, gamma, sizeof(double) * N);
    double gnorm = norm2(g, N);
    const double gnorm_init = gnorm;
    hess_compute(&P, g, 0, 0.0, density, &H);
    double t1 = now_ms();
    obj_vals[0] = f;
    mar_errs[0] = gnorm;
    run_times[0] = t1 - t0;
    densities[0] = density_T2H(density, n, m);
    int h = 1, i;
    double alpha = 1.0;
    for (i = 0; i < max_iter; i++) {
        if (verbose >= 1) printf("iter = %d, objval = %g, ||grad|| = %g\n", i, f, gnorm);
        t0 = now_ms();
        if (gnorm < tol || !isfinite(f)) break;
        const double shift = (gnorm < shift_max) ? gnorm : shift_max;
        for (long k = 0; k < N; k++) rhs[k] = -g[k];
        int low_rank = 0;
        if (splr) {
            for (long k = 0; k < N; k++) {
                y[k] = g[k] - g_pre[k];
                s[k] = gamma[k] - gamma_pre[k];
            }
            const double ys = dot(y, s, N), yy = dot(y, y, N);
            low_rank = (i > 0) && (ys > eps * yy);
        }
        if (low_rank)
            lin_solve_low_rank(&P, &H, opts, rhs, shift, y, s, direc);
        else
            lin_solve(&P, &H, opts, rhs, shift, direc);
        int recompute_T = 1, ntrials = 0;
        double last_step = 0.0;
        double init = 1.5 * alpha < 1.0 ? 1.5 * alpha : 1.0;
        alpha = orc_line_search_wolfe(N, erot_fg, &P, init, gamma, direc, f, g, opts->ls_max_iter,
                                      &recompute_T, x, xg, &last_step, &ntrials);
        double hdr[ORC_TRACE_HDR] = {splr ? 3 : 2, i, f, gnorm, shift, density, init, alpha,
                                     recompute_T, low_rank, ntrials, last_step, 0, 0, 0, 0};
        if (g_trace) memcpy(gamma_old, gamma, sizeof(double) * N);
        if (splr) memcpy(gamma_pre, gamma, sizeof(double) * N);
        for (long k = 0; k < N; k++) gamma[k] += alpha * direc[k];
        if (splr) memcpy(g_pre, g, sizeof(double) * N);
        const double* tnew = recompute_T ? gamma : x;
        if (g_trace) {
            trace_record(hdr, N, gamma_old, tpt, direc, low_rank ? y : NULL, low_rank ? s : NULL);
        }
        f = orc_dual_obj_grad(n, m, M, a, b, reg, gamma, g, P.T, recompute_T);
        memcpy(tpt, tnew, sizeof(double) * N);
        const double gnorm_pre = gnorm;
        gnorm = norm2(g, N);
        if (splr) {
            const int bad_move = (gnorm_pre < gnorm_init) && (gnorm > gnorm_pre);
            density *= bad_move ? 1.1 : 0.99;
        } else {
            const int bad_move = (gnorm_pre < gnorm_init) && (gnorm > 1.1 * gnorm_pre);
            density *= bad_move ? 2.0 : 0.9;
        }
        if (density < density_min) density = density_min;
        if (density > density_max) density = density_max;
        hess_compute(&P, g, 0, 0.0, density, &H);
        t1 = now_ms();
        obj_vals[h] = f;
        mar_errs[h] = gnorm;
        run_times[h] = run_times[h - 1] + (t1 - t0);
        densities[h] = density_T2H(density, n, m);
        h++;
    }
    hess_free(&H);
    free(P.T);
    free(direc);
    free(g);
    free(g_pre);
    free(gamma_pre);
    free(y);
    free(s);
    free(rhs);
    free(x);
    free(xg);
    free(tpt);
    free(gamma_old);
    *hist_len = h;
    return i;
}

int orc_sinkhorn_sparse_newton(int n, int m, const double* M, const double* a, const double* b,
                               double reg, double tol, int max_iter, int verbose,
                               const orc_opts* opts, double* dual, double* obj_vals,
                               double* mar_errs, double* run_times, double* densities,
                               int* hist_len)
{
    return newton_family(0, n, m, M, a, b, reg, tol, max_iter, verbose, opts, dual, obj_vals,
                         mar_errs, run_times, densities, hist_len);
}

int orc_sinkhorn_splr(int n, int m, const double* M, const double* a, const double* b,
                      double reg, double tol, int max_iter, int verbose, const orc_opts* opts,
                      double* dual, double* obj_vals, double* mar_errs, double* run_times,
                      double* densities, int* hist_len)
{
    return newton_family(1, n, m, M, a, b, reg, tol, max_iter, verbose, opts, dual, obj_vals,
                         mar_errs, run_times, densities, hist_len);
}

/* ------------------------------------------------------------------------ */
/* Dense Newton, L-BFGS dual and APDAGD drivers                              */
/* ------------------------------------------------------------------------ */

/* Dense dual Hessian at gamma (N x N row-major, N = n + m - 1): diagonal
 * [r; c~] / reg, off-diagonal block T[:, 0:m-1] / reg and its transpose.
 * Restates Problem::dual_obj_grad_densehess, src/sinkhorn_problem_obj.cpp:396-418
 * (r = grad + a, c~ = grad + b~ as there). */
static void dense_hess(const prob_t* P, const double* grad, double* H)
{
    const int n = P->n, m = P->m;
    const long N = n + m - 1;
    memset(H, 0, sizeof(double) * (size_t)N * N);
    for (int i = 0; i < n; i++) H[i * N + i] = (grad[i] + P->a[i]) / P->reg;
    for (int j = 0; j < m - 1; j++)
        H[(n + j) * N + n + j] = (grad[n + j] + P->b[j]) / P->reg;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m - 1; j++) {
            const double v = P->T[(long)i * m + j] / P->reg;
            H[(long)i * N + n + j] = v;
            H[(long)(n + j) * N + i] = v;
        }
}

/* Solve A x = rhs for symmetric positive definite A (row-major, overwritten
 * by its Cholesky factor).  The reference uses Eigen::LLT
 * (src/sinkhorn_newton.cpp:87-88).  Returns 0, or -2 if A is not PD. */
static int chol_solve(long N, double* A, const double* rhs, double* x)
{
    for (long j = 0; j < N; j++) {
        double s = A[j * N + j];
        for (long k = 0; k < j; k++) s -= A[j * N + k] * A[j * N + k];
        if (!(s > 0.0)) return -2;
        const double d = sqrt(s);
        A[j * N + j] = d;
        for (long i = j + 1; i < N; i++) {
            double t = A[i * N + j];
            for (long k = 0; k < j; k++) t -= A[i * N + k] * A[j * N + k];
            A[i * N + j] = t / d;
        }
    }
    for (long i = 0; i < N; i++) {
        double t = rhs[i];
        for (long k = 0; k < i; k++) t -= A[i * N + k] * x[k];
        x[i] = t / A[i * N + i];
    }
    for (long i = N - 1; i >= 0; i--) {
        double t = x[i];
        for (long k = i + 1; k < N; k++) t -= A[k * N + i] * x[k];
        x[i] = t / A[i * N + i];
    }
    return 0;
}

int orc_newton_direction(int n, int m, const double* M, const double* a, const double* b,
                         double reg, const double* gamma, double* direc)
{
    const long N = n + m - 1;
    prob_t P = {n, m, M, a, b, reg, (double*)malloc(sizeof(double) * (size_t)n * m)};
    double* g = (double*)malloc(sizeof(double) * N);
    double* H = (double*)malloc(sizeof(double) * (size_t)N * N);
    double* rhs = (double*)malloc(sizeof(double) * N);
    orc_dual_obj_grad(n, m, M, a, b, reg, gamma, g, P.T, 1);
    dense_hess(&P, g, H);
    for (long k = 0; k < N; k++) rhs[k] = -g[k];
    const int rc = chol_solve(N, H, rhs, direc);
    free(P.T);
    free(g);
    free(H);
    free(rhs);
    return rc;
}

/* Newton's method with the dense Hessian, Cholesky direction and Armijo
 * backtracking (theta = 0.5, kappa = 0.5, <= 20 trials).  Restates
 * sinkhorn_newton_internal, src/sinkhorn_newton.cpp:17-129.  Returns niter, or
 * -2 when the Hessian is not positive definite (the reference's LLT would then
 * return a meaningless direction; the GPU driver reports an error). */
int orc_sinkhorn_newton(int n, int m, const double* M, const double* a, const double* b,
                        double reg, double tol, int max_iter, int verbose,
                        double* gamma, double* obj_vals, double* mar_errs, double* run_times,
                        int* hist_len)
{
    const long N = n + m - 1;
    const double theta = 0.5, kappa = 0.5;
    const int nlinesearch = 20;
    prob_t P = {n, m, M, a, b, reg, (double*)malloc(sizeof(double) * (size_t)n * m)};
    double* g = (double*)malloc(sizeof(double) * N);
    double* H = (double*)malloc(sizeof(double) * (size_t)N * N);
    double* rhs = (double*)malloc(sizeof(double) * N);
    double* direc = (double*)malloc(sizeof(double) * N);
    double* newgamma = (double*)malloc(sizeof(double) * N);
    initial_gamma(&P, gamma);
    double t0 = now_ms();
    double f = orc_dual_obj_grad(n, m, M, a, b, reg, gamma, g, P.T, 1);
    dense_hess(&P, g, H);
    double gnorm = norm2(g, N);
    double t1 = now_ms();
    obj_vals[0] = f;
    mar_errs[0] = gnorm;
    run_times[0] = t1 - t0;
    int h = 1, i, rc = 0;
    for (i = 0; i < max_iter; i++) {
        if (verbose >= 1) printf("iter = %d, objval = %g, ||grad|| = %g\n", i, f, gnorm);
        t0 = now_ms();
        if (gnorm < tol || !isfinite(f)) break;
        for (long k = 0; k < N; k++) rhs[k] = -g[k];
        if (chol_solve(N, H, rhs, direc) != 0) {
            rc = -2;
            break;
        }
        double step = 1.0;
        const double thresh = theta * dot(g, direc, N);
        for (int k = 0; k < nlinesearch; k++) {
            for (long q = 0; q < N; q++) newgamma[q] = gamma[q] + step * direc[q];
            const double newf = orc_dual_obj(n, m, M, a, b, reg, newgamma);
            if (newf <= f + step * thresh) break;
            step *= kappa;
        }
        memcpy(gamma, newgamma, sizeof(double) * N);
        f = orc_dual_obj_grad(n, m, M, a, b, reg, gamma, g, P.T, 1);
        dense_hess(&P, g, H);
        gnorm = norm2(g, N);
        t1 = now_ms();
        obj_vals[h] = f;
        mar_errs[h] = gnorm;
        run_times[h] = run_times[h - 1] + (t1 - t0);
        h++;
    }
    free(P.T);
    free(g);
    free(H);
    free(rhs);
    free(direc);
    free(newgamma);
    *hist_len = h;
    return rc != 0 ? rc : i;
}

/* Limited-memory BFGS inverse-Hessian operator with mm correction pairs kept
 * in a ring.  Restates BFGSMat (L-BFGS part), src/LBFGSpp/LBFGSpp/BFGSMat.h:
 * reset :65-83, add_correction :86-97 (theta = y'y / s'y of the newest pair),
 * apply_Hv (two-loop recursion, H0 = I / theta) :276-302. */
typedef struct {
    int mm, ncorr, ptr;
    long N;
    double theta;
    double *S, *Y, *ys, *al;
} lbfgs_mat;

static void lbfgs_reset(lbfgs_mat* B, long N, int mm)
{
    B->mm = mm;
    B->N = N;
    B->theta = 1.0;
    B->ncorr = 0;
    B->ptr = mm;
    B->S = (double*)malloc(sizeof(double) * (size_t)N * mm);
    B->Y = (double*)malloc(sizeof(double) * (size_t)N * mm);
    B->ys = (double*)malloc(sizeof(double) * mm);
    B->al = (double*)malloc(sizeof(double) * mm);
}

static void lbfgs_free(lbfgs_mat* B)
{
    free(B->S);
    free(B->Y);
    free(B->ys);
    free(B->al);
}

static void lbfgs_add(lbfgs_mat* B, const double* s, const double* y)
{
    const int loc = B->ptr % B->mm;
    double* sl = B->S + (long)loc * B->N;
    double* yl = B->Y + (long)loc * B->N;
    memcpy(sl, s, sizeof(double) * B->N);
    memcpy(yl, y, sizeof(double) * B->N);
    const double ys = dot(sl, yl, B->N);
    B->ys[loc] = ys;
    B->theta = dot(yl, yl, B->N) / ys;
    if (B->ncorr < B->mm) B->ncorr++;
    B->ptr = loc + 1;
}

static void lbfgs_apply_hv(lbfgs_mat* B, const double* v, double a, double* res)
{
    const long N = B->N;
    for (long k = 0; k < N; k++) res[k] = a * v[k];
    int j = B->ptr % B->mm;
    for (int i = 0; i < B->ncorr; i++) {
        j = (j + B->mm - 1) % B->mm;
        const double* sj = B->S + (long)j * N;
        const double* yj = B->Y + (long)j * N;
        B->al[j] = dot(sj, res, N) / B->ys[j];
        for (long k = 0; k < N; k++) res[k] -= B->al[j] * yj[k];
    }
    for (long k = 0; k < N; k++) res[k] /= B->theta;
    for (int i = 0; i < B->ncorr; i++) {
        const double* sj = B->S + (long)j * N;
        const double* yj = B->Y + (long)j * N;
        const double bet = dot(yj, res, N) / B->ys[j];
        for (long k = 0; k < N; k++) res[k] += (B->al[j] - bet) * sj[k];
        j = (j + 1) % B->mm;
    }
}

/* Exposed for the tests: the two-loop product H v after adding the pairs
 * (S[k], Y[k]), k = 0..npairs-1, in order (skipping none), memory mm. */
void orc_lbfgs_apply(long N, int mm, int npairs, const double* S, const double* Y,
                     const double* v, double* res)
{
    lbfgs_mat B;
    lbfgs_reset(&B, N, mm);
    for (int k = 0; k < npairs; k++) lbfgs_add(&B, S + (long)k * N, Y + (long)k * N);
    lbfgs_apply_hv(&B, v, 1.0, res);
    lbfgs_free(&B);
}

/* Step candidate of the zoom phase: minimiser of the quadratic through
 * (lo, f_lo), (hi, f_hi) with slope dg_lo at lo, replaced by the midpoint
 * when not finite, outside (lo, hi) or within 1% of the interval of an end.
 * Restates LineSearchNocedalWright::quad_interp,
 * src/LBFGSpp/LBFGSpp/LineSearchNocedalWright.h:26-55. */
static double nw_quad_interp(double lo, double hi, double flo, double fhi, double dglo)
{
    const double fdiff = fhi - flo, sdiff = hi - lo, smid = (hi + lo) / 2.0;
    double c = fdiff * lo - smid * sdiff * dglo;
    c = c / (fdiff - sdiff * dglo);
    const double end_dist = fmin(fabs(c - lo), fabs(c - hi));
    const int near_end = end_dist < 0.01 * fabs(sdiff);
    const int bisect = !isfinite(c) || c <= fmin(lo, hi) || c >= fmax(lo, hi) || near_end;
    return bisect ? smid : c;
}

/* Strong-Wolfe line search (ftol 1e-4, wolfe 0.9, <= 20 evaluations,
 * expansion x2): bracketing, then zoom by quadratic interpolation.  Restates
 * LineSearchNocedalWright::LineSearch,
 * src/LBFGSpp/LBFGSpp/LineSearchNocedalWright.h:58-283.  On entry *fx, grad,
 * *dg are at xp; on exit x, *fx, grad, *dg, *step describe the accepted point
 * and *last_f is the last value evaluated.  Returns 0, -3 (direction is not a
 * descent direction), -4 (insufficient precision) or -5 (no decrease). */
/* f and grad of the EROT dual at x (the L-BFGS objective). */
static double erot_objgrad(void* ctx, const double* x, double* grad)
{
    const prob_t* P = (const prob_t*)ctx;
    return orc_dual_obj_grad(P->n, P->m, P->M, P->a, P->b, P->reg, x, grad, NULL, 1);
}

static int nw_line_search(long N, orc_objgrad_fn fn, void* ctx, const double* xp,
                          const double* drt, double* step, double* fx, double* grad, double* dg,
                          double* x, double* last_f, double* xlo, double* glo)
{
    const int max_ls = 20;
    const double ftol = 1e-4, wolfe = 0.9, expansion = 2.0;
    const double fx_init = *fx, dg_init = *dg;
    if (dg_init > 0.0) return -3;
    const double test_decr = ftol * dg_init, test_curv = -wolfe * dg_init;
    double step_hi, fx_hi, step_lo = 0.0, fx_lo = fx_init, dg_lo = dg_init;
    memcpy(xlo, xp, sizeof(double) * N);
    memcpy(glo, grad, sizeof(double) * N);
    int iter = 0;
    for (;;) {
        for (long k = 0; k < N; k++) x[k] = xp[k] + (*step) * drt[k];
        *fx = fn(ctx, x, grad);
        *last_f = *fx;
        *dg = dot(grad, drt, N);
        if (*fx - fx_init > (*step) * test_decr || (0.0 < step_lo && *fx >= fx_lo)) {
            step_hi = *step;
            fx_hi = *fx;
            break;
        }
        if (fabs(*dg) <= test_curv) return 0;
        step_hi = step_lo;
        fx_hi = fx_lo;
        step_lo = *step;
        fx_lo = *fx;
        dg_lo = *dg;
        memcpy(xlo, x, sizeof(double) * N);
        memcpy(glo, grad, sizeof(double) * N);
        if (*dg >= 0.0) break;
        iter++;
        if (iter >= max_ls) {
            /* every trial decreased f: keep the latest (== the _lo state) */
            return 0;
        }
        *step *= expansion;
    }
    for (;;) {
        *step = nw_quad_interp(step_lo, step_hi, fx_lo, fx_hi, dg_lo);
        for (long k = 0; k < N; k++) x[k] = xp[k] + (*step) * drt[k];
        *fx = fn(ctx, x, grad);
        *last_f = *fx;
        *dg = dot(grad, drt, N);
        if (*fx - fx_init > (*step) * test_decr || *fx >= fx_lo) {
            if (*step == step_hi) return -4;
            step_hi = *step;
            fx_hi = *fx;
        } else {
            if (fabs(*dg) <= test_curv) return 0;
            if (*dg * (step_hi - step_lo) >= 0.0) {
                step_hi = step_lo;
                fx_hi = fx_lo;
            }
            if (*step == step_lo) return -4;
            step_lo = *step;
            fx_lo = *fx;
            dg_lo = *dg;
            memcpy(xlo, x, sizeof(double) * N);
            memcpy(glo, grad, sizeof(double) * N);
        }
        iter++;
        if (iter >= max_ls) {
            if (step_lo <= 0.0) return -5;
            *step = step_lo;
            *fx = fx_lo;
            *dg = dg_lo;
            memcpy(x, xlo, sizeof(double) * N);
            memcpy(grad, glo, sizeof(double) * N);
            return 0;
        }
    }
}

/* Exposed for the tests: one line search from gamma along direc with the
 * initial step; returns the status, fills the accepted step, f and gradient. */
int orc_nw_line_search(int n, int m, const double* M, const double* a, const double* b,
                       double reg, const double* gamma, const double* direc, double init_step,
                       double* step, double* f, double* grad, double* last_f)
{
    const long N = n + m - 1;
    prob_t P = {n, m, M, a, b, reg, NULL};
    double* x = (double*)malloc(sizeof(double) * N);
    double* xlo = (double*)malloc(sizeof(double) * N);
    double* glo = (double*)malloc(sizeof(double) * N);
    *f = orc_dual_obj_grad(n, m, M, a, b, reg, gamma, grad, NULL, 1);
    double dg = dot(grad, direc, N);
    *step = init_step;
    const int rc = nw_line_search(N, erot_objgrad, &P, gamma, direc, step, f, grad, &dg, x, last_f,
                                  xlo, glo);
    free(x);
    free(xlo);
    free(glo);
    return rc;
}

/* LBFGSSolver::minimize, src/LBFGSpp/LBFGS.h:64-189 (memory 6, parameters
 * src/LBFGSpp/LBFGSpp/Param.h:154-171, no past-delta test) on the objective
 * fn over N variables from x: stops when ||grad|| <= eps or ||grad|| <=
 * eps_rel ||x||, or after max_iter iterations (0 = unlimited).  it(ictx, x,
 * last_f, ||grad||) is the iterate() callback, called with the initial point
 * and after every line search; last_f is the most recent objective evaluated
 * (after a search that fell back to an earlier trial, not the accepted
 * point's value).  Returns minimize()'s niter (1 when the initial point
 * passes) or a negative line-search status. */
int orc_lbfgs_minimize(long N, orc_objgrad_fn fn, void* fctx, double eps_abs, double eps_rel,
                       int max_iter, orc_iter_fn it, void* ictx, double* x)
{
    const double eps = 2.220446049250313e-16;
    double* grad = (double*)malloc(sizeof(double) * N);
    double* xp = (double*)malloc(sizeof(double) * N);
    double* gradp = (double*)malloc(sizeof(double) * N);
    double* drt = (double*)malloc(sizeof(double) * N);
    double* vs = (double*)malloc(sizeof(double) * N);
    double* vy = (double*)malloc(sizeof(double) * N);
    double* xlo = (double*)malloc(sizeof(double) * N);
    double* glo = (double*)malloc(sizeof(double) * N);
    lbfgs_mat B;
    lbfgs_reset(&B, N, 6);
    int k = 0, rc = 0;
    double fx = fn(fctx, x, grad);
    double last_f = fx;
    double gnorm = norm2(grad, N);
    it(ictx, x, last_f, gnorm);
    if (gnorm <= eps_abs || gnorm <= eps_rel * norm2(x, N)) {
        k = 1;
        goto done;
    }
    for (long q = 0; q < N; q++) drt[q] = -grad[q];
    double step = 1.0 / norm2(drt, N);
    k = 1;
    for (;;) {
        memcpy(xp, x, sizeof(double) * N);
        memcpy(gradp, grad, sizeof(double) * N);
        double dg = dot(grad, drt, N);
        rc = nw_line_search(N, fn, fctx, xp, drt, &step, &fx, grad, &dg, x, &last_f, xlo, glo);
        if (rc != 0) break;
        gnorm = norm2(grad, N);
        it(ictx, x, last_f, gnorm);
        if (gnorm <= eps_abs || gnorm <= eps_rel * norm2(x, N)) break;
        if (max_iter != 0 && k >= max_iter) break;
        for (long q = 0; q < N; q++) {
            vs[q] = x[q] - xp[q];
            vy[q] = grad[q] - gradp[q];
        }
        if (dot(vs, vy, N) > eps * dot(vy, vy, N)) lbfgs_add(&B, vs, vy);
        lbfgs_apply_hv(&B, grad, -1.0, drt);
        step = 1.0;
        k++;
    }
done:
    lbfgs_free(&B);
    free(grad);
    free(xp);
    free(gradp);
    free(drt);
    free(vs);
    free(vy);
    free(xlo);
    free(glo);
    return rc != 0 ? rc : k;
}

typedef struct {
    int h, hist_cap, verbose;
    double t0;
    double *obj_vals, *mar_errs, *run_times;
} erot_hist;

/* iterate() of src/sinkhorn_lbfgs_dual.cpp:56-76 */
static void erot_lbfgs_iter(void* ctx, const double* x, double last_f, double gnorm)
{
    (void)x;
    erot_hist* H = (erot_hist*)ctx;
    const double t1 = now_ms();
    if (H->h < H->hist_cap) {
        H->obj_vals[H->h] = last_f;
        H->mar_errs[H->h] = gnorm;
        H->run_times[H->h] = (H->h > 0 ? H->run_times[H->h - 1] : 0.0) + (t1 - H->t0);
        H->h++;
    }
    if (H->verbose >= 1) printf("iter = %d, objval = %g, ||grad|| = %g\n", H->h - 1, last_f, gnorm);
    H->t0 = now_ms();
}

/* L-BFGS on the dual (epsilon = tol, epsilon_rel = 0, max_iterations =
 * max_iter).  Restates sinkhorn_lbfgs_dual_internal,
 * src/sinkhorn_lbfgs_dual.cpp:84-125: obj_vals holds the last objective
 * evaluated, mar_errs the accepted point's ||grad||.  At most hist_cap
 * history entries are kept. */
int orc_sinkhorn_lbfgs_dual(int n, int m, const double* M, const double* a, const double* b,
                            double reg, double tol, int max_iter, int verbose, int hist_cap,
                            double* x, double* obj_vals, double* mar_errs, double* run_times,
                            int* hist_len)
{
    prob_t P = {n, m, M, a, b, reg, NULL};
    initial_gamma(&P, x);
    erot_hist H = {0, hist_cap, verbose, now_ms(), obj_vals, mar_errs, run_times};
    const int rc = orc_lbfgs_minimize(n + m - 1, erot_objgrad, &P, tol, 0.0, max_iter,
                                      erot_lbfgs_iter, &H, x);
    *hist_len = H.h;
    return rc;
}

/* Accelerated primal-dual adaptive gradient descent.  Restates
 * sinkhorn_apdagd_internal, src/sinkhorn_apdagd.cpp:14-183: dual variables
 * lambda = -gamma, adaptive L (halved per iteration, doubled per failed
 * test, <= 20 inner trials), primal plan = running combination
 * (1 - tau) plan + tau T(-lambda~); stop when primal - dual gap <= tol and
 * the full marginal error (all m columns) <= tol.  entropy = sum P (1 - log P),
 * so an exact zero in P makes it NaN (0 * inf), as in the reference.
 * plan (n x m) receives the primal plan; x (N) the final gamma = -lambda~. */
int orc_sinkhorn_apdagd(int n, int m, const double* M, const double* a, const double* b,
                        double reg, double tol, int max_iter, int verbose,
                        double* x, double* plan, double* obj_vals, double* mar_errs,
                        double* run_times, int* hist_len)
{
    const long N = n + m - 1;
    const double L0 = 1.0;
    const int max_inner = 20;
    prob_t P = {n, m, M, a, b, reg, NULL};
    double* grad = (double*)malloc(sizeof(double) * N);
    double* du = (double*)malloc(sizeof(double) * N);      /* dual */
    double* du_new = (double*)malloc(sizeof(double) * N);
    double* dp = (double*)malloc(sizeof(double) * N);      /* dualp */
    double* dp_new = (double*)malloc(sizeof(double) * N);
    double* dt = (double*)malloc(sizeof(double) * N);      /* dualt */
    double* Tn = (double*)malloc(sizeof(double) * (size_t)n * m);
    double* rs = (double*)malloc(sizeof(double) * n);
    double* cs = (double*)malloc(sizeof(double) * m);
    initial_gamma(&P, x);
    for (long k = 0; k < N; k++) du[k] = dp[k] = dt[k] = -x[k];
    orc_get_plan(n, m, M, reg, x, plan);
    double t0 = now_ms();
    double f = orc_dual_obj_grad(n, m, M, a, b, reg, x, grad, NULL, 1);
    double gnorm = norm2(grad, N);
    double t1 = now_ms();
    obj_vals[0] = f;
    mar_errs[0] = gnorm;
    run_times[0] = t1 - t0;
    int h = 1, i;
    double L = L0, pbeta = 0.0;
    for (i = 0; i < max_iter; i++) {
        t0 = now_ms();
        if (verbose >= 2)
            printf("\n[line search]=================================================\n");
        L *= 0.5;
        double ptau = 0.0, phi_new = 0.0;
        for (int j = 0; j < max_inner; j++) {
            const double palpha = (1.0 + sqrt(4.0 * L * pbeta + 1.0)) / (2.0 * L);
            const double pbeta_new = pbeta + palpha;
            ptau = palpha / pbeta_new;
            for (long k = 0; k < N; k++) {
                dt[k] = (1.0 - ptau) * du[k] + ptau * dp[k];
                x[k] = -dt[k];
            }
            f = orc_dual_obj_grad(n, m, M, a, b, reg, x, grad, NULL, 1);
            const double phit = -f;
            for (long k = 0; k < N; k++) {
                dp_new[k] = dp[k] + palpha * grad[k];
                du_new[k] = (1.0 - ptau) * du[k] + ptau * dp_new[k];
                x[k] = -du_new[k];
            }
            f = orc_dual_obj(n, m, M, a, b, reg, x);
            phi_new = -f;
            double gd = 0.0, dd = 0.0;
            for (long k = 0; k < N; k++) {
                const double d = du_new[k] - dt[k];
                gd += grad[k] * d;
                dd += d * d;
            }
            const double rhs = phit + gd - 0.5 * L * dd;
            if (verbose >= 2) {
                printf("║ j = %d, palpha = %g, pbeta = %g\n", j, palpha, pbeta_new);
                printf("║ phit = %g, phi_new = %g, rhs = %g\n", phit, phi_new, rhs);
                printf("║-------------------------------------------------------------\n");
            }
            if (phi_new >= rhs || j >= max_inner - 1) {
                pbeta = pbeta_new;
                memcpy(dp, dp_new, sizeof(double) * N);
                memcpy(du, du_new, sizeof(double) * N);
                break;
            }
            L *= 2.0;
        }
        for (long k = 0; k < N; k++) x[k] = -dt[k];
        orc_get_plan(n, m, M, reg, x, Tn);
        double ent = 0.0, mp = 0.0;
        memset(rs, 0, sizeof(double) * n);
        memset(cs, 0, sizeof(double) * m);
        for (int r = 0; r < n; r++)
            for (int c = 0; c < m; c++) {
                const long q = (long)r * m + c;
                const double v = (1.0 - ptau) * plan[q] + ptau * Tn[q];
                plan[q] = v;
                ent += v * (1.0 - log(v));
                mp += M[q] * v;
                rs[r] += v;
                cs[c] += v;
            }
        const double primal_obj = mp - reg * ent;
        double e2 = 0.0;
        for (int r = 0; r < n; r++) e2 += (rs[r] - a[r]) * (rs[r] - a[r]);
        for (int c = 0; c < m; c++) e2 += (cs[c] - b[c]) * (cs[c] - b[c]);
        const double mar_err = sqrt(e2);
        t1 = now_ms();
        obj_vals[h] = f;
        mar_errs[h] = mar_err;
        run_times[h] = run_times[h - 1] + (t1 - t0);
        h++;
        if (verbose >= 1)
            printf("iter = %d, primal_obj = %g, phi_new = %g, mar_err = %g\n", i, primal_obj,
                   phi_new, mar_err);
        if (primal_obj - phi_new <= tol && mar_err <= tol) break;
    }
    free(grad);
    free(du);
    free(du_new);
    free(dp);
    free(dp_new);
    free(dt);
    free(Tn);
    free(rs);
    free(cs);
    *hist_len = h;
    return i;
}

/* ------------------------------------------------------------------------ */
/* Feasible rounding and the primal value                                    */
/* ------------------------------------------------------------------------ */

/* Rounding of a nonnegative n x m plan onto the transport polytope of (a, b):
 * scale rows down to a, then columns down to b, then add the rank-one
 * correction err_a err_b' / |err_a|_1.  Restates approx_proj,
 * src/approx_proj.h:6-26 (Altschuler, Weed and Rigollet 2017, Algorithm 2),
 * including its order of operations: err_b is taken from the column sums of
 * the scaled matrix, and an exactly feasible input makes |err_a|_1 = 0, so
 * the correction (and the result) is NaN as in the reference. */
void orc_approx_proj(int n, int m, const double* plan, const double* a, const double* b,
                     double* out)
{
    double* x = (double*)malloc(sizeof(double) * n);
    double* y = (double*)malloc(sizeof(double) * m);
    double* ea = (double*)malloc(sizeof(double) * n);
    double* eb = (double*)calloc(m, sizeof(double));
    for (int i = 0; i < n; i++) {
        double r = 0.0;
        for (int j = 0; j < m; j++) r += plan[(long)i * m + j];
        x[i] = fmin(a[i] / r, 1.0);
    }
    for (long q = 0; q < (long)n * m; q++) out[q] = plan[q] * x[q / m];
    for (int j = 0; j < m; j++) y[j] = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++) y[j] += out[(long)i * m + j];
    for (int j = 0; j < m; j++) y[j] = fmin(b[j] / y[j], 1.0);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++) out[(long)i * m + j] *= y[j];
    double l1 = 0.0;
    for (int i = 0; i < n; i++) {
        double r = 0.0;
        for (int j = 0; j < m; j++) r += out[(long)i * m + j];
        ea[i] = a[i] - r;
        l1 += fabs(ea[i]);
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++) eb[j] += out[(long)i * m + j];
    for (int j = 0; j < m; j++) eb[j] = b[j] - eb[j];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++) out[(long)i * m + j] += ea[i] * eb[j] / l1;
    free(x);
    free(y);
    free(ea);
    free(eb);
}

/* Primal objective of the rounded plan P = approx_proj(T(gamma)):
 * <P, M> - reg sum P (1 - log P).  Restates Problem::primal_val,
 * src/sinkhorn_problem_obj.cpp:300-319. */
double orc_primal_val(int n, int m, const double* M, const double* a, const double* b,
                      double reg, const double* gamma)
{
    double* T = (double*)malloc(sizeof(double) * (size_t)n * m);
    double* P = (double*)malloc(sizeof(double) * (size_t)n * m);
    orc_get_plan(n, m, M, reg, gamma, T);
    orc_approx_proj(n, m, T, a, b, P);
    double ent = 0.0, mp = 0.0;
    for (long q = 0; q < (long)n * m; q++) {
        ent += P[q] * (1.0 - log(P[q]));
        mp += P[q] * M[q];
    }
    free(T);
    free(P);
    return mp - reg * ent;
}
