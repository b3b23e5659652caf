// qrot.cpp -- host side of the QROT (quadratically regularised OT) path: the
// dual objective and gradient, optimal_alpha / optimal_beta by Newton steps
// on the piecewise-linear root equations, and the BCD, gradient-descent and
// L-BFGS drivers, all over device vectors (gamma = [alpha (n); beta (m)]).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <memory>
#include <vector>

#include "driver_work.hpp"
#include "qrot.hpp"

namespace regot {

int set_error(int code, const char* what);

namespace {

QrotScratch& scratch(regot_problem* p)
{
    if (!p->qrot) p->qrot = std::make_unique<QrotScratch>(p->n, p->m);
    return *p->qrot;
}

}  // namespace

double qrot_obj_grad(regot_problem* p, const double* gamma, double* grad, double tau,
                     double* gnorm2)
{
    const int n = p->n, m = p->m;
    hipStream_t s = p->stream;
    QrotScratch& w = scratch(p);
    const int rows = qrot_rows(n, m);
    const int nstrip = ceil_div(m, 256), nblk = ceil_div(n, rows);
    const double* alpha = gamma;
    const double* beta = gamma + n;
    launch_qrot_tile(p->bulk, p->dM, p->ldm, n, m, alpha, beta, 1.0, nullptr, nullptr,
                     grad ? w.rowpart.p : nullptr, grad ? w.colpart.p : nullptr, w.sqpart, s);
    if (grad) {
        launch_qrot_fold(w.rowpart, nstrip, n, n, w.rsum, s);
        launch_qrot_fold(w.colpart, nblk, m, m, w.csum, s);
    }
    launch_qrot_sum1(w.sqpart, nstrip * nblk, w.sq, s);
    launch_qrot_finish(alpha, beta, p->a, p->b, n, m, p->reg, tau, w.sq, w.rsum, w.csum, grad,
                       w.fpart, w.gpart, w.out, s);
    HIP_CHECK(hipMemcpyAsync(w.hout.p, w.out.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (gnorm2) *gnorm2 = w.hout.p[1];
    return w.hout.p[0];
}

// Newton's method from the right on sum_k (x - y_k)_+ = reg w per column
// (cols) or row: the first step takes every y (x = mean + reg w / count), each
// later one the y below the current x, until the active set stops shrinking.
// The iterates decrease monotonically and every step but the last drops at
// least one breakpoint from the active set, so an equation with `nbreak`
// breakpoints (n per column equation, m per row equation) settles within
// nbreak + 1 steps; the root is (reg w + sum of the active y) / (their
// count), the formula of piecewise_linear_equation,
// src/qrot_problem.cpp:365-394.
void qrot_optimal(regot_problem* p, bool cols, const double* pot, double* out)
{
    const int n = p->n, m = p->m;
    hipStream_t s = p->stream;
    QrotScratch& w = scratch(p);
    const int count = cols ? m : n;
    const int nbreak = cols ? n : m;
    launch_qrot_pl_init(out, w.kprev, w.done, count, w.open, s);
    const double* wt = cols ? p->b.p : p->a.p;
    for (int step = 0; step <= nbreak + 1; step++) {
        launch_qrot_pl_step(p->bulk, p->dM, p->ldm, n, m, cols, pot, wt, p->reg, out, w.kprev,
                            w.done, w.spart, w.kpart, w.open, s);
        HIP_CHECK(hipMemcpyAsync(w.hopen.p, w.open.p, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        if (w.hopen.p[0] == 0) return;
    }
    fail(E_INTERNAL, "QROT root iteration did not settle");
}

// ---- Generalised Hessian (src/qrot_hess.cpp, src/qrot_cg.cpp) ----

namespace {

// k dots of length len on the stream, results on the host
void qdots(QrotHess& H, hipStream_t s, long len, int k, const double* const* a,
           const double* const* b, double* out)
{
    vec_dots(a, b, len, k, H.dsc, s);
    HIP_CHECK(hipMemcpyAsync(H.hs.p, H.dsc.p, k * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (int q = 0; q < k; q++) out[q] = H.hs.p[q];
}

double qsum(QrotHess& H, hipStream_t s, long len, const double* x)
{
    const double* a[1] = {x};
    const double* b[1] = {H.ones.p};
    double r;
    qdots(H, s, len, 1, a, b, &r);
    return r;
}

// out (n) = B x = sigma x - tau (1'x) 1, x (m)
void apply_B(regot_problem* p, QrotHess& H, const double* x, double tau, double* out)
{
    const int n = p->n, m = p->m;
    launch_qrot_sig(H.mask, n, m, true, false, nullptr, x, H.rowpart4, nullptr, p->stream);
    launch_qrot_fold(H.rowpart4, 4 * ceil_div(m, 256), n, n, out, p->stream);
    if (tau > 0.0) {
        const double c = tau * qsum(H, p->stream, m, x);
        vec_axpby(out, 1.0, out, -c, H.ones, n, p->stream);
    }
}

// out (m) = C x = sigma' x - tau (1'x) 1, x (n)
void apply_C(regot_problem* p, QrotHess& H, const double* x, double tau, double* out)
{
    const int n = p->n, m = p->m;
    QrotScratch& w = scratch(p);
    launch_qrot_sig(H.mask, n, m, false, true, x, nullptr, nullptr, w.colpart, p->stream);
    launch_qrot_fold(w.colpart, ceil_div(n, qrot_rows(n, m)), m, m, out, p->stream);
    if (tau > 0.0) {
        const double c = tau * qsum(H, p->stream, n, x);
        vec_axpby(out, 1.0, out, -c, H.ones, m, p->stream);
    }
}

// out (n) = A^-1 x, A = diag(h1 + shift) + tau 11' (src/qrot_hess.cpp:130-146)
void solve_A(regot_problem* p, QrotHess& H, const double* x, double shift, double tau,
             double* out)
{
    const int n = p->n;
    hipStream_t s = p->stream;
    double c = 0.0;
    if (tau > 0.0) {
        if (H.hinv_shift != shift) {
            launch_qrot_recip(H.h1, shift, H.hinv, n, s);
            H.hinv_shift = shift;
        }
        const double* a[2] = {H.hinv.p, x};
        const double* b[2] = {H.ones.p, H.hinv.p};
        double d[2];
        qdots(H, s, n, 2, a, b, d);
        c = d[1] / (1.0 / tau + d[0]);
    }
    launch_qrot_ainv(x, H.h1, shift, c, out, n, s);
}

// out (m) = Delta x = h2 .* x + shift x - C A^-1 B x + tau (1'x) 1
void apply_Delta(regot_problem* p, QrotHess& H, const double* x, double shift, double tau,
                 double* out)
{
    const int m = p->m;
    apply_B(p, H, x, tau, H.tn);
    solve_A(p, H, H.tn, shift, tau, H.tn2);
    apply_C(p, H, H.tn2, tau, H.tm2);
    const double c = tau > 0.0 ? tau * qsum(H, p->stream, m, x) : 0.0;
    launch_qrot_delta_fin(x, H.h2, shift, H.tm2, c, out, m, p->stream);
}

// out (n + m) = (H + tau vv') x: [sigma z + h1 .* w; sigma' w + h2 .* z] +
// tau (1'w - 1'z) v, x = [w; z] (src/qrot_hess.cpp:193-214 with shift 0)
void apply_H(regot_problem* p, QrotHess& H, const double* x, double tau, double* out)
{
    const int n = p->n, m = p->m;
    hipStream_t s = p->stream;
    QrotScratch& w = scratch(p);
    launch_qrot_sig(H.mask, n, m, true, false, nullptr, x + n, H.rowpart4, nullptr, s);
    launch_qrot_fold(H.rowpart4, 4 * ceil_div(m, 256), n, n, H.tn, s);
    launch_qrot_sig(H.mask, n, m, false, true, x, nullptr, nullptr, w.colpart, s);
    launch_qrot_fold(w.colpart, ceil_div(n, qrot_rows(n, m)), m, m, H.tm2, s);
    double c = 0.0;
    if (tau > 0.0) c = tau * (qsum(H, s, n, x) - qsum(H, s, m, x + n));
    launch_qrot_hfin(x, H.h1, H.tn, c, out, n, s);
    launch_qrot_hfin(x + n, H.h2, H.tm2, -c, out + n, m, s);
}

}  // namespace

void apply_qrot_hess(regot_problem* p, QrotHess& H, const double* x, double tau, double* out)
{
    apply_H(p, H, x, tau, out);
}

QrotHess& qrot_hess_build(regot_problem* p, const double* gamma)
{
    const int n = p->n, m = p->m;
    hipStream_t s = p->stream;
    QrotScratch& w = scratch(p);
    if (!w.hess) {
        w.hess = std::make_unique<QrotHess>(n, m);
        launch_fill(w.hess->ones, 1.0, n > m ? n : m, s);
    }
    QrotHess& H = *w.hess;
    launch_qrot_mask(p->bulk, p->dM, p->ldm, n, m, gamma, gamma + n, H.mask, s);
    launch_qrot_sig(H.mask, n, m, true, true, nullptr, nullptr, H.rowpart4, w.colpart, s);
    launch_qrot_fold(H.rowpart4, 4 * ceil_div(m, 256), n, n, H.h1, s);
    launch_qrot_fold(w.colpart, ceil_div(n, qrot_rows(n, m)), m, m, H.h2, s);
    H.hinv_shift = -1.0;
    return H;
}

// Block elimination of src/qrot_cg.cpp:9-51 and Eigen's ConjugateGradient
// (identity preconditioner) on Delta: x = 0, r = p = rhs; stop when ||r||^2 <
// max(tol^2 ||rhs||^2, DBL_MIN) or after max_iter steps.
int qrot_hess_solve(regot_problem* p, QrotHess& H, const double* rhs, double shift, double tau,
                    double tol, int max_iter, double* out, const double* x0)
{
    const int n = p->n, m = p->m;
    hipStream_t s = p->stream;
    double* x = out + n;
    solve_A(p, H, rhs, shift, tau, H.v);
    apply_C(p, H, H.v, tau, H.tm);
    vec_axpby(H.r, 1.0, rhs + n, -1.0, H.tm, m, s);     // r = z - C v (the CG rhs)
    double rhs2;
    {
        const double* a[1] = {H.r.p};
        qdots(H, s, m, 1, a, a, &rhs2);
    }
    // from the guess (which may be x itself): r = rhs - Delta x0
    double rr = rhs2;
    if (x0 && rhs2 > 0.0) {
        if (x0 != x) vec_copy(x, x0, m, s);
        apply_Delta(p, H, x, shift, tau, H.q);
        vec_axpby(H.r, 1.0, H.r, -1.0, H.q, m, s);
        const double* a[1] = {H.r.p};
        qdots(H, s, m, 1, a, a, &rr);
    } else {
        launch_fill(x, 0.0, m, s);
    }
    vec_copy(H.pv, H.r, m, s);
    int it = 0;
    if (rhs2 > 0.0) {
        const double thr = std::max(tol * tol * rhs2, 2.2250738585072014e-308);
        double abs_new = rr;
        if (rr >= thr) {
            while (it < max_iter) {
                apply_Delta(p, H, H.pv, shift, tau, H.q);
                double pq;
                {
                    const double* a[1] = {H.pv.p};
                    const double* b[1] = {H.q.p};
                    qdots(H, s, m, 1, a, b, &pq);
                }
                const double al = abs_new / pq;
                vec_axpby(x, 1.0, x, al, H.pv, m, s);
                vec_axpby(H.r, 1.0, H.r, -al, H.q, m, s);
                double rr;
                {
                    const double* a[1] = {H.r.p};
                    qdots(H, s, m, 1, a, a, &rr);
                }
                if (rr < thr) break;
                const double bt = rr / abs_new;
                abs_new = rr;
                vec_axpby(H.pv, 1.0, H.r, bt, H.pv, m, s);
                it++;
            }
        }
    }
    // r1 = A^-1 (w - B r2)
    apply_B(p, H, x, tau, H.tn);
    vec_axpby(H.tn, 1.0, rhs, -1.0, H.tn, n, s);
    solve_A(p, H, H.tn, shift, tau, out);
    return it;
}

// GRSSN (src/qrot_grssn.cpp:16-133): direction (H + shift I + tau vv')^-1 (-g)
// by qrot_hess_solve (method 0: tol 1e-8 and at most 2m CG steps, the
// reference's CG; other methods name its direct factorisations, an exact
// solve, run here as CG to 1e-12 with at most 10m + 1000 steps; cg_tol /
// cg_max_iter override), Armijo backtracking from 1 (theta = kappa = 0.5, 20
// trials, the last one taken); test ||g|| < tol before each step.
static void solve_qrot_grssn(regot_problem* p, double tol, int max_iter, int verbose,
                             const regot_opts& o, regot_result* res)
{
    const int n = p->n, m = p->m;
    const long N = static_cast<long>(n) + m;
    hipStream_t s = p->stream;
    const double reg = p->reg, tau = o.tau, shift = o.shift;
    const double cg_tol = o.cg_tol > 0.0 ? o.cg_tol : (o.method == 0 ? 1e-8 : 1e-12);
    const int cg_max = o.cg_max_iter > 0 ? o.cg_max_iter : (o.method == 0 ? 2 * m : 10 * m + 1000);
    Work w(p, N);
    w.og = [&](const double* x, double* g, double* gn2) {
        return qrot_obj_grad(p, x, g, tau, gn2);
    };
    if (o.x0) {
        HIP_CHECK(hipMemcpyAsync(w.gamma, o.x0, N * sizeof(double), hipMemcpyHostToDevice, s));
    } else {
        launch_fill(w.gamma, 0.0, n, s);
        qrot_optimal(p, true, w.gamma, w.gamma.p + n);
    }
    res->qrot = true;
    Timer timer;
    timer.tic();
    auto record = [&](const double* x, double f, double gnorm) {
        const double prim = qrot_round_plan(p, x, x + n, nullptr, 0);
        const double dur = timer.toc();
        res->obj_vals.push_back(f);
        res->prim_vals.push_back(prim);
        res->mar_errs.push_back(gnorm / reg);
        res->run_times.push_back((res->run_times.empty() ? 0.0 : res->run_times.back()) + dur);
        timer.tic();
    };
    double gnorm;
    double f = w.obj_grad(w.gamma, w.g, &gnorm);
    record(w.gamma, f, gnorm);
    int it;
    for (it = 0; it < max_iter; it++) {
        if (verbose >= 1) {
            printf("iter = %d, objval = %g, ||grad|| = %g\n", it, f, gnorm);
            fflush(stdout);
        }
        if (gnorm < tol) break;
        QrotHess& H = qrot_hess_build(p, w.gamma);
        vec_axpby(w.rhs, -1.0, w.g, 0.0, nullptr, N, s);
        const int cgi = qrot_hess_solve(p, H, w.rhs, shift, tau, cg_tol, cg_max, w.direc);
        w.cg_iters += cgi;
        w.alg_bytes += w.pass_bytes() + (2.0 * (cgi + 2)) * (n * (m / 8.0));
        const double thresh = 0.5 * w.dot(w.g, w.direc);
        double step = 1.0;
        for (int k = 0; k < 20; k++) {
            w.move(w.x, w.gamma, step, w.direc);
            const double fn = w.obj(w.x);
            if (fn <= f + step * thresh) break;
            step *= 0.5;
        }
        std::swap(w.gamma, w.x);
        f = w.obj_grad(w.gamma, w.g, &gnorm);
        record(w.gamma, f, gnorm);
    }
    res->niter = it;
    w.report(res);
    res->dual.resize(N);
    HIP_CHECK(hipMemcpy(res->dual.data(), w.gamma.p, N * sizeof(double), hipMemcpyDeviceToHost));
}

// ASSN (src/qrot_assn.cpp:16-170): Newton direction with shift mu = lam ||g||
// (CG as GRSSN), u = gamma + d taken when ||g(u)|| <= nu xi (xi: the last
// accepted gradient norm, ||g_0|| at the start); otherwise rho = -g(u)'d /
// ||d||^2 chooses the projection point v = gamma + (g(u)'d / ||g(u)||^2) g(u)
// (rho >= eta1 and ||g(v)|| <= ||g||) or the gradient step gamma - 0.5 g, and
// lam becomes (lam_min + lam) / 2, 2 lam or 4 lam (rho >= eta2, >= eta1,
// below).  nu = 0.99, eta1 = 0.25, eta2 = 0.75, lam_min = 0.01, lam_0 = 1.
static void solve_qrot_assn(regot_problem* p, double tol, int max_iter, int verbose,
                            const regot_opts& o, regot_result* res)
{
    const int n = p->n, m = p->m;
    const long N = static_cast<long>(n) + m;
    hipStream_t s = p->stream;
    const double reg = p->reg, tau = o.tau;
    const double nu = 0.99, eta1 = 0.25, eta2 = 0.75, lam_min = 0.01;
    const double cg_tol = o.cg_tol > 0.0 ? o.cg_tol : (o.method == 0 ? 1e-8 : 1e-12);
    const int cg_max = o.cg_max_iter > 0 ? o.cg_max_iter : (o.method == 0 ? 2 * m : 10 * m + 1000);
    Work w(p, N);
    w.og = [&](const double* x, double* g, double* gn2) {
        return qrot_obj_grad(p, x, g, tau, gn2);
    };
    if (o.x0) {
        HIP_CHECK(hipMemcpyAsync(w.gamma, o.x0, N * sizeof(double), hipMemcpyHostToDevice, s));
    } else {
        launch_fill(w.gamma, 0.0, n, s);
        qrot_optimal(p, true, w.gamma, w.gamma.p + n);
    }
    res->qrot = true;
    Timer timer;
    timer.tic();
    auto record = [&](const double* x, double f, double gnorm) {
        const double prim = qrot_round_plan(p, x, x + n, nullptr, 0);
        const double dur = timer.toc();
        res->obj_vals.push_back(f);
        res->prim_vals.push_back(prim);
        res->mar_errs.push_back(gnorm / reg);
        res->run_times.push_back((res->run_times.empty() ? 0.0 : res->run_times.back()) + dur);
        timer.tic();
    };
    // u = w.x, g(u) = w.xg, v = w.y, g(v) = w.sv
    double gnorm;
    double f = w.obj_grad(w.gamma, w.g, &gnorm);
    record(w.gamma, f, gnorm);
    double xi = 0.0, lam = 1.0;
    int it;
    for (it = 0; it < max_iter; it++) {
        if (verbose >= 1) {
            printf("iter = %d, objval = %g, ||grad|| = %g\n", it, f, gnorm);
            fflush(stdout);
        }
        if (gnorm < tol) break;
        if (it == 0) xi = gnorm;
        QrotHess& H = qrot_hess_build(p, w.gamma);
        vec_axpby(w.rhs, -1.0, w.g, 0.0, nullptr, N, s);
        const int cgi = qrot_hess_solve(p, H, w.rhs, lam * gnorm, tau, cg_tol, cg_max, w.direc);
        w.cg_iters += cgi;
        w.alg_bytes += w.pass_bytes() + (2.0 * (cgi + 2)) * (n * (m / 8.0));
        w.move(w.x, w.gamma, 1.0, w.direc);
        double gunorm;
        w.obj_grad(w.x, w.xg, &gunorm);
        if (gunorm <= nu * xi) {
            std::swap(w.gamma, w.x);
            xi = gunorm;
        } else {
            const double* xs[2] = {w.xg.p, w.direc.p};
            const double* ys[2] = {w.direc.p, w.direc.p};
            double d[2];
            w.dots(2, xs, ys, d);
            const double numer = -d[0];
            const double rho = numer / d[1];
            w.move(w.y, w.gamma, (-numer) / (gunorm * gunorm), w.xg);
            double gvnorm;
            w.obj_grad(w.y, w.sv, &gvnorm);
            if (rho >= eta1 && gvnorm <= gnorm)
                std::swap(w.gamma, w.y);
            else
                w.move(w.gamma, w.gamma, -0.5, w.g);
            if (rho >= eta2)
                lam = 0.5 * (lam_min + lam);
            else if (rho >= eta1)
                lam *= 2.0;
            else
                lam *= 4.0;
        }
        f = w.obj_grad(w.gamma, w.g, &gnorm);
        record(w.gamma, f, gnorm);
    }
    res->niter = it;
    w.report(res);
    res->dual.resize(N);
    HIP_CHECK(hipMemcpy(res->dual.data(), w.gamma.p, N * sizeof(double), hipMemcpyDeviceToHost));
}

// S5N (src/qrot_s5n.cpp:16-196): shift = mu (primal value + f / reg), the
// duality gap of the rounded plan (delta = 1); direction by the CG solve,
// warm-started from the last direction's beta block after a rejected step;
// step from line_selection2 (src/qrot_problem.cpp:337-361) over {1, 0.5,
// 0.1}, then {0.05, 0.01, 1e-3, 1e-4} on alternate failures; rho = (f -
// f_new) / (-g's - s'Hs / 2) accepts the step when > 0 and moves mu by x4
// (rho < 0.25) or x0.5 down to 1e-3 (rho > 0.75); H, f and g are recomputed
// only after an accepted step; every iteration records f, the primal value
// and ||g|| / reg.
static void solve_qrot_s5n(regot_problem* p, double tol, int max_iter, int verbose,
                           const regot_opts& o, regot_result* res)
{
    const int n = p->n, m = p->m;
    const long N = static_cast<long>(n) + m;
    hipStream_t s = p->stream;
    const double reg = p->reg, tau = o.tau;
    const double cg_tol = o.cg_tol > 0.0 ? o.cg_tol : (o.method == 0 ? 1e-8 : 1e-12);
    const int cg_max = o.cg_max_iter > 0 ? o.cg_max_iter : (o.method == 0 ? 2 * m : 10 * m + 1000);
    const double alphas[3] = {1.0, 0.5, 0.1};
    const double alphas_small[4] = {0.05, 0.01, 0.001, 1e-4};
    Work w(p, N);
    w.og = [&](const double* x, double* g, double* gn2) {
        return qrot_obj_grad(p, x, g, tau, gn2);
    };
    if (o.x0) {
        HIP_CHECK(hipMemcpyAsync(w.gamma, o.x0, N * sizeof(double), hipMemcpyHostToDevice, s));
    } else {
        launch_fill(w.gamma, 0.0, n, s);
        qrot_optimal(p, true, w.gamma, w.gamma.p + n);
    }
    launch_fill(w.direc, 0.0, N, s);
    res->qrot = true;
    Timer timer;
    timer.tic();
    auto record = [&](const double* x, double f, double gnorm) {
        const double prim = qrot_round_plan(p, x, x + n, nullptr, 0);
        const double dur = timer.toc();
        res->obj_vals.push_back(f);
        res->prim_vals.push_back(prim);
        res->mar_errs.push_back(gnorm / reg);
        res->run_times.push_back((res->run_times.empty() ? 0.0 : res->run_times.back()) + dur);
        timer.tic();
    };
    // line_selection2: the first candidate below cur, else the best one
    auto select = [&](const double* cand, int nc, double cur, double* objval) {
        double best = 1.0;
        *objval = HUGE_VAL;
        for (int k = 0; k < nc; k++) {
            w.move(w.x, w.gamma, cand[k], w.direc);
            const double fk = w.obj(w.x);
            if (fk < *objval) {
                best = cand[k];
                *objval = fk;
            }
            if (*objval < cur) break;
        }
        return best;
    };
    double gnorm;
    double f = w.obj_grad(w.gamma, w.g, &gnorm);
    QrotHess* H = &qrot_hess_build(p, w.gamma);
    record(w.gamma, f, gnorm);
    double mu = o.mu0, rho = 1.0;
    bool use_small = true;
    int it;
    for (it = 0; it < max_iter; it++) {
        const double dgap = res->prim_vals.back() + f / reg;
        if (verbose >= 1) {
            printf("iter = %d, objval = %g, ||grad|| = %g, ldgap = %g\n", it, f, gnorm,
                   std::log10(dgap));
            fflush(stdout);
        }
        if (gnorm < tol) break;
        const double shift = mu * std::pow(dgap, 1.0);
        vec_axpby(w.rhs, -1.0, w.g, 0.0, nullptr, N, s);
        const int cgi = qrot_hess_solve(p, *H, w.rhs, shift, tau, cg_tol, cg_max, w.direc,
                                        rho <= 0.0 ? w.direc.p + n : nullptr);
        w.cg_iters += cgi;
        w.alg_bytes += (2.0 * (cgi + 2)) * (n * (m / 8.0));
        double newf;
        double alpha = select(alphas, 3, f, &newf);
        if (newf >= f) {
            if (use_small) alpha = select(alphas_small, 4, f, &newf);
            use_small = !use_small;
        } else {
            use_small = true;
        }
        vec_axpby(w.step, alpha, w.direc, 0.0, nullptr, N, s);
        apply_qrot_hess(p, *H, w.step, tau, w.Hstep);
        const double* xs[2] = {w.g.p, w.step.p};
        const double* ys[2] = {w.step.p, w.Hstep.p};
        double d[2];
        w.dots(2, xs, ys, d);
        rho = (f - newf) / (-d[0] - 0.5 * d[1]);
        if (rho > 0.0) vec_axpby(w.gamma, 1.0, w.gamma, 1.0, w.step, N, s);
        if (rho < 0.25)
            mu *= 4.0;
        else if (rho > 0.75)
            mu = std::max(0.5 * mu, 0.001);
        if (rho > 0.0) {
            f = w.obj_grad(w.gamma, w.g, &gnorm);
            H = &qrot_hess_build(p, w.gamma);
        }
        record(w.gamma, f, gnorm);
    }
    res->niter = it;
    w.report(res);
    res->dual.resize(N);
    HIP_CHECK(hipMemcpy(res->dual.data(), w.gamma.p, N * sizeof(double), hipMemcpyDeviceToHost));
}

// PDAAM on the QROT dual (src/qrot_pdaam.cpp:14-188): lambda = -gamma; per
// iteration L_new = L / 2, doubled per failed test (at most 20 trials, the
// last one accepted); alpha_new = (1 + sqrt(4 alpha^2 L L_new + 1)) / (2
// L_new), tau = 1 / (alpha_new L_new), lambda~ = (1 - tau) lambda + tau
// lambda'; at gamma = -lambda~ the block with the larger gradient norm is
// minimised exactly (optimal_alpha or optimal_beta) and accepted when phi_new
// >= phi~ + ||g||^2 / (2 L_new reg^2); then lambda' += (alpha / reg) g and
// the device plan P (kept in the result) becomes (alpha_old / alpha)^2 (L_old
// / L) P + tau plan(-lambda~) in one pass with its sums.  Histories and the
// stop test as APDAGD; the dual is -lambda~.
static void solve_qrot_pdaam(regot_problem* p, double tol, int max_iter, int verbose,
                             const regot_opts& o, regot_result* res)
{
    const int n = p->n, m = p->m;
    const long N = static_cast<long>(n) + m;
    hipStream_t s = p->stream;
    const double reg = p->reg;
    Work w(p, N);
    w.og = [&](const double* x, double* g, double* gn2) {
        return qrot_obj_grad(p, x, g, 0.0, gn2);
    };
    if (o.x0) {
        HIP_CHECK(hipMemcpyAsync(w.gamma, o.x0, N * sizeof(double), hipMemcpyHostToDevice, s));
    } else {
        launch_fill(w.gamma, 0.0, n, s);
        qrot_optimal(p, true, w.gamma, w.gamma.p + n);
    }
    double* dual = w.x.p;
    double* dualp = w.xg.p;
    double* dualt = w.rhs.p;
    double* gamma = w.gamma.p;
    double* grad = w.g.p;
    vec_axpby(dual, -1.0, gamma, 0.0, nullptr, N, s);
    vec_copy(dualp, dual, N, s);
    vec_copy(dualt, dual, N, s);

    QrotScratch& sc = scratch(p);
    const int rows = qrot_rows(n, m);
    const long nstrip = ceil_div(m, 256), nblk = ceil_div(n, rows);
    DBuf<double> wpart(2 * nstrip * nblk + ceil_div(n > m ? n : m, 256)), pout(3);
    HBuf<double> hp;
    hp.ensure(3);
    res->plan.alloc(static_cast<size_t>(n) * m);
    res->plan_ld = m;
    res->plan_device = p->device;
    auto plan_pass = [&](double cold, double tau, bool init) {
        launch_qrot_apd_plan(p->bulk, p->dM, p->ldm, n, m, gamma, gamma + n, reg, cold, tau, init,
                             res->plan.p, m, p->a, p->b, sc.rowpart, sc.colpart, sc.rsum, sc.csum,
                             wpart, pout, s);
        HIP_CHECK(hipMemcpyAsync(hp.p, pout.p, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        w.alg_bytes += w.pass_bytes() + 16.0 * n * m;
    };
    // ||grad_alpha||^2 and ||grad_beta||^2
    auto block_norms = [&](double* lr, double* mr) {
        const double* xs[1] = {grad};
        vec_dots(xs, xs, n, 1, w.dscal.p, s);
        const double* xt[1] = {grad + n};
        vec_dots(xt, xt, m, 1, w.dscal.p + 1, s);
        HIP_CHECK(hipMemcpyAsync(w.hscal.p, w.dscal.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        *lr = w.hscal.p[0];
        *mr = w.hscal.p[1];
    };
    plan_pass(0.0, 1.0, true);
    res->qrot = true;
    Timer timer;
    timer.tic();
    double gnorm;
    double f = w.obj_grad(gamma, grad, &gnorm);
    double dur = timer.toc();
    res->obj_vals.push_back(f);
    res->prim_vals.push_back(qrot_round_plan(p, gamma, gamma + n, nullptr, 0));
    res->mar_errs.push_back(gnorm / reg);
    res->run_times.push_back(dur);
    double L = 1.0, palpha = 0.0;
    int it;
    for (it = 0; it < max_iter; it++) {
        timer.tic();
        double L_new = 0.5 * L, phi_new = 0.0, c_old = 0.0, c_new = 0.0;
        for (int j = 0; j < 20; j++) {
            const double palpha_new =
                (1.0 + std::sqrt(4.0 * palpha * palpha * L * L_new + 1.0)) / (2.0 * L_new);
            const double ptau = 1.0 / (palpha_new * L_new);
            vec_axpby(dualt, 1.0 - ptau, dual, ptau, dualp, N, s);
            vec_axpby(gamma, -1.0, dualt, 0.0, nullptr, N, s);
            f = w.obj_grad(gamma, grad, nullptr);
            const double phit = -f / reg;
            double lr, mr;
            block_norms(&lr, &mr);
            if (lr > mr)
                qrot_optimal(p, false, gamma + n, gamma);
            else
                qrot_optimal(p, true, gamma, gamma + n);
            f = w.obj(gamma);
            phi_new = -f / reg;
            const double rhs = phit + (lr + mr) / (2.0 * L_new * reg * reg);
            if (phi_new >= rhs || j >= 19) {
                const double ar = palpha / palpha_new, lrat = L / L_new;
                c_old = ar * ar * lrat;
                c_new = ptau;
                palpha = palpha_new;
                L = L_new;
                vec_axpby(dual, -1.0, gamma, 0.0, nullptr, N, s);
                break;
            }
            L_new *= 2.0;
        }
        vec_axpby(dualp, 1.0, dualp, palpha / reg, grad, N, s);
        vec_axpby(gamma, -1.0, dualt, 0.0, nullptr, N, s);
        plan_pass(c_old, c_new, false);
        const double primal_obj = hp.p[0] + 0.5 * reg * hp.p[1];
        const double mar_err = std::sqrt(hp.p[2]);
        dur = timer.toc();
        res->obj_vals.push_back(f);
        res->prim_vals.push_back(qrot_round_plan(p, gamma, gamma + n, nullptr, 0));
        res->mar_errs.push_back(mar_err);
        res->run_times.push_back(res->run_times.back() + dur);
        if (verbose >= 1) {
            printf("iter = %d, primal_obj = %g, phi_new = %g, mar_err = %g\n", it, primal_obj,
                   phi_new, mar_err);
            fflush(stdout);
        }
        if (primal_obj - phi_new <= tol && mar_err <= tol) break;
    }
    res->niter = it;
    w.report(res);
    res->dual.resize(N);
    HIP_CHECK(hipMemcpy(res->dual.data(), gamma, N * sizeof(double), hipMemcpyDeviceToHost));
}

// Drivers.  Semantics: src/qrot_bcd.cpp:15-101 (tau forced to 0),
// src/qrot_gd.cpp:15-121, src/qrot_lbfgs_dual.cpp:85-126 (epsilon =
// epsilon_rel = tol).  Histories: obj_vals f, prim_vals the primal value of
// the rounded plan, mar_errs ||grad|| / reg, run_times cumulative ms; niter + 1
// entries for BCD and GD (test ||grad|| < tol before each update), one per
// iterate() call for L-BFGS.
// L-BFGS on the semi-dual (src/qrot_lbfgs_semi_dual.cpp:78-118): over alpha
// alone, from alpha = 0 or x0's head; every evaluation finds beta =
// optimal_beta(alpha) and takes the dual objective there (tau = 0) with the
// gradient's alpha block; every iterate records the primal value at [alpha;
// optimal_beta(alpha)]; the result is [alpha; optimal_beta(alpha)].
static void solve_qrot_semi_dual(regot_problem* p, double tol, int max_iter, int verbose,
                                 const regot_opts& o, regot_result* res)
{
    const int n = p->n, m = p->m;
    const long N = static_cast<long>(n) + m;
    hipStream_t s = p->stream;
    const double reg = p->reg;
    Work w(p, n);
    DBuf<double> gfull(N), gtmp(N), gpr(N);
    w.og = [&](const double* x, double* g, double* gn2) {
        vec_copy(gfull, x, n, s);
        qrot_optimal(p, true, gfull, gfull.p + n);
        const double f = qrot_obj_grad(p, gfull, g ? gtmp.p : nullptr, 0.0, nullptr);
        if (g) {
            vec_copy(g, gtmp, n, s);
            if (gn2) *gn2 = w.dot(g, g);
        }
        return f;
    };
    if (o.x0)
        HIP_CHECK(hipMemcpyAsync(w.gamma, o.x0, n * sizeof(double), hipMemcpyHostToDevice, s));
    else
        launch_fill(w.gamma, 0.0, n, s);
    res->qrot = true;
    Timer timer;
    timer.tic();
    auto iterate = [&](const double* x, double last_f, double gnorm) {
        const double dur = timer.toc();
        vec_copy(gpr, x, n, s);
        qrot_optimal(p, true, gpr, gpr.p + n);
        const double prim = qrot_round_plan(p, gpr, gpr.p + n, nullptr, 0);
        res->obj_vals.push_back(last_f);
        res->prim_vals.push_back(prim);
        res->mar_errs.push_back(gnorm / reg);
        res->run_times.push_back((res->run_times.empty() ? 0.0 : res->run_times.back()) + dur);
        if (verbose >= 1) {
            printf("iter = %d, objval = %g, ||grad|| = %g\n",
                   static_cast<int>(res->obj_vals.size()) - 1, last_f, gnorm);
            fflush(stdout);
        }
        timer.tic();
    };
    const double* xfinal = w.gamma.p;
    res->niter = lbfgs_minimize(w, tol, tol, max_iter, iterate, &xfinal);
    vec_copy(gfull, xfinal, n, s);
    qrot_optimal(p, true, gfull, gfull.p + n);
    w.report(res);
    res->dual.resize(N);
    HIP_CHECK(hipMemcpy(res->dual.data(), gfull.p, N * sizeof(double), hipMemcpyDeviceToHost));
}

// APDAGD on the QROT dual (src/qrot_apdagd.cpp:14-181): dual variables
// lambda = -gamma, L halved per iteration and doubled per failed test (at most
// 20 inner trials, the last one accepted); the primal plan P (device, n x m
// FP64, kept in the result) starts at plan(gamma_0) and becomes (1 - tau) P +
// tau plan(-lambda~) in one fused pass with its sums; stop when <M, P> +
// reg/2 ||P||^2 - phi <= tol and the full marginal error <= tol.  History:
// f of the last objective evaluated, the primal value of the rounded plan at
// -lambda~, mar_errs the full marginal error of P (the first entry ||g|| /
// reg); niter + 2 entries after a stop.  The dual is -lambda~.
static void solve_qrot_apdagd(regot_problem* p, double tol, int max_iter, int verbose,
                              const regot_opts& o, regot_result* res)
{
    const int n = p->n, m = p->m;
    const long N = static_cast<long>(n) + m;
    hipStream_t s = p->stream;
    const double reg = p->reg;
    Work w(p, N);
    w.og = [&](const double* x, double* g, double* gn2) {
        return qrot_obj_grad(p, x, g, 0.0, gn2);
    };
    if (o.x0) {
        HIP_CHECK(hipMemcpyAsync(w.gamma, o.x0, N * sizeof(double), hipMemcpyHostToDevice, s));
    } else {
        launch_fill(w.gamma, 0.0, n, s);
        qrot_optimal(p, true, w.gamma, w.gamma.p + n);
    }
    double* dual = w.x.p;
    double* dualp = w.xg.p;
    double* dualt = w.rhs.p;
    double* dual_new = w.y.p;
    double* dualp_new = w.sv.p;
    double* gamma = w.gamma.p;
    double* grad = w.g.p;
    double* diff = w.step.p;
    vec_axpby(dual, -1.0, gamma, 0.0, nullptr, N, s);
    vec_copy(dualp, dual, N, s);
    vec_copy(dualt, dual, N, s);

    QrotScratch& sc = scratch(p);
    const int rows = qrot_rows(n, m);
    const long nstrip = ceil_div(m, 256), nblk = ceil_div(n, rows);
    DBuf<double> wpart(2 * nstrip * nblk + ceil_div(n > m ? n : m, 256)), pout(3);
    HBuf<double> hp;
    hp.ensure(3);
    res->plan.alloc(static_cast<size_t>(n) * m);
    res->plan_ld = m;
    res->plan_device = p->device;
    auto plan_pass = [&](double tau, bool init) {
        launch_qrot_apd_plan(p->bulk, p->dM, p->ldm, n, m, gamma, gamma + n, reg, 1.0 - tau,
                             tau, init,
                             res->plan.p, m, p->a, p->b, sc.rowpart, sc.colpart, sc.rsum, sc.csum,
                             wpart, pout, s);
        HIP_CHECK(hipMemcpyAsync(hp.p, pout.p, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        w.alg_bytes += w.pass_bytes() + 16.0 * n * m;
    };
    plan_pass(1.0, true);
    res->qrot = true;
    Timer timer;
    timer.tic();
    double gnorm;
    double f = w.obj_grad(gamma, grad, &gnorm);
    double dur = timer.toc();
    res->obj_vals.push_back(f);
    res->prim_vals.push_back(qrot_round_plan(p, gamma, gamma + n, nullptr, 0));
    res->mar_errs.push_back(gnorm / reg);
    res->run_times.push_back(dur);
    double L = 1.0, pbeta = 0.0;
    int it;
    for (it = 0; it < max_iter; it++) {
        timer.tic();
        L *= 0.5;
        double ptau = 0.0, phi_new = 0.0;
        for (int j = 0; j < 20; j++) {
            const double palpha = (1.0 + std::sqrt(4.0 * L * pbeta + 1.0)) / (2.0 * L);
            const double pbeta_new = pbeta + palpha;
            ptau = palpha / pbeta_new;
            vec_axpby(dualt, 1.0 - ptau, dual, ptau, dualp, N, s);
            vec_axpby(gamma, -1.0, dualt, 0.0, nullptr, N, s);
            f = w.obj_grad(gamma, grad, nullptr);
            const double phit = -f / reg;
            vec_axpby(dualp_new, 1.0, dualp, palpha / reg, grad, N, s);
            vec_axpby(dual_new, 1.0 - ptau, dual, ptau, dualp_new, N, s);
            vec_axpby(gamma, -1.0, dual_new, 0.0, nullptr, N, s);
            f = w.obj(gamma);
            phi_new = -f / reg;
            vec_axpby(diff, 1.0, dual_new, -1.0, dualt, N, s);
            const double* xs[2] = {grad, diff};
            const double* ys[2] = {diff, diff};
            double d[2];
            w.dots(2, xs, ys, d);
            const double rhs = phit + d[0] / reg - 0.5 * L * d[1];
            if (phi_new >= rhs || j >= 19) {
                pbeta = pbeta_new;
                std::swap(dualp, dualp_new);
                std::swap(dual, dual_new);
                break;
            }
            L *= 2.0;
        }
        vec_axpby(gamma, -1.0, dualt, 0.0, nullptr, N, s);
        plan_pass(ptau, false);
        const double primal_obj = hp.p[0] + 0.5 * reg * hp.p[1];
        const double mar_err = std::sqrt(hp.p[2]);
        dur = timer.toc();
        res->obj_vals.push_back(f);
        res->prim_vals.push_back(qrot_round_plan(p, gamma, gamma + n, nullptr, 0));
        res->mar_errs.push_back(mar_err);
        res->run_times.push_back(res->run_times.back() + dur);
        if (verbose >= 1) {
            printf("iter = %d, primal_obj = %g, phi_new = %g, mar_err = %g\n", it, primal_obj,
                   phi_new, mar_err);
            fflush(stdout);
        }
        if (primal_obj - phi_new <= tol && mar_err <= tol) break;
    }
    res->niter = it;
    w.report(res);
    res->dual.resize(N);
    HIP_CHECK(hipMemcpy(res->dual.data(), gamma, N * sizeof(double), hipMemcpyDeviceToHost));
}

void solve_qrot(regot_problem* p, int solver, double tol, int max_iter, int verbose,
                const regot_opts& o, regot_result* res)
{
    if (solver == 3) {
        solve_qrot_semi_dual(p, tol, max_iter, verbose, o, res);
        return;
    }
    if (solver == 4) {
        solve_qrot_apdagd(p, tol, max_iter, verbose, o, res);
        return;
    }
    if (solver == 5) {
        solve_qrot_grssn(p, tol, max_iter, verbose, o, res);
        return;
    }
    if (solver == 6) {
        solve_qrot_assn(p, tol, max_iter, verbose, o, res);
        return;
    }
    if (solver == 7) {
        solve_qrot_s5n(p, tol, max_iter, verbose, o, res);
        return;
    }
    if (solver == 8) {
        solve_qrot_pdaam(p, tol, max_iter, verbose, o, res);
        return;
    }
    const int n = p->n, m = p->m;
    const long N = static_cast<long>(n) + m;
    hipStream_t s = p->stream;
    const double tau = solver == 0 ? 0.0 : o.tau;
    const double reg = p->reg;
    Work w(p, N);
    w.og = [&](const double* x, double* g, double* gn2) {
        return qrot_obj_grad(p, x, g, tau, gn2);
    };
    if (o.x0) {
        HIP_CHECK(hipMemcpyAsync(w.gamma, o.x0, N * sizeof(double), hipMemcpyHostToDevice, s));
    } else {
        launch_fill(w.gamma, 0.0, n, s);
        qrot_optimal(p, true, w.gamma, w.gamma.p + n);
    }
    res->qrot = true;
    Timer timer;
    timer.tic();
    auto record = [&](const double* x, double f, double gnorm) {
        const double prim = qrot_round_plan(p, x, x + n, nullptr, 0);
        const double dur = timer.toc();
        res->obj_vals.push_back(f);
        res->prim_vals.push_back(prim);
        res->mar_errs.push_back(gnorm / reg);
        res->run_times.push_back((res->run_times.empty() ? 0.0 : res->run_times.back()) + dur);
        timer.tic();
    };
    const double* xfinal = w.gamma.p;
    if (solver == 2) {
        auto iterate = [&](const double* x, double last_f, double gnorm) {
            record(x, last_f, gnorm);
            if (verbose >= 1) {
                printf("iter = %d, objval = %g, ||grad|| = %g\n",
                       static_cast<int>(res->obj_vals.size()) - 1, last_f, gnorm);
                fflush(stdout);
            }
        };
        res->niter = lbfgs_minimize(w, tol, tol, max_iter, iterate, &xfinal);
    } else {
        double gnorm;
        double f = w.obj_grad(w.gamma, w.g, &gnorm);
        record(w.gamma, f, gnorm);
        int it;
        for (it = 0; it < max_iter; it++) {
            if (verbose >= 1) {
                printf("iter = %d, objval = %g, ||grad|| = %g\n", it, f, gnorm);
                fflush(stdout);
            }
            if (gnorm < tol) break;
            if (solver == 0) {
                qrot_optimal(p, false, w.gamma.p + n, w.gamma.p);
                qrot_optimal(p, true, w.gamma.p, w.gamma.p + n);
            } else {
                // step 1 halved until f(gamma - step g) <= f - step 0.5 ||g||^2,
                // at most 20 trials, the last one taken
                double step = 1.0;
                const double thresh = 0.5 * gnorm * gnorm;
                for (int k = 0; k < 20; k++) {
                    w.move(w.x, w.gamma, -step, w.g);
                    const double fn = w.obj(w.x);
                    if (fn <= f - step * thresh) break;
                    step *= 0.5;
                }
                vec_copy(w.gamma, w.x, N, s);
            }
            f = w.obj_grad(w.gamma, w.g, &gnorm);
            record(w.gamma, f, gnorm);
        }
        res->niter = it;
    }
    w.report(res);
    res->dual.resize(N);
    HIP_CHECK(hipMemcpy(res->dual.data(), xfinal, N * sizeof(double), hipMemcpyDeviceToHost));
}

}  // namespace regot

using namespace regot;

#define CAPI_BEGIN try {
#define CAPI_END                                                                          \
    }                                                                                     \
    catch (const regot::Error& e) { return set_error(e.code, e.what()); }                 \
    catch (const std::bad_alloc&) { return set_error(E_NOMEM, "host allocation failed"); } \
    catch (const std::exception& e) { return set_error(E_INTERNAL, e.what()); }            \
    return OK;

namespace regot {
// QROT plan of a host dual (n + m) into host memory, in row chunks.
void qrot_plan_host(regot_problem* p, const double* dual, double* plan)
{
    const int n = p->n, m = p->m;
    hipStream_t s = p->stream;
    DBuf<double> g(static_cast<size_t>(n) + m);
    HIP_CHECK(hipMemcpyAsync(g.p, dual, (static_cast<size_t>(n) + m) * sizeof(double),
                             hipMemcpyHostToDevice, s));
    const long chunk = std::max<long>(1, (64L << 20) / (8L * m));
    DBuf<double> buf(static_cast<size_t>(std::min<long>(chunk, n)) * m);
    for (long r0 = 0; r0 < n; r0 += chunk) {
        const int nr = static_cast<int>(std::min<long>(chunk, n - r0));
        launch_qrot_plan(p->bulk, p->dM, p->ldm, n, m, g.p, g.p + n, p->reg, buf.p, r0, nr, s);
        HIP_CHECK(hipMemcpyAsync(plan + r0 * m, buf.p, sizeof(double) * nr * m,
                                 hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
}
}  // namespace regot

extern "C" {

int regot_qrot_solve(regot_problem* p, int solver, double tol, int max_iter, int verbose,
                     const regot_opts* opts, regot_result** out)
{
    CAPI_BEGIN
    if (!p || !out) fail(E_INVALID, "null argument");
    *out = nullptr;
    if (solver < 0 || solver > 8) fail(E_INVALID, "unknown QROT solver %d", solver);
    if (max_iter < 0) fail(E_INVALID, "max_iter must be >= 0");
    HIP_CHECK(hipSetDevice(p->device));
    regot_opts o;
    if (opts) {
        o = *opts;
    } else {
        o = regot_opts{};
        regot_opts_default(&o, 0);
        o.method = 0;   // QROTSolverOpts' defaults: CG, mu0 = 1
        o.mu0 = 1.0;
    }
    if (solver == 7 && !(o.mu0 > 0.0)) fail(E_INVALID, "mu0 must be > 0");
    if (solver == 5 && !(o.shift >= 0.0)) fail(E_INVALID, "shift must be >= 0");
    auto res = std::make_unique<regot_result>();
    res->solver = 100 + solver;
    res->n = p->n;
    res->m = p->m;
    solve_qrot(p, solver, tol, max_iter, verbose, o, res.get());
    *out = res.release();
    CAPI_END
}

int regot_qrot_obj_grad(regot_problem* p, const double* gamma, double tau, double* f,
                        double* grad)
{
    CAPI_BEGIN
    if (!p || !gamma || !f) fail(E_INVALID, "null argument");
    HIP_CHECK(hipSetDevice(p->device));
    const long N = static_cast<long>(p->n) + p->m;
    DBuf<double> g(N), gr(N);
    HIP_CHECK(hipMemcpyAsync(g.p, gamma, N * sizeof(double), hipMemcpyHostToDevice, p->stream));
    *f = qrot_obj_grad(p, g.p, grad ? gr.p : nullptr, tau, nullptr);
    if (grad) HIP_CHECK(hipMemcpy(grad, gr.p, N * sizeof(double), hipMemcpyDeviceToHost));
    CAPI_END
}

int regot_qrot_optimal_beta(regot_problem* p, const double* alpha, double* beta)
{
    CAPI_BEGIN
    if (!p || !alpha || !beta) fail(E_INVALID, "null argument");
    HIP_CHECK(hipSetDevice(p->device));
    DBuf<double> a(p->n), b(p->m);
    HIP_CHECK(hipMemcpyAsync(a.p, alpha, p->n * sizeof(double), hipMemcpyHostToDevice, p->stream));
    qrot_optimal(p, true, a.p, b.p);
    HIP_CHECK(hipMemcpy(beta, b.p, p->m * sizeof(double), hipMemcpyDeviceToHost));
    CAPI_END
}

int regot_qrot_optimal_alpha(regot_problem* p, const double* beta, double* alpha)
{
    CAPI_BEGIN
    if (!p || !alpha || !beta) fail(E_INVALID, "null argument");
    HIP_CHECK(hipSetDevice(p->device));
    DBuf<double> a(p->n), b(p->m);
    HIP_CHECK(hipMemcpyAsync(b.p, beta, p->m * sizeof(double), hipMemcpyHostToDevice, p->stream));
    qrot_optimal(p, false, b.p, a.p);
    HIP_CHECK(hipMemcpy(alpha, a.p, p->n * sizeof(double), hipMemcpyDeviceToHost));
    CAPI_END
}

int regot_qrot_primal_val(regot_problem* p, const double* gamma, double* value)
{
    CAPI_BEGIN
    if (!p || !gamma || !value) fail(E_INVALID, "null argument");
    HIP_CHECK(hipSetDevice(p->device));
    const long N = static_cast<long>(p->n) + p->m;
    DBuf<double> g(N);
    HIP_CHECK(hipMemcpyAsync(g.p, gamma, N * sizeof(double), hipMemcpyHostToDevice, p->stream));
    *value = qrot_round_plan(p, g.p, g.p + p->n, nullptr, 0);
    CAPI_END
}

int regot_qrot_hess_solve(regot_problem* p, const double* gamma, const double* rhs,
                          double shift, double tau, double tol, int max_iter, double* out,
                          int* cg_iters)
{
    CAPI_BEGIN
    if (!p || !gamma || !rhs || !out) fail(E_INVALID, "null argument");
    if (!(shift >= 0.0) || !(tau >= 0.0)) fail(E_INVALID, "shift and tau must be >= 0");
    HIP_CHECK(hipSetDevice(p->device));
    const long N = static_cast<long>(p->n) + p->m;
    DBuf<double> g(N), r(N), x(N);
    HIP_CHECK(hipMemcpyAsync(g.p, gamma, N * sizeof(double), hipMemcpyHostToDevice, p->stream));
    HIP_CHECK(hipMemcpyAsync(r.p, rhs, N * sizeof(double), hipMemcpyHostToDevice, p->stream));
    QrotHess& H = qrot_hess_build(p, g.p);
    const int it = qrot_hess_solve(p, H, r.p, shift, tau, tol,
                                   max_iter > 0 ? max_iter : 2 * p->m, x.p);
    HIP_CHECK(hipMemcpy(out, x.p, N * sizeof(double), hipMemcpyDeviceToHost));
    if (cg_iters) *cg_iters = it;
    CAPI_END
}

int regot_result_prim_vals(const regot_result* r, double* out)
{
    CAPI_BEGIN
    if (!r || !out) fail(E_INVALID, "null argument");
    std::copy(r->prim_vals.begin(), r->prim_vals.end(), out);
    CAPI_END
}

int regot_result_dual_len(const regot_result* r)
{
    return r ? static_cast<int>(r->dual.size()) : -1;
}

}  // extern "C"
