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
// The iterates decrease monotonically, so at most count + 1 steps run; the
// root is (reg w + sum of the active y) / (their count), the formula of
// piecewise_linear_equation, src/qrot_problem.cpp:365-394.
void qrot_optimal(regot_problem* p, bool cols, const double* pot, double* out)
{
    const int n = p->n, m = p->m;
    hipStream_t s = p->stream;
    QrotScratch& w = scratch(p);
    const int count = cols ? m : n;
    launch_qrot_pl_init(out, w.kprev, w.done, count, w.open, s);
    const double* wt = cols ? p->b.p : p->a.p;
    for (int step = 0; step <= count + 1; step++) {
        launch_qrot_pl_step(p->bulk, p->dM, p->ldm, n, m, cols, pot, wt, p->reg, out, w.kprev,
                            w.done, w.spart, w.kpart, w.open, s);
        HIP_CHECK(hipMemcpyAsync(w.hopen.p, w.open.p, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        if (w.hopen.p[0] == 0) return;
    }
    fail(E_INTERNAL, "QROT root iteration did not settle");
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
        launch_qrot_apd_plan(p->bulk, p->dM, p->ldm, n, m, gamma, gamma + n, reg, tau, init,
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
    if (solver < 0 || solver > 4) fail(E_INVALID, "unknown QROT solver %d", solver);
    if (max_iter < 0) fail(E_INVALID, "max_iter must be >= 0");
    HIP_CHECK(hipSetDevice(p->device));
    regot_opts o;
    if (opts) {
        o = *opts;
    } else {
        o = regot_opts{};
        regot_opts_default(&o, 0);
    }
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
