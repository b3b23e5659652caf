// newton.cpp -- host drivers of the Newton-type EROT solvers (SSNS, sparse
// Newton, SPLR, and Newton with the full Hessian).  Control flow, schedules and line searches run on the host
// in FP64 and mirror the reference drivers step for step; every vector lives
// on the device, every O(n m) evaluation and every Hessian product or CG
// solve is a HIP kernel, and only scalars come back to the host.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

#include "driver_work.hpp"

namespace regot {
namespace {

// Minimiser of the quadratic through fa, ga (at a) and fb (at b);
// src/sinkhorn_problem_ls.cpp:72-77.
double quad_fg(double a, double b, double fa, double ga, double fb)
{
    const double ba = b - a;
    const double w = 0.5 * ba * ga / (fa - fb + ba * ga);
    return a + w * ba;
}

// Minimiser of the quadratic with slopes ga (at a) and gb (at b); :84-88.
double quad_gg(double a, double b, double ga, double gb)
{
    const double w = ga / (ga - gb);
    return a + w * (b - a);
}

// Local minimiser of the cubic through fa, ga, fb, gb; :93-153.
double cubic_min(double a, double b, double fa, double fb, double ga, double gb, bool& exists)
{
    const double apb = a + b, ba = b - a, ba2 = ba * ba;
    const double fba = fb - fa, gba = gb - ga;
    const double z3 = (ga + gb) * ba - 2.0 * fba;
    const double z2 = 0.5 * (gba * ba2 - 3.0 * apb * z3);
    const double z1 = fba * ba2 - apb * z2 - (a * apb + b * b) * z3;
    const double eps = std::numeric_limits<double>::epsilon();
    if (std::fabs(z3) < eps * std::fabs(z2) || std::fabs(z3) < eps * std::fabs(z1)) {
        exists = (z2 * ba > 0.0);
        return exists ? (-0.5 * z1 / z2) : b;
    }
    const double u = z2 / (3.0 * z3), v = z1 / z2;
    const double vu = v / u;
    exists = (vu <= 1.0);
    if (!exists) return b;
    double r1, r2;
    if (std::fabs(u) >= std::fabs(v)) {
        const double w = 1.0 + std::sqrt(1.0 - vu);
        r1 = -u * w;
        r2 = -v / w;
    } else {
        const double sd = std::sqrt(std::fabs(u)) * std::sqrt(std::fabs(v)) * std::sqrt(1 - u / v);
        r1 = -u - sd;
        r2 = -u + sd;
    }
    return (z3 * ba > 0.0) ? std::max(r1, r2) : std::min(r1, r2);
}

// Next trial from the bracket (Case 1/3 of More-Thuente); :157-223.
double step_select(double al, double au, double at, double fl, double fu, double ft, double gl,
                   double gu, double gt)
{
    if (al == au) return al;
    if (!std::isfinite(ft) || !std::isfinite(gt)) return (al + at) / 2.0;
    bool ac_ok;
    const double ac = cubic_min(al, at, fl, ft, gl, gt, ac_ok);
    const double aq = quad_fg(al, at, fl, gl, ft);
    if (ft > fl) {
        if (!ac_ok) return aq;
        return (std::fabs(ac - al) < std::fabs(aq - al)) ? ac : ((aq + ac) / 2.0);
    }
    const double as = quad_gg(al, at, gl, gt);
    if (gt * gl < 0.0) return (std::fabs(ac - at) >= std::fabs(as - at)) ? ac : as;
    const double dl = 1.1, du = 0.66;
    if (std::fabs(gt) < std::fabs(gl)) {
        const bool take_ac =
            ac_ok && ((ac - at) * (at - al) > 0.0) && (std::fabs(ac - at) < std::fabs(as - at));
        const double res = take_ac ? ac : as;
        const double lim = at + du * (au - at);
        return (at > al) ? std::min(lim, res) : std::max(lim, res);
    }
    if (!std::isfinite(au) || !std::isfinite(fu) || !std::isfinite(gu)) return at + dl * (at - al);
    bool ae_ok;
    const double ae = cubic_min(at, au, ft, fu, gt, gu, ae_ok);
    const double lim = at + du * (au - at);
    return (at > al) ? std::min(lim, ae) : std::max(lim, ae);
}

// Strong-Wolfe search (c1 = 1e-4, c2 = 0.9, at most max_iter trials -- the
// reference's 20 when max_iter <= 0 -- steps in [1e-10, 2]), the
// More-Thuente variant of Problem::line_search_wolfe,
// src/sinkhorn_problem_ls.cpp:226-470.  Each trial is one dual_obj_grad pass
// at gamma + step d into w.x / w.xg.  Returns the step; *reuse_last is the
// reference's recompute_T == false: the caller keeps the last trial's T (its
// f in *f_last, gradient in w.xg, norm in *gn_last, taken at *step_last),
// which is the returned step's except when the trials ran out.
double wolfe(Work& w, double init_step, double f0, double dg0, int max_iter, bool* reuse_last,
             double* f_last, double* gn_last, double* step_last)
{
    const double c1 = 1e-4, c2 = 0.9;
    if (max_iter <= 0) max_iter = 20;
    const double step_min = 1e-10, step_max = 2.0;
    *reuse_last = false;
    init_step = std::min(std::max(init_step, step_min), step_max);
    double step = init_step;
    double fx = f0;
    if (dg0 > 0.0) return init_step;
    const double test_decr = c1 * dg0, test_curv = -c2 * dg0;
    const double inf = std::numeric_limits<double>::infinity();
    double I_lo = 0.0, I_hi = inf, fI_lo = 0.0, fI_hi = inf;
    double gI_lo = (1.0 - c1) * dg0, gI_hi = inf, psiI_lo = 0.0;
    bool bracketed = false, safeguard = true;
    double I_width = inf, I_width_prev = inf;
    int shrink_fail = 0;
    const double delta_max = 1.1, delta_min = 7.0 / 12, shrink = 0.66;
    for (int it = 0; it < max_iter; it++) {
        w.move(w.x, w.gamma, step, w.direc);
        fx = w.obj_grad(w.x, w.xg, gn_last);
        const double dg = w.dot(w.xg, w.direc);
        *step_last = step;
        *f_last = fx;
        const double psit = fx - f0 - step * test_decr;
        const double dpsit = dg - test_decr;
        if (psit <= 0.0 && std::fabs(dg) <= test_curv) {
            *reuse_last = true;
            return step;
        }
        if (step <= step_min && (psit > 0.0 || dpsit >= 0.0)) {
            *reuse_last = true;
            return step;
        }
        if (step >= step_max && (psit <= 0.0 && dpsit < 0.0)) {
            *reuse_last = true;
            return step;
        }
        const double ft = psit, gt = dpsit;
        if (safeguard && (psit <= 0.0 && dpsit < 0.0)) safeguard = false;
        double new_step;
        const bool case2 = (psit <= psiI_lo) && (dpsit * (I_lo - step) > 0.0);
        if (case2) {
            new_step = std::min(step_max, step + delta_max * (step - I_lo));
        } else {
            new_step = step_select(I_lo, I_hi, step, fI_lo, fI_hi, ft, gI_lo, gI_hi, gt);
            new_step = std::min(std::max(new_step, step_min), step_max);
            if (safeguard) {
                const double upper = std::max(step_min, delta_min * step);
                new_step = std::min(std::max(new_step, step_min), upper);
            }
        }
        if (psit > psiI_lo) {
            I_hi = step;
            fI_hi = ft;
            gI_hi = gt;
        } else if (case2) {
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
        if (!bracketed && !case2)
            bracketed = (std::min(I_lo, I_hi) >= step_min && std::max(I_lo, I_hi) <= step_max);
        if (bracketed) {
            I_width_prev = I_width;
            I_width = std::fabs(I_hi - I_lo);
            if (I_width_prev < inf && I_width > shrink * I_width_prev)
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
    // Trials exhausted (:448-469): the step the last update produced if
    // psi(last evaluated f, that step) <= psi(I_lo), else I_lo (or the initial
    // step when I_lo is not positive).  As in the reference, the first answer
    // keeps the last trial's T (recompute_T stays false there) even though the
    // returned step was not itself evaluated.
    const double psit = fx - f0 - step * test_decr;
    if (psit <= psiI_lo) {
        *reuse_last = true;
        return step;
    }
    if (I_lo <= 0.0) return init_step;
    return I_lo;
}


}  // namespace

// SSNS.  Mirrors sinkhorn_ssns_internal, src/sinkhorn_ssns.cpp:16-234:
// trust-region-like mu control (rho thresholds 0.25 / 0.75, mu x4 / x0.5,
// floor 1e-3), shift = mu ||g||, step candidates {1, 0.5, 0.25, 0.1},
// delta = 0.01 ||g||, threshold sparsification with the previous density as
// the hint.  After an accepted step the gradient and the sparsified Hessian
// come from the T of the last candidate evaluated, and f from that T with the
// accepted potentials, as the reference's dual_obj_grad(gamma, g, T, false)
// after line_selection (src/sinkhorn_ssns.cpp:185-187).  That candidate is
// the accepted one whenever a candidate lowered f; otherwise the last is 0.1,
// which a convex dual makes the best of the four up to exact ties.
void solve_ssns(regot_problem* p, double tol, int max_iter, int verbose, const regot_opts& o,
                regot_result* res)
{
    const double rho_t1 = 0.25, rho_t2 = 0.75, mu_s1 = 4.0, mu_s2 = 0.5, kappa = 0.001,
                 nu0 = 0.01;
    const double cand[4] = {1.0, 0.5, 0.25, 0.1};
    PreciseScope ps(p);
    Work w(p);
    w.init(o);
    double mu = o.mu0, rho = 1.0;
    Timer timer;
    timer.tic();
    double gnorm;
    double f = w.obj_grad(w.gamma, w.g, &gnorm);
    double delta = nu0 * gnorm;
    w.threshold(delta, 0.001);
    res->obj_vals.push_back(f);
    res->mar_errs.push_back(gnorm);
    res->run_times.push_back(timer.toc());
    res->densities.push_back(hess_of(p).density());
    int i;
    for (i = 0; i < max_iter; i++) {
        if (verbose >= 1) {
            printf("iter = %d, objval = %g, ||grad|| = %g\n", i, f, gnorm);
            fflush(stdout);
        }
        timer.tic();
        if (gnorm < tol) break;
        const double shift = mu * gnorm;
        if (verbose >= 2) {
            printf("[params]===================================================\n"
                   "║ rho = %g, mu = %g, shift = %g\n"
                   "║----------------------------------------------------------\n",
                   rho, mu, shift);
        }
        Timer inner;
        inner.tic();
        double t_lin = 0.0, t_ls = 0.0, t_rho = 0.0, t_grad = 0.0, t_hess = 0.0;
        vec_axpby(w.rhs, -1.0, w.g, 0.0, nullptr, w.N, w.s);
        w.lin_solve(o, w.rhs, shift, w.direc);
        t_lin = inner.lap();
        // line selection, src/sinkhorn_problem_ls.cpp:13-41
        double best = 1.0, newf = std::numeric_limits<double>::infinity(), c_last = 1.0;
        for (double a : cand) {
            w.move(w.x, w.gamma, a, w.direc);
            const double fa = w.obj(w.x);
            c_last = a;
            if (fa < newf) {
                best = a;
                newf = fa;
            }
            if (newf < f) break;
        }
        t_ls = inner.lap();
        vec_axpby(w.step, best, w.direc, 0.0, nullptr, w.N, w.s);
        hess_apply_hs(p, w.step, w.Hstep);
        const double* xs[3] = {w.g.p, w.step.p, w.step.p};
        const double* ys[3] = {w.step.p, w.Hstep.p, w.step.p};
        double d[3];
        w.dots(verbose >= 2 ? 3 : 2, xs, ys, d);
        t_rho = inner.lap();
        const double numer = f - newf;
        const double denom = -d[0] - 0.5 * d[1];
        rho = numer / denom;
        if (verbose >= 2) {
            printf("[update]---------------------------------------------------\n"
                   "║ alpha = %g, ||step|| = %g\n"
                   "║ rho_numer = %g, rho_denom = %g\n",
                   best, std::sqrt(d[2]), numer, denom);
        }
        if (rho > 0.0) vec_axpy(w.gamma, 1.0, w.step, w.N, w.s);
        if (rho < rho_t1)
            mu = mu_s1 * mu;
        else if (rho > rho_t2)
            mu = std::max(mu_s2 * mu, kappa);
        if (rho > 0.0) {
            const double hint = hess_of(p).density();
            if (c_last == best) {
                f = w.obj_grad(w.gamma, w.g, &gnorm);
                t_grad = inner.lap();
                delta = nu0 * gnorm;
                w.threshold(delta, hint);
            } else {
                // w.x still holds gamma_old + c_last d, the last candidate
                const double fx = w.obj_grad(w.x, w.g, &gnorm);
                f = fx - (best - c_last) * w.dot_ab(w.direc);
                t_grad = inner.lap();
                delta = nu0 * gnorm;
                w.threshold(delta, hint, w.x);
            }
            t_hess = inner.lap();
            if (verbose >= 2) {
                printf("║----------------------------------------------------------\n"
                       "[sparse]---------------------------------------------------\n"
                       "║ delta = %g, den = %g, den_new = %g\n"
                       "║----------------------------------------------------------\n"
                       "[timing]---------------------------------------------------\n"
                       "║ lin_solve = %g, grad = %g, hess = %g\n"
                       "║ line_search = %g, rho = %g\n",
                       delta, hint, hess_of(p).density(), t_lin, t_grad, t_hess, t_ls, t_rho);
            }
        }
        if (verbose >= 2) {
            printf("===========================================================\n\n");
            fflush(stdout);
        }
        res->obj_vals.push_back(f);
        res->mar_errs.push_back(gnorm);
        res->run_times.push_back(res->run_times.back() + timer.toc());
        res->densities.push_back(hess_of(p).density());
    }
    res->niter = i;
    w.report(res);
    res->dual.resize(w.N);
    HIP_CHECK(hipMemcpy(res->dual.data(), w.gamma.p, w.N * sizeof(double), hipMemcpyDeviceToHost));
}

// Sparse Newton and SPLR.  Mirror sinkhorn_sparse_newton_internal,
// src/sinkhorn_sparse_newton.cpp:24-170, and sinkhorn_splr_internal,
// src/sinkhorn_splr.cpp:24-199: density starts at 0.1 density_max within
// [0.01, 1] x density_max; shift = min(||g||, shift_max); Wolfe search from
// min(1, 1.5 alpha_prev); sparse Newton's density x2 on a bad move (||g||
// above 1.1 x the previous once below the initial), else x0.9; SPLR's x1.1 /
// x0.99 and its BFGS correction when y's > 1e-6 y'y (not in the first
// iteration).  densities report the target density_T2H(density).
void solve_newton_family(regot_problem* p, bool splr, double tol, int max_iter, int verbose,
                         const regot_opts& o, regot_result* res)
{
    const int n = p->n, m = p->m;
    const double density_max = o.density, density_min = 0.01 * o.density;
    double density = 0.1 * density_max;
    const double shift_max = o.shift;
    const double eps = 1e-6;
    auto density_T2H = [&](double denT) {
        const double nnz = 2.0 * denT * n * (m - 1) + n + m - 1;
        return nnz / static_cast<double>(n + m - 1) / static_cast<double>(n + m - 1);
    };
    PreciseScope ps(p);
    Work w(p);
    w.init(o);
    launch_fill(w.gamma_pre, 0.0, w.N, w.s);
    launch_fill(w.g_pre, 0.0, w.N, w.s);
    Timer timer;
    timer.tic();
    double gnorm;
    double f = w.obj_grad(w.gamma, w.g, &gnorm);
    const double gnorm_init = gnorm;
    w.topk(density);
    res->obj_vals.push_back(f);
    res->mar_errs.push_back(gnorm);
    res->run_times.push_back(timer.toc());
    res->densities.push_back(density_T2H(density));
    double alpha = 1.0;
    int i;
    for (i = 0; i < max_iter; i++) {
        if (verbose >= 1) {
            printf("iter = %d, objval = %g, ||grad|| = %g\n", i, f, gnorm);
            fflush(stdout);
        }
        timer.tic();
        if (gnorm < tol || !std::isfinite(f)) break;
        Timer inner;
        inner.tic();
        const double shift = std::min(gnorm, shift_max);
        vec_axpby(w.rhs, -1.0, w.g, 0.0, nullptr, w.N, w.s);
        bool low_rank = false;
        double ys_ = 0.0, yy_ = 0.0;
        if (splr) {
            vec_axpby(w.y, 1.0, w.g, -1.0, w.g_pre, w.N, w.s);
            vec_axpby(w.sv, 1.0, w.gamma, -1.0, w.gamma_pre, w.N, w.s);
            const double* xs[2] = {w.y.p, w.y.p};
            const double* ys[2] = {w.sv.p, w.y.p};
            double d[2];
            w.dots(2, xs, ys, d);
            ys_ = d[0];
            yy_ = d[1];
            low_rank = (i > 0) && (d[0] > eps * d[1]);
        }
        if (low_rank)
            w.lin_solve_low_rank(o, w.rhs, shift, w.y, w.sv, w.direc);
        else
            w.lin_solve(o, w.rhs, shift, w.direc);
        const double t_lin = inner.lap();
        const double dg0 = w.dot(w.g, w.direc);
        bool reuse = false;
        double f_last = 0.0, gn_last = 0.0, step_last = 0.0;
        alpha = wolfe(w, std::min(1.0, 1.5 * alpha), f, dg0, o.ls_max_iter, &reuse, &f_last,
                      &gn_last, &step_last);
        if (splr) vec_copy(w.gamma_pre, w.gamma, w.N, w.s);
        vec_axpy(w.gamma, alpha, w.direc, w.N, w.s);
        if (splr) vec_copy(w.g_pre, w.g, w.N, w.s);
        const double t_ls = inner.lap();
        const double gnorm_pre = gnorm;
        const double* tpt = nullptr;      // T of the new gamma
        if (reuse) {
            // the last trial's T (src/sinkhorn_sparse_newton.cpp:124-127):
            // gradient and Hessian from it, f from it with the new potentials
            vec_copy(w.g, w.xg, w.N, w.s);
            f = f_last;
            gnorm = gn_last;
            if (step_last != alpha) {
                f = f_last - (alpha - step_last) * w.dot_ab(w.direc);
                tpt = w.x;                 // gamma_old + step_last d
            }
        } else {
            f = w.obj_grad(w.gamma, w.g, &gnorm);
        }
        const double t_grad = inner.lap();
        if (splr) {
            const bool bad = (gnorm_pre < gnorm_init) && (gnorm > gnorm_pre);
            density *= bad ? 1.1 : 0.99;
        } else {
            const bool bad = (gnorm_pre < gnorm_init) && (gnorm > 1.1 * gnorm_pre);
            density *= bad ? 2.0 : 0.9;
        }
        density = std::min(density_max, std::max(density_min, density));
        w.topk(density, tpt);
        const double t_hess = inner.lap();
        if (verbose >= 2) {
            // src/sinkhorn_splr.cpp:167-180, src/sinkhorn_sparse_newton.cpp:142-150
            if (splr)
                printf("[lowrank]---------------------------------------------------\n"
                       "║ ys = %g, yy = %g\n"
                       "║ low_rank = %d\n"
                       "===========================================================\n",
                       ys_, yy_, low_rank ? 1 : 0);
            printf("[timing]---------------------------------------------------\n"
                   "║ lin_solve = %g, line_search = %g\n"
                   "║ grad = %g, hess = %g\n"
                   "===========================================================\n\n",
                   t_lin, t_ls, t_grad, t_hess);
            fflush(stdout);
        }
        res->obj_vals.push_back(f);
        res->mar_errs.push_back(gnorm);
        res->run_times.push_back(res->run_times.back() + timer.toc());
        res->densities.push_back(density_T2H(density));
    }
    res->niter = i;
    w.report(res);
    res->dual.resize(w.N);
    HIP_CHECK(hipMemcpy(res->dual.data(), w.gamma.p, w.N * sizeof(double), hipMemcpyDeviceToHost));
}

// Newton's method with the full dual Hessian.  Mirrors sinkhorn_newton_internal,
// src/sinkhorn_newton.cpp:17-129: direction -H^{-1} g, Armijo backtracking on
// the objective (theta = 0.5, kappa = 0.5, at most 20 trials, the last trial
// taken when none passes), then f, g and H at the new point; stop on
// ||g|| < tol or a non-finite f.  The reference forms H densely
// (src/sinkhorn_problem_obj.cpp:396-418) and factors it with Eigen::LLT; here
// H is the sparsified Hessian with every entry kept (top-k at density 1) and
// the direction comes from the Jacobi-preconditioned device CG on its Schur
// complement, run to a 1e-12 relative residual (cg_tol / cg_max_iter
// override), which reproduces the exact solve to that accuracy.
void solve_newton_dense(regot_problem* p, double tol, int max_iter, int verbose,
                        const regot_opts& o, regot_result* res)
{
    const double theta = 0.5, kappa = 0.5;
    const int nlinesearch = 20;
    const double cg_tol = o.cg_tol > 0 ? o.cg_tol : 1e-12;
    const int cg_max = o.cg_max_iter > 0 ? o.cg_max_iter : 10 * p->m + 1000;
    PreciseScope ps(p);
    Work w(p);
    w.init(o);
    Timer timer;
    timer.tic();
    double gnorm;
    double f = w.obj_grad(w.gamma, w.g, &gnorm);
    w.topk(1.0);
    res->obj_vals.push_back(f);
    res->mar_errs.push_back(gnorm);
    res->run_times.push_back(timer.toc());
    int i;
    for (i = 0; i < max_iter; i++) {
        if (verbose >= 1) {
            printf("iter = %d, objval = %g, ||grad|| = %g\n", i, f, gnorm);
            fflush(stdout);
        }
        timer.tic();
        if (gnorm < tol || !std::isfinite(f)) break;
        vec_axpby(w.rhs, -1.0, w.g, 0.0, nullptr, w.N, w.s);
        {
            const auto t0 = Clock::now();
            w.count_cg(hess_cg(p, w.rhs, 0.0, cg_tol, cg_max, w.direc, nullptr, nullptr, true));
            w.phase_ms[2] += ms_since(t0);
            w.phase_n[2] += 1;
        }
        double step = 1.0;
        const double thresh = theta * w.dot(w.g, w.direc);
        for (int k = 0; k < nlinesearch; k++) {
            w.move(w.x, w.gamma, step, w.direc);
            const double newf = w.obj(w.x);
            if (newf <= f + step * thresh) break;
            step *= kappa;
        }
        vec_copy(w.gamma, w.x, w.N, w.s);
        f = w.obj_grad(w.gamma, w.g, &gnorm);
        w.topk(1.0);
        res->obj_vals.push_back(f);
        res->mar_errs.push_back(gnorm);
        res->run_times.push_back(res->run_times.back() + timer.toc());
    }
    res->niter = i;
    w.report(res);
    res->dual.resize(w.N);
    HIP_CHECK(hipMemcpy(res->dual.data(), w.gamma.p, w.N * sizeof(double), hipMemcpyDeviceToHost));
}

}  // namespace regot
