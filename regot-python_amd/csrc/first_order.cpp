// first_order.cpp -- host drivers of the first-order EROT solvers: L-BFGS on
// the dual and APDAGD.  Control flow and the line-search / step-size logic
// run on the host in FP64 on scalars; every vector lives on the device and
// every O(n m) evaluation (objective, gradient, the APDAGD plan pass) is a
// HIP kernel.  Objectives and gradients are evaluated in FP64 sums (precise
// passes) because both line searches difference objective values.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <utility>
#include <vector>

#include "driver_work.hpp"

namespace regot {
namespace {

// Line-search failures; the texts are LBFGSpp's exception messages
// (src/LBFGSpp/LBFGSpp/LineSearchNocedalWright.h:95-96,216-217,233-234,251-252).
[[noreturn]] void ls_fail(int which)
{
    static const char* msg[] = {
        "the moving direction increases the objective function value",
        "the line search routine failed, possibly due to insufficient numeric precision",
        "the line search routine failed, unable to sufficiently decrease the function value"};
    fail(E_SOLVER, "%s", msg[which]);
}

// Candidate of the zoom phase: minimiser of the quadratic through
// (lo, f_lo), (hi, f_hi) with slope dg_lo at lo; the midpoint when that is not
// finite, not inside (lo, hi) or within 1% of the interval of an end.
// LineSearchNocedalWright::quad_interp,
// src/LBFGSpp/LBFGSpp/LineSearchNocedalWright.h:26-55.
double nw_quad_interp(double lo, double hi, double flo, double fhi, double dglo)
{
    const double fdiff = fhi - flo, sdiff = hi - lo, smid = (hi + lo) / 2.0;
    double c = fdiff * lo - smid * sdiff * dglo;
    c = c / (fdiff - sdiff * dglo);
    const double end_dist = std::min(std::fabs(c - lo), std::fabs(c - hi));
    const bool near_end = end_dist < 0.01 * std::fabs(sdiff);
    const bool bisect = !std::isfinite(c) || c <= std::min(lo, hi) || c >= std::max(lo, hi) ||
                        near_end;
    return bisect ? smid : c;
}

// Device state of the L-BFGS iteration.  x / grad and their "_lo" copies are
// swapped by pointer, as the line search swaps its vectors.
struct LbfgsState {
    double* x;
    double* grad;
    double* xlo;
    double* glo;
    double* xp;
    double* drt;
    double last_f = 0.0;   // the most recent objective evaluation
    double gnorm = 0.0;    // ||grad|| at x (from the evaluation that formed grad)
};

// Strong-Wolfe search (ftol 1e-4, wolfe 0.9, at most 20 evaluations, x2
// expansion): bracketing, then zoom.  LineSearchNocedalWright::LineSearch,
// src/LBFGSpp/LBFGSpp/LineSearchNocedalWright.h:58-283.  On entry fx, grad,
// dg (and st.gnorm) hold the values at xp; on return x, fx, grad, dg, step
// and st.gnorm describe the accepted point.
void nw_line_search(Work& w, LbfgsState& st, double& step, double& fx, double& dg)
{
    const int max_ls = 20;
    const double ftol = 1e-4, wolfe = 0.9, expansion = 2.0;
    const double fx_init = fx, dg_init = dg;
    if (dg_init > 0.0) ls_fail(0);
    const double test_decr = ftol * dg_init, test_curv = -wolfe * dg_init;
    double step_hi = 0.0, fx_hi = 0.0;
    double step_lo = 0.0, fx_lo = fx_init, dg_lo = dg_init;
    vec_copy(st.xlo, st.xp, w.N, w.s);
    vec_copy(st.glo, st.grad, w.N, w.s);
    double gn = st.gnorm, gn_lo = st.gnorm;
    auto eval = [&]() {
        w.move(st.x, st.xp, step, st.drt);
        fx = w.obj_grad(st.x, st.grad, &gn);
        st.last_f = fx;
        dg = w.dot(st.grad, st.drt);
    };
    int iter = 0;
    for (;;) {
        eval();
        if (fx - fx_init > step * test_decr || (0.0 < step_lo && fx >= fx_lo)) {
            step_hi = step;
            fx_hi = fx;
            break;
        }
        if (std::fabs(dg) <= test_curv) {
            st.gnorm = gn;
            return;
        }
        step_hi = step_lo;
        fx_hi = fx_lo;
        step_lo = step;
        fx_lo = fx;
        dg_lo = dg;
        gn_lo = gn;
        std::swap(st.xlo, st.x);
        std::swap(st.glo, st.grad);
        if (dg >= 0.0) break;
        iter++;
        if (iter >= max_ls) {
            // every trial lowered f: the latest one (now in the _lo slots)
            std::swap(st.x, st.xlo);
            std::swap(st.grad, st.glo);
            st.gnorm = gn_lo;
            return;
        }
        step *= expansion;
    }
    for (;;) {
        step = nw_quad_interp(step_lo, step_hi, fx_lo, fx_hi, dg_lo);
        eval();
        if (fx - fx_init > step * test_decr || fx >= fx_lo) {
            if (step == step_hi) ls_fail(1);
            step_hi = step;
            fx_hi = fx;
        } else {
            if (std::fabs(dg) <= test_curv) {
                st.gnorm = gn;
                return;
            }
            if (dg * (step_hi - step_lo) >= 0.0) {
                step_hi = step_lo;
                fx_hi = fx_lo;
            }
            if (step == step_lo) ls_fail(1);
            step_lo = step;
            fx_lo = fx;
            dg_lo = dg;
            gn_lo = gn;
            std::swap(st.xlo, st.x);
            std::swap(st.glo, st.grad);
        }
        iter++;
        if (iter >= max_ls) {
            if (step_lo <= 0.0) ls_fail(2);
            step = step_lo;
            fx = fx_lo;
            dg = dg_lo;
            std::swap(st.x, st.xlo);
            std::swap(st.grad, st.glo);
            st.gnorm = gn_lo;
            return;
        }
    }
}

}  // namespace

// LBFGSSolver::minimize, src/LBFGSpp/LBFGS.h:64-189 (memory 6, parameters
// src/LBFGSpp/LBFGSpp/Param.h:154-171, no past-delta test) on w's objective
// from w.gamma: stops when ||grad|| <= eps_abs or ||grad|| <= eps_rel ||x||,
// or after max_iter iterations (0 = unlimited).  iterate(x, last_f, ||grad||)
// is the iterate() callback (the initial point, then every accepted point);
// last_f is the most recent objective evaluated.  Returns minimize()'s niter
// (1 when the initial point passes); *xout is the device vector holding the
// final point.  The two-loop product runs on scalars: one many-workgroup pass
// forms every dot product it needs (the candidate pair against itself, the
// kept pairs and the gradient; the kept pairs against the gradient), Gram
// matrices of the kept pairs carry the rest, and one pass forms the direction
// as a linear combination of the gradient and the pairs (mathematically the
// recursion's result; the rounding differs).
int lbfgs_minimize(Work& w, double eps_abs, double eps_rel, int max_iter,
                   const std::function<void(const double*, double, double)>& iterate,
                   const double** xout)
{
    const int mm = 6;
    const double eps = std::numeric_limits<double>::epsilon();
    const long N = w.N;
    DBuf<double> S(static_cast<size_t>(mm) * N), Y(static_cast<size_t>(mm) * N);
    DBuf<double> dpart(vec_mdots_scratch(N)), dout(32);
    HBuf<double> hout;
    hout.ensure(32);
    auto slot_s = [&](int i) { return S.p + static_cast<long>(i) * N; };
    auto slot_y = [&](int i) { return Y.p + static_cast<long>(i) * N; };
    // Gram matrices of the correction pairs (by ring slot): SY[i][j] = s_i'y_j,
    // YY[i][j] = y_i'y_j; ys[i] = SY[i][i]
    double SY[6][6] = {}, YY[6][6] = {};
    std::vector<double> ys(mm, 0.0);
    int ncorr = 0, ptr = mm;
    double theta = 1.0;
    LbfgsState st{w.gamma.p, w.g.p, w.lr0.p, w.lr1.p, w.gamma_pre.p, w.direc.p};
    double gnorm;
    double fx = w.obj_grad(st.x, st.grad, &gnorm);
    st.last_f = fx;
    st.gnorm = gnorm;
    auto converged = [&]() {
        if (gnorm <= eps_abs) return true;
        if (eps_rel > 0.0) return gnorm <= eps_rel * std::sqrt(w.dot(st.x, st.x));
        return gnorm <= 0.0;
    };
    int k = 0;
    iterate(st.x, st.last_f, gnorm);
    if (converged()) {
        k = 1;
    } else {
        vec_axpby(st.drt, -1.0, st.grad, 0.0, nullptr, N, w.s);
        double step = 1.0 / gnorm;   // 1 / ||drt||, drt = -grad
        double dg = -gnorm * gnorm;  // grad'drt
        k = 1;
        for (;;) {
            vec_copy(st.xp, st.x, N, w.s);
            vec_copy(w.g_pre, st.grad, N, w.s);
            nw_line_search(w, st, step, fx, dg);
            gnorm = st.gnorm;
            iterate(st.x, st.last_f, gnorm);
            if (converged()) break;
            if (max_iter != 0 && k >= max_iter) break;
            // Candidate pair s = x - xp, y = grad - gradp (temporaries: the
            // ring keeps its slots unless the pair passes s'y > eps y'y), and in
            // the same pass every dot product the two-loop recursion needs:
            // the candidate's against itself, the kept pairs and grad, and
            // the kept pairs' against grad.
            const double* v = st.grad;
            vec_axpby(w.sv, 1.0, st.x, -1.0, st.xp, N, w.s);
            vec_axpby(w.y, 1.0, st.grad, -1.0, w.g_pre, N, w.s);
            const int loc = ptr % mm;
            std::vector<int> act;
            for (int c = 0, j = ptr % mm; c < ncorr; c++) {
                j = (j + mm - 1) % mm;
                act.push_back(j);
            }
            const double* xs[32];
            const double* zs[32];
            int np = 0;
            auto add = [&](const double* x_, const double* z_) {
                xs[np] = x_;
                zs[np] = z_;
                return np++;
            };
            const int i_sy = add(w.sv, w.y), i_yy = add(w.y, w.y), i_sv = add(w.sv, v),
                      i_yv = add(w.y, v), i_vv = add(v, v);
            int i_sYi[6], i_Siy[6], i_yYi[6], i_Sv[6], i_Yv[6];
            for (int i : act) {
                if (i != loc) {
                    i_sYi[i] = add(w.sv, slot_y(i));
                    i_Siy[i] = add(slot_s(i), w.y);
                    i_yYi[i] = add(w.y, slot_y(i));
                }
                i_Sv[i] = add(slot_s(i), v);
                i_Yv[i] = add(slot_y(i), v);
            }
            vec_mdots(xs, zs, np, N, dpart, dout, w.s);
            HIP_CHECK(hipMemcpyAsync(hout.p, dout.p, np * sizeof(double), hipMemcpyDeviceToHost,
                                     w.s));
            HIP_CHECK(hipStreamSynchronize(w.s));
            const double* d = hout.p;
            double Sv[6] = {}, Yv[6] = {};
            for (int i : act) {
                Sv[i] = d[i_Sv[i]];
                Yv[i] = d[i_Yv[i]];
            }
            if (d[i_sy] > eps * d[i_yy]) {
                // BFGSMat::add_correction: the pair takes slot loc
                vec_copy(slot_s(loc), w.sv, N, w.s);
                vec_copy(slot_y(loc), w.y, N, w.s);
                for (int i : act) {
                    if (i == loc) continue;
                    SY[loc][i] = d[i_sYi[i]];
                    SY[i][loc] = d[i_Siy[i]];
                    YY[loc][i] = YY[i][loc] = d[i_yYi[i]];
                }
                SY[loc][loc] = ys[loc] = d[i_sy];
                YY[loc][loc] = d[i_yy];
                Sv[loc] = d[i_sv];
                Yv[loc] = d[i_yv];
                theta = d[i_yy] / d[i_sy];
                if (ncorr < mm) ncorr++;
                ptr = loc + 1;
            }
            // BFGSMat::apply_Hv(grad, -1, drt) on the scalars: loop 1 (newest
            // first) alpha_j = s_j'q / ys_j with q = -grad - sum alpha_i y_i;
            // r = q / theta; loop 2 (oldest first) beta_j = y_j'r / ys_j,
            // r += (alpha_j - beta_j) s_j.  The inner products of q and r
            // expand into the dots above and the Gram matrices.
            const double a = -1.0;
            double alpha[6] = {}, beta[6] = {};
            int order[6], nc = 0;
            for (int c = 0, j = ptr % mm; c < ncorr; c++) {
                j = (j + mm - 1) % mm;
                double sq = a * Sv[j];
                for (int q = 0; q < nc; q++) sq -= alpha[order[q]] * SY[j][order[q]];
                alpha[j] = sq / ys[j];
                order[nc++] = j;
            }
            double done2_coef[6] = {};
            int done2[6], nd = 0;
            for (int c = nc - 1; c >= 0; c--) {
                const int j = order[c];
                double yq = a * Yv[j];
                for (int q = 0; q < nc; q++) yq -= alpha[order[q]] * YY[j][order[q]];
                double yr = yq / theta;
                for (int q = 0; q < nd; q++) yr += done2_coef[q] * SY[done2[q]][j];
                beta[j] = yr / ys[j];
                done2[nd] = j;
                done2_coef[nd++] = alpha[j] - beta[j];
            }
            // drt = (a / theta) grad - sum (alpha_i / theta) y_i + sum (alpha_i - beta_i) s_i
            const double* vs[13];
            double cf[13];
            int nv = 0;
            vs[nv] = v;
            cf[nv++] = a / theta;
            dg = (a / theta) * d[i_vv];
            for (int q = 0; q < nc; q++) {
                const int j = order[q];
                vs[nv] = slot_y(j);
                cf[nv++] = -alpha[j] / theta;
                vs[nv] = slot_s(j);
                cf[nv++] = alpha[j] - beta[j];
                dg += (-alpha[j] / theta) * Yv[j] + (alpha[j] - beta[j]) * Sv[j];
            }
            vec_lincomb(vs, cf, nv, N, st.drt, w.s);
            step = 1.0;
            k++;
        }
    }
    *xout = st.x;
    return k;
}

// L-BFGS on the dual: epsilon = tol, epsilon_rel = 0, max_iterations =
// max_iter.  Mirrors sinkhorn_lbfgs_dual_internal,
// src/sinkhorn_lbfgs_dual.cpp:84-125, with the iterate() callback
// (src/sinkhorn_lbfgs_dual.cpp:56-76): obj_vals records the most recent
// objective evaluation, mar_errs the accepted point's ||grad||.
void solve_lbfgs_dual(regot_problem* p, double tol, int max_iter, int verbose,
                      const regot_opts& o, regot_result* res)
{
    PreciseScope ps(p);
    Work w(p);
    w.init(o);
    Timer timer;
    timer.tic();
    auto iterate = [&](const double*, double last_f, double gnorm) {
        const double dur = timer.toc();
        res->obj_vals.push_back(last_f);
        res->mar_errs.push_back(gnorm);
        res->run_times.push_back((res->run_times.empty() ? 0.0 : res->run_times.back()) + dur);
        if (verbose >= 1) {
            printf("iter = %d, objval = %g, ||grad|| = %g\n",
                   static_cast<int>(res->obj_vals.size()) - 1, last_f, gnorm);
            fflush(stdout);
        }
        timer.tic();
    };
    const double* x = nullptr;
    res->niter = lbfgs_minimize(w, tol, 0.0, max_iter, iterate, &x);
    w.report(res);
    res->dual.resize(w.N);
    HIP_CHECK(hipMemcpy(res->dual.data(), x, w.N * sizeof(double), hipMemcpyDeviceToHost));
}

// Accelerated primal-dual adaptive gradient descent.  Mirrors
// sinkhorn_apdagd_internal, src/sinkhorn_apdagd.cpp:14-183: dual variables
// lambda = -gamma, L halved per iteration and doubled per failed test (at most
// 20 inner trials, the last one accepted), the primal plan the running
// combination (1 - tau) P + tau T(-lambda~) kept on the device, convergence
// when the primal-dual gap and the full marginal error (all m columns) are
// both <= tol, tested after the history entry.  The final dual is
// -lambda~; the plan stays in the result (device) and is copied out on request.
void solve_apdagd(regot_problem* p, double tol, int max_iter, int verbose, const regot_opts& o,
                  regot_result* res)
{
    const double L0 = 1.0;
    const int max_inner = 20;
    const int n = p->n, m = p->m;
    PreciseScope ps(p);
    Work w(p);
    w.init(o);
    // lambda (du), lambda' (dp), lambda~ (dt) and their new values
    double* du = w.direc.p;
    double* du_new = w.step.p;
    double* dp = w.Hstep.p;
    double* dp_new = w.rhs.p;
    double* dt = w.y.p;
    double* diff = w.sv.p;
    double* x = w.gamma.p;
    double* grad = w.g.p;
    vec_axpby(du, -1.0, x, 0.0, nullptr, w.N, w.s);
    vec_copy(dp, du, w.N, w.s);
    vec_copy(dt, du, w.N, w.s);
    // plan pass buffers
    const long ldp = (m + 31) / 32 * 32;
    res->plan.alloc(static_cast<size_t>(n) * ldp);
    res->plan_ld = ldp;
    res->plan_device = p->device;
    long nrp, ncp, nsc, npt;
    apd_scratch_sizes(n, m, &nrp, &ncp, &nsc, &npt);
    DBuf<double> rowpart(nrp), colpart(ncp), scal(nsc), part(npt), out(3);
    HBuf<double> hout;
    hout.ensure(3);
    DBuf<double> alpha(n), beta(m);
    auto plan_pass = [&](double tau, bool init) {
        launch_split_gamma(x, n, m, alpha, beta, w.s);
        launch_apd_plan(p->bulk, p->dM, p->ldm, n, m, alpha, beta, p->a, p->b, p->reg, tau, init,
                        res->plan, ldp, rowpart, colpart, scal, part, out, w.s);
        // read M (and P unless init), write P
        w.alg_bytes += static_cast<double>(n) * m * ((p->bulk == 0 ? 4.0 : 8.0) + (init ? 8 : 16));
        HIP_CHECK(hipMemcpyAsync(hout.p, out.p, 3 * sizeof(double), hipMemcpyDeviceToHost, w.s));
        HIP_CHECK(hipStreamSynchronize(w.s));
    };
    plan_pass(1.0, true);   // result.get_plan(gamma) at the initial point
    Timer timer;
    timer.tic();
    double gnorm;
    double f = w.obj_grad(x, grad, &gnorm);
    res->obj_vals.push_back(f);
    res->mar_errs.push_back(gnorm);
    res->run_times.push_back(timer.toc());
    int i;
    double L = L0, pbeta = 0.0;
    for (i = 0; i < max_iter; i++) {
        timer.tic();
        if (verbose >= 2)
            printf("\n[line search]=================================================\n");
        L *= 0.5;
        double ptau = 0.0, phi_new = 0.0;
        for (int j = 0; j < max_inner; j++) {
            const double palpha = (1.0 + std::sqrt(4.0 * L * pbeta + 1.0)) / (2.0 * L);
            const double pbeta_new = pbeta + palpha;
            ptau = palpha / pbeta_new;
            vec_axpby(dt, 1.0 - ptau, du, ptau, dp, w.N, w.s);
            vec_axpby(x, -1.0, dt, 0.0, nullptr, w.N, w.s);
            f = w.obj_grad(x, grad, nullptr);
            const double phit = -f;
            vec_axpby(dp_new, 1.0, dp, palpha, grad, w.N, w.s);
            vec_axpby(du_new, 1.0 - ptau, du, ptau, dp_new, w.N, w.s);
            vec_axpby(x, -1.0, du_new, 0.0, nullptr, w.N, w.s);
            f = w.obj(x);
            phi_new = -f;
            vec_axpby(diff, 1.0, du_new, -1.0, dt, w.N, w.s);
            const double* xs[2] = {grad, diff};
            const double* zs[2] = {diff, diff};
            double d[2];
            w.dots(2, xs, zs, d);
            const double rhs = phit + d[0] - 0.5 * L * d[1];
            if (verbose >= 2) {
                printf("║ j = %d, palpha = %g, pbeta = %g\n", j, palpha, pbeta_new);
                printf("║ phit = %g, phi_new = %g, rhs = %g\n", phit, phi_new, rhs);
                printf("║-------------------------------------------------------------\n");
            }
            if (phi_new >= rhs || j >= max_inner - 1) {
                pbeta = pbeta_new;
                std::swap(dp, dp_new);
                std::swap(du, du_new);
                break;
            }
            L *= 2.0;
        }
        vec_axpby(x, -1.0, dt, 0.0, nullptr, w.N, w.s);
        plan_pass(ptau, false);
        const double primal_obj = hout.p[1] - p->reg * hout.p[2];
        const double mar_err = std::sqrt(hout.p[0]);
        res->obj_vals.push_back(f);
        res->mar_errs.push_back(mar_err);
        res->run_times.push_back(res->run_times.back() + timer.toc());
        if (verbose >= 1) {
            printf("iter = %d, primal_obj = %g, phi_new = %g, mar_err = %g\n", i, primal_obj,
                   phi_new, mar_err);
            fflush(stdout);
        }
        if (primal_obj - phi_new <= tol && mar_err <= tol) break;
    }
    res->niter = i;
    w.report(res);
    res->dual.resize(w.N);
    HIP_CHECK(hipMemcpy(res->dual.data(), x, w.N * sizeof(double), hipMemcpyDeviceToHost));
}

}  // namespace regot
