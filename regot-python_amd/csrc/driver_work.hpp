// driver_work.hpp -- device workspace and helpers shared by the host drivers
// of the second-order and first-order EROT solvers (newton.cpp,
// first_order.cpp): every vector lives on the device, every O(n m) evaluation
// is a kernel sequence, and only scalars come back to the host.
#pragma once

#include <chrono>
#include <cmath>
#include <functional>
#include <limits>

#include "drivers.hpp"
#include "hessian.hpp"

namespace regot {

using Clock = std::chrono::steady_clock;

inline double ms_since(Clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// Sets the handle's precise flag for the lifetime of a solve: objectives and
// gradients the line searches compare are then FP64 sums even in FP32 bulk
// (the sparsified Hessian keeps the bulk precision).
struct PreciseScope {
    regot_problem* p;
    bool was;
    explicit PreciseScope(regot_problem* pp) : p(pp), was(pp->precise) { p->precise = true; }
    ~PreciseScope() { p->precise = was; }
};

// Device workspace of one Newton-type solve.
struct Work {
    regot_problem* p;
    int n, m;
    long N;
    hipStream_t s;
    DBuf<double> gamma, g, direc, step, Hstep, x, xg, rhs, y, sv, g_pre, gamma_pre, lr0, lr1;
    DBuf<double> dscal;
    HBuf<double> hscal;
    long long cg_iters = 0;
    // algorithmic HBM bytes of the work done: n m s per objective or gradient
    // evaluation (one logical read of M), n m s + 2 nnz 12 per Hessian build
    // (CSR + CSC written), 2 nnz 12 per Hessian product (s = 4 or 8 bytes)
    double alg_bytes = 0.0;
    // per-phase wall time (each phase ends in a host synchronisation)
    double phase_ms[3] = {0.0, 0.0, 0.0};
    long long phase_n[3] = {0, 0, 0};
    // Objective other than the EROT dual (e.g. QROT): og(x, grad, &||grad||^2)
    // returns f; grad and the norm pointer may be null.  Vectors then have
    // N = nvar entries.
    std::function<double(const double*, double*, double*)> og;

    explicit Work(regot_problem* pp, long nvar = -1)
        : p(pp), n(pp->n), m(pp->m), N(nvar > 0 ? nvar : pp->n + pp->m - 1), s(pp->stream)
    {
        for (DBuf<double>* b : {&gamma, &g, &direc, &step, &Hstep, &x, &xg, &rhs, &y, &sv, &g_pre,
                                &gamma_pre, &lr0, &lr1})
            b->alloc(N);
        dscal.alloc(8);
        hscal.ensure(8);
    }

    // Reference init (e.g. src/sinkhorn_ssns.cpp:53-62): alpha = 0,
    // beta = optimal_beta(alpha), gamma normalised; or the caller's x0.
    void init(const regot_opts& o)
    {
        if (o.x0) {
            HIP_CHECK(hipMemcpyAsync(gamma, o.x0, N * sizeof(double), hipMemcpyHostToDevice, s));
            return;
        }
        launch_fill(p->alpha, 0.0, n, s);
        ColEpi c;
        c.beta_out = p->beta;
        p->col_pass(p->alpha, nullptr, c);
        vec_normalise(p->alpha, p->beta, n, m, gamma, s);
    }

    double dot(const double* a, const double* b)
    {
        const double* xs[1] = {a};
        const double* ys[1] = {b};
        vec_dots(xs, ys, N, 1, dscal, s);
        HIP_CHECK(hipMemcpyAsync(hscal.p, dscal.p, sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        return hscal.p[0];
    }

    // Several dots in one launch.
    void dots(int k, const double* const* a, const double* const* b, double* out)
    {
        vec_dots(a, b, N, k, dscal, s);
        HIP_CHECK(hipMemcpyAsync(hscal.p, dscal.p, k * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        for (int q = 0; q < k; q++) out[q] = hscal.p[q];
    }

    // [a; b_0..m-2] . v, for objectives formed from one point's T with
    // another point's potentials (f = reg sum T - gamma'[a; b~])
    double dot_ab(const double* v)
    {
        const double* xs[2] = {v, v + n};
        const double* ys[2] = {p->a.p, p->b.p};
        vec_dots(xs, ys, n, 1, dscal, s);
        vec_dots(xs + 1, ys + 1, m - 1, 1, dscal.p + 1, s);
        HIP_CHECK(hipMemcpyAsync(hscal.p, dscal.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        return hscal.p[0] + hscal.p[1];
    }

    // out = x + t d
    void move(double* out, const double* x0, double t, const double* d)
    {
        vec_axpby(out, 1.0, x0, t, d, N, s);
    }

    double obj(const double* at)
    {
        const auto t0 = Clock::now();
        const double f = og ? og(at, nullptr, nullptr) : p->dev_obj(at);
        alg_bytes += pass_bytes();
        phase_ms[0] += ms_since(t0);
        phase_n[0] += 1;
        return f;
    }

    double obj_grad(const double* at, double* grad, double* gnorm)
    {
        const auto t0 = Clock::now();
        double gn2 = 0.0;
        const double f = og ? og(at, grad, &gn2) : p->dev_obj_grad(at, grad, &gn2);
        alg_bytes += pass_bytes();
        if (gnorm) *gnorm = std::sqrt(gn2);
        phase_ms[0] += ms_since(t0);
        phase_n[0] += 1;
        return f;
    }

    // The sparsified Hessian of the T formed at `tpt` (gamma unless given),
    // with the diagonals from the gradient g, which is that T's.
    void threshold(double delta, double hint, const double* tpt = nullptr)
    {
        const auto t0 = Clock::now();
        hess_threshold(p, tpt ? tpt : gamma.p, g, delta, hint, true);
        alg_bytes += pass_bytes() + build_bytes();
        phase_ms[1] += ms_since(t0);
        phase_n[1] += 1;
    }

    void topk(double density, const double* tpt = nullptr)
    {
        const auto t0 = Clock::now();
        hess_topk(p, tpt ? tpt : gamma.p, g, density, true);
        alg_bytes += pass_bytes() + build_bytes();
        phase_ms[1] += ms_since(t0);
        phase_n[1] += 1;
    }

    double pass_bytes() const
    {
        return static_cast<double>(n) * m * (p->bulk == 0 ? 4.0 : 8.0);
    }

    // what a Hessian build writes: CSR + CSC (12 bytes per entry each), or
    // the kept matrix of a matrix-free build (4 bytes per entry)
    double build_bytes()
    {
        const SparseHessian& H = hess_of(p);
        if (H.mf) return static_cast<double>(n) * (m - 1) * 4.0;
        return 2.0 * H.nnz * 12.0;
    }

    // one product with sigma and sigma': the CSR + CSC reads, or a
    // matrix-free pass (the FP32 kept matrix)
    double product_bytes()
    {
        const SparseHessian& H = hess_of(p);
        if (H.mf) return static_cast<double>(n) * (m - 1) * 4.0;
        return 2.0 * H.nnz * 12.0;
    }

    // CG iterations of one solve: one product with Delta each, plus the two
    // sigma products of the block elimination
    void count_cg(int iters)
    {
        cg_iters += iters;
        alg_bytes += (iters + 1) * product_bytes();
    }

    void report(regot_result* res) const
    {
        res->cg_iters = cg_iters;
        res->alg_bytes = alg_bytes;
        for (int q = 0; q < 3; q++) {
            res->phase_ms[q] = phase_ms[q];
            res->phase_n[q] = phase_n[q];
        }
    }

    // (Hs + shift I)^{-1} rhs.  method 0 is the reference's CG (relative
    // residual 1e-8, Eigen's default cap of 2 * HessianCG::cols() = 2m, no
    // preconditioner).  The reference's default method (SimplicialLDLT) is an
    // exact solve; a Jacobi-preconditioned CG stands in for it on the GPU with a
    // 1e-10 residual and a cap high enough that the tolerance, not the cap,
    // ends it on ill-conditioned Schur complements.
    void lin_solve(const regot_opts& o, const double* r, double shift, double* out)
    {
        const double tol = o.cg_tol > 0 ? o.cg_tol : (o.method == 0 ? 1e-8 : 1e-10);
        const int maxit = o.cg_max_iter > 0 ? o.cg_max_iter
                                            : (o.method == 0 ? 2 * m : 10 * m + 1000);
        const auto t0 = Clock::now();
        count_cg(hess_cg(p, r, shift, tol, maxit, out, nullptr, nullptr, o.method != 0));
        phase_ms[2] += ms_since(t0);
        phase_n[2] += 1;
    }

    // Inverse-BFGS correction of (Hs + shift I)^{-1} applied to r, with the
    // curvature pair (s, y).  src/sinkhorn_linsolve.cpp:160-225 (two solves, then
    // res = A^-1 r + ((1 + y'A^-1y / y's) s'r/y's - y'A^-1r / y's) s - (s'r/y's) A^-1 y).
    void lin_solve_low_rank(const regot_opts& o, const double* r, double shift, const double* yv,
                            const double* sv_, double* out)
    {
        // the two solves share the system: their CGs run in the same launches
        // (each with the iterates of its own lin_solve)
        const double tol = o.cg_tol > 0 ? o.cg_tol : (o.method == 0 ? 1e-8 : 1e-10);
        const int maxit = o.cg_max_iter > 0 ? o.cg_max_iter
                                            : (o.method == 0 ? 2 * m : 10 * m + 1000);
        const auto t0 = Clock::now();
        int it2[2] = {0, 0};
        hess_cg2(p, r, yv, shift, tol, maxit, lr0, lr1, it2, nullptr, o.method != 0);
        count_cg(it2[0]);
        count_cg(it2[1]);
        phase_ms[2] += ms_since(t0);
        phase_n[2] += 2;
        const double* xs[4] = {sv_, yv, yv, yv};
        const double* ys[4] = {r, sv_, lr1.p, lr0.p};
        double d[4];
        dots(4, xs, ys, d);
        const double rs = d[0], ys_ = d[1], yiAy = d[2], yiAr = d[3];
        const double rs_ys = rs / ys_;
        const double cs = (1.0 + yiAy / ys_) * rs_ys - yiAr / ys_;
        // out = lr0 + cs s - rs_ys lr1
        vec_axpby(out, 1.0, lr0, cs, sv_, N, s);
        vec_axpy(out, -rs_ys, lr1, N, s);
    }
};

struct Timer {
    Clock::time_point t0;
    void tic() { t0 = Clock::now(); }
    double toc() { return ms_since(t0); }
    // ms since the last tic or lap, then restart (the reference's
    // Timer::toc(name), src/timer.h:38-45)
    double lap()
    {
        const auto t = Clock::now();
        const double d = std::chrono::duration<double, std::milli>(t - t0).count();
        t0 = t;
        return d;
    }
};

// L-BFGS (first_order.cpp) on w's objective from w.gamma; see there.
int lbfgs_minimize(Work& w, double eps_abs, double eps_rel, int max_iter,
                   const std::function<void(const double*, double, double)>& iterate,
                   const double** xout);

}  // namespace regot
