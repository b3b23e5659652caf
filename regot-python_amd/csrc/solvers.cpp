// solvers.cpp -- host C++ drivers of the EROT solvers and the solver-level
// C-ABI.  The drivers own the control flow (tests, schedules, histories) in
// FP64; every O(n m) step runs in the HIP kernels.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <deque>
#include <vector>

#include "drivers.hpp"
#include "problem.hpp"
#include "qrot.hpp"
#include "regot_hip.h"
#include "sparse.hpp"

namespace regot {

int set_error(int code, const char* what);

namespace {

// gamma = [alpha + beta_{m-1}; beta_{0..m-2} - beta_{m-1}]; src/sinkhorn_bcd.cpp:50-51
void normalise_host(int n, int m, const double* alpha, const double* beta, double* gamma)
{
    for (int i = 0; i < n; i++) gamma[i] = alpha[i] + beta[m - 1];
    for (int j = 0; j < m - 1; j++) gamma[n + j] = beta[j] - beta[m - 1];
}

// Plain Sinkhorn / BCD with the whole loop on the GPU.
//
// Semantics follow sinkhorn_bcd_internal, src/sinkhorn_bcd.cpp:14-124: init
// alpha = 0, beta = optimal_beta(alpha); at iterate k test ||g_k|| < tol first,
// then alpha <- optimal_alpha(beta), beta <- optimal_beta(alpha).
//
// Each iteration is three launches on one stream:
//   row pass        beta_k -> alpha_{k+1} (double buffer) and the row sums at
//                   iterate k; each workgroup writes its partials of
//                   ||r - a||^2, sum r and alpha'a;
//   column partial  per-tile log2 sums of column j at alpha_{k+1};
//   column combine  every workgroup reduces the row partials and takes the
//                   stop decision for iterate k itself (same data, same
//                   order); unless stopping it writes beta_{k+1} and the
//                   tile sums of the new column residual (g_{k+1}'s column
//                   part, zero up to FP64 rounding).  Workgroup 0 records
//                   f_k, ||g_k|| and a timestamp and stops or advances k.
// Once the stop flag is set every later kernel returns at once, so the host
// keeps two batches of iterations in flight and only polls the flag.
//
// Single-read form (bcd_fused.hip, FP32 bulk): the row pass and the column
// partial become one pass whose workgroups hold whole rows; the combine then
// sums plain column sums.  It needs beta_k to come from an exact beta update,
// so with a warm start iteration 0 keeps the two-pass form.
//
// Precise finish (FP32 bulk).  The FP32 passes resolve each row or column
// sum to about 2^-23 of its size in the worst case (a row dominated by one
// term carries that term's exp2 rounding), so ||g|| can have a floor near
// 2^-23 ||[a; b]||, below which a tolerance cannot be met or trusted.  When
// tol lies below twice that bound, the FP32 loop stops at the bound instead;
// it also stops when ||g|| has made no new minimum for kStallWindow
// iterations (ctl[0] = 2: an unexpected floor).  Either way the driver then
// re-projects beta_k = optimal_beta(alpha_k) with FP64 exp and sums and
// continues from iterate k (its history entry re-evaluated) with the
// two-pass iteration in FP64 arithmetic over the FP32 M, to tol or max_iter,
// as the reference runs FP64 throughout after its upcast
// (src/solvers.cpp:359-373).
constexpr int kStallWindow = 24;

void solve_bcd(regot_problem* p, double tol, int max_iter, int verbose, const regot_opts& o,
               regot_result* res)
{
    const int n = p->n, m = p->m;
    hipStream_t s = p->stream;
    // With kernel profiling on, the iterations are enqueued exactly as
    // without it; each pass's events carry its iteration, and the launches
    // past the stop decision (which return at once) are dropped at the end.
    const int K = o.bcd_batch > 0 ? o.bcd_batch : 8;
    const int cap = std::max(4096, 4 * K + 8);

    if (!p->bcd) p->bcd = std::make_unique<BcdWork>();
    BcdWork& w = *p->bcd;
    const int nrb = row_pass_blocks(p->bulk, n);
    // measured faster than the two-pass form at every eligible shape tried
    // (2048^2: 1.3x, 8192^2 and 32768 x 8192: 1.8x per solve)
    const bool fused_ok = o.bcd_fused != 0 && p->bcd_fused_ok();
    const int nfb = fused_ok ? bcd_fused_blocks(n, p->ncu) : 0;
    if (fused_ok) w.fpart.ensure(static_cast<size_t>(nfb) * m);
    w.ap0.ensure(n);
    w.ap1.ensure(n);
    w.beta.ensure(m);
    w.gcol.ensure(m);
    w.rpart.ensure(3L * std::max(nrb, nfb));
    w.cpart.ensure(4L * p->ntiles);
    w.ctl.ensure(4);
    w.bm.ensure(1);
    w.stall.ensure(4);
    w.hf.ensure(cap);
    w.hg.ensure(cap);
    w.hts.ensure(cap + 1);
    w.hc.ensure(8);
    for (auto& e : w.ev)
        if (!e) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));

    const bool fp32 = p->bulk == REGOT_F32;
    bool precise = false;
    const double tol_fp32 =
        fp32 ? std::max(tol, 0x1p-22 * std::sqrt(p->sumsq_a + p->sumsq_b)) : tol;
    struct FlagScope {
        bool& f;
        bool was;
        FlagScope(bool& ff, bool v) : f(ff), was(ff) { f = v; }
        ~FlagScope() { f = was; }
    };

    std::vector<double> ah(n), bh(m);

    ColEpi c0;
    c0.resid = w.gcol;
    c0.tilepart = w.cpart;
    {
        FlagScope fs(p->precise, precise);
        if (o.x0) {
            HIP_CHECK(hipMemcpyAsync(p->gamma, o.x0, (n + m - 1) * sizeof(double),
                                     hipMemcpyHostToDevice, s));
            launch_split_gamma(p->gamma, n, m, w.ap0, w.beta, s);
            c0.colpot = w.beta;
            p->col_pass(w.ap0, w.beta, c0);
        } else {
            launch_fill(w.ap0, 0.0, n, s);
            c0.beta_out = w.beta;
            p->col_pass(w.ap0, nullptr, c0);
        }
    }
    launch_stamp(w.hts.d + cap, s);

    BcdState st{};
    st.a = p->a;
    st.beta = w.beta;
    st.ctl = w.ctl;
    st.hist_f = w.hf.d;
    st.hist_g = w.hg.d;
    st.hist_ts = w.hts.d;
    st.hist_cap = cap;
    st.rpart = w.rpart;
    st.bm = w.bm;
    st.cpart = w.cpart;
    st.nrb = nrb;
    st.ntiles = p->ntiles;
    st.sum_a = p->sum_a;
    st.sum_b1 = p->sum_b1;
    st.tol = tol;
    st.reg = p->reg;
    st.max_iter = max_iter;
    st.n = n;
    st.m = m;
    st.stall = w.stall;

    const double tick_ms = p->wall_khz > 0 ? 1.0 / p->wall_khz : 0.0;
    int h_read = 0, niter = 0;
    long long ts_init = 0;

    // verbose >= 2: the reference's per-iteration [timing] block
    // (src/sinkhorn_bcd.cpp:99-106) from HIP events around each pass: alpha =
    // the row pass (or the single-read pass), beta = the column pass (or the
    // single-read pass's combine), grad = 0 (both passes produce the gradient
    // terms).  The events use the profiling slots; without profiling on, they
    // are dropped after the solve.
    const bool user_prof = p->prof;
    size_t cur[4] = {0, 0, 0, 0};
    if (verbose >= 2) {
        p->prof = true;
        for (int q = 0; q < 4; q++) cur[q] = p->prof_slot[q].tagged.size();
    }
    auto pass_ms = [&](int kind, int it) {
        const auto& tg = p->prof_slot[kind].tagged;
        while (cur[kind] < tg.size() && tg[cur[kind]].first < it) cur[kind]++;
        double t = 0.0;
        for (size_t q = cur[kind]; q < tg.size() && tg[q].first == it; q++) t += tg[q].second;
        return t;
    };
    auto print_timing = [&](int it) {
        p->prof_collect();
        printf("[timing]=================================================\n"
               "║ alpha = %g, beta = %g, grad = %g\n"
               "=========================================================\n\n",
               pass_ms(0, it) + pass_ms(2, it), pass_ms(1, it) + pass_ms(3, it), 0.0);
    };

    // Whether an FP32 loop that stopped at iterate k with status ctl[0] hands
    // over to the FP64 finish (iterate k is then evaluated again)
    auto handover = [&](int status, int k) {
        return fp32 && !precise && k < max_iter && (status == 2 || tol_fp32 > tol);
    };

    // One run of the on-device loop from iterate k0 until its stop decision;
    // returns ctl[0] (1: converged or max_iter, 2: stalled).
    auto run_loop = [&](int k0) -> int {
        const bool fused = fused_ok && !precise;
        FlagScope fs(p->bcd_precise, precise);
        st.stall_window = (fp32 && !precise) ? kStallWindow : 0;
        st.tol = precise ? tol : tol_fp32;
        // ctl = {0, k0, 0, 0}; both stall slots = {+inf, k0} (fill kernels,
        // no memset nodes and no host synchronisation)
        launch_fill_int(w.ctl, 0, 4, s);
        launch_fill_int(w.ctl.p + 1, k0, 1, s);
        launch_fill(w.stall, static_cast<double>(k0), 4, s);
        launch_fill(w.stall.p, HUGE_VAL, 1, s);
        launch_fill(w.stall.p + 2, HUGE_VAL, 1, s);
        // alpha_k lives in ap[k & 1]; iteration k writes alpha_{k+1} to the other
        int k_enq = k0;
        auto enqueue_iteration = [&]() {
            const int kk = k_enq++;
            p->prof_tag = kk;
            double* ak = (kk & 1) ? w.ap1.p : w.ap0.p;
            double* ak1 = (kk & 1) ? w.ap0.p : w.ap1.p;
            st.k = kk;
            ColEpi c;
            c.beta_out = w.beta;
            c.resid = w.gcol;
            c.tilepart = w.cpart;
            if (fused && (kk > 0 || !o.x0)) {
                st.nrb = nfb;
                p->bcd_fused_iteration(ak, ak1, w.beta, c, st, nfb, w.fpart);
                return;
            }
            st.nrb = nrb;
            RowEpi e;
            e.loga = p->loga;
            e.rowpot = ak;
            e.alpha_out = ak1;
            p->row_pass_bcd(w.beta, e, st);
            p->col_pass(ak1, w.beta, c, w.ctl, &st);
        };
        auto launch_batch = [&](int slot) {
            for (int it = 0; it < K && k_enq <= max_iter; it++) enqueue_iteration();
            HIP_CHECK(hipMemcpyAsync(w.hc.p + 4 * slot, w.ctl.p, 4 * sizeof(int),
                                     hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipEventRecord(w.ev[slot], s));
        };
        std::deque<int> q;
        int status = 0;
        launch_batch(0);
        q.push_back(0);
        launch_batch(1);
        q.push_back(1);
        for (;;) {
            const int slot = q.front();
            q.pop_front();
            HIP_CHECK(hipEventSynchronize(w.ev[slot]));
            const int* c = w.hc.p + 4 * slot;
            const bool done = c[0] != 0;
            const int avail = done ? c[2] + 1 : c[1];
            if (h_read == 0 && avail > 0) ts_init = w.hts.h[cap];
            for (; h_read < avail; h_read++) {
                const int h = h_read % cap;
                res->obj_vals.push_back(w.hf.h[h]);
                res->mar_errs.push_back(w.hg.h[h]);
                res->run_times.push_back(static_cast<double>(w.hts.h[h] - ts_init) * tick_ms);
                if (verbose >= 1 && h_read < max_iter && !(done && h_read == c[2] &&
                                                            handover(c[0], c[2]))) {
                    printf("iter = %d, objval = %g, ||grad|| = %g\n", h_read, w.hf.h[h],
                           w.hg.h[h]);
                    if (verbose >= 2 && !(done && h_read == c[2])) print_timing(h_read);
                }
            }
            if (verbose >= 1) fflush(stdout);
            if (done) {
                niter = c[2];
                status = c[0];
                break;
            }
            launch_batch(slot);
            q.push_back(slot);
        }
        HIP_CHECK(hipStreamSynchronize(s));
        return status;
    };

    try {
        int k0 = 0;
        for (;;) {
            const int status = run_loop(k0);
            if (!handover(status, niter)) break;
            // FP32 stopped at iterate niter short of tol: drop its entry,
            // re-project beta in FP64 and go on from there in FP64 arithmetic
            res->obj_vals.pop_back();
            res->mar_errs.pop_back();
            res->run_times.pop_back();
            h_read = niter;
            k0 = niter;
            precise = true;
            res->precise_from = k0;
            if (p->prof) {
                // the FP32 launches past iterate k0 returned at once
                p->prof_drop_after(k0);
                for (int q = 0; q < 4; q++)
                    cur[q] = std::min(cur[q], p->prof_slot[q].tagged.size());
            }
            FlagScope fs(p->precise, true);
            ColEpi c;
            c.beta_out = w.beta;
            c.resid = w.gcol;
            c.tilepart = w.cpart + 2L * p->ntiles * (k0 & 1);
            p->col_pass((k0 & 1) ? w.ap1.p : w.ap0.p, w.beta, c);
        }
    } catch (...) {
        p->prof_tag = -1;
        p->prof = user_prof;
        (void)hipStreamSynchronize(s);
        throw;
    }
    p->prof_tag = -1;
    if (verbose >= 2 && !user_prof) {
        p->prof_collect();
        for (auto& slot : p->prof_slot) slot.tagged.clear();
        p->prof = false;
    }
    if (p->prof) p->prof_commit(niter);

    HIP_CHECK(hipMemcpyAsync(ah.data(), (niter & 1) ? w.ap1.p : w.ap0.p, n * sizeof(double),
                             hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(bh.data(), w.beta.p, m * sizeof(double), hipMemcpyDeviceToHost, s));
    p->sync();
    res->niter = niter;
    res->dual.resize(n + m - 1);
    normalise_host(n, m, ah.data(), bh.data(), res->dual.data());
}

}  // namespace
}  // namespace regot

using namespace regot;

#define CAPI_BEGIN try {
#define CAPI_END                                                                          \
    }                                                                                     \
    catch (const regot::Error& e) { return set_error(e.code, e.what()); }                 \
    catch (const std::bad_alloc&) { return set_error(E_NOMEM, "host allocation failed"); } \
    catch (const std::exception& e) { return set_error(E_INTERNAL, e.what()); }            \
    return OK;

extern "C" {

void regot_opts_default(regot_opts* o, int solver)
{
    if (!o) return;
    // SinkhornSolverOpts defaults, src/sinkhorn_solvers.h:35-37, with the
    // per-solver method override of src/solvers.cpp:415,433,451.
    (void)solver;
    o->method = 1;
    o->mu0 = 1.0;
    o->shift = 1e-3;
    o->density = 0.01;
    o->cg_tol = 0.0;
    o->cg_max_iter = 0;
    o->x0 = nullptr;
    o->bcd_batch = 0;
    o->bcd_fused = 1;
    o->tau = 0.0;
    o->ls_max_iter = 0;
}

int regot_solve(regot_problem* p, int solver, double tol, int max_iter, int verbose,
                const regot_opts* opts, regot_result** out)
{
    CAPI_BEGIN
    if (!p) fail(E_INVALID, "null problem handle");
    if (!out) fail(E_INVALID, "out is null");
    *out = nullptr;
    if (max_iter < 0) fail(E_INVALID, "max_iter must be >= 0");
    HIP_CHECK(hipSetDevice(p->device));
    regot_opts o;
    if (opts)
        o = *opts;
    else
        regot_opts_default(&o, solver);
    auto res = std::make_unique<regot_result>();
    res->solver = solver;
    res->n = p->n;
    res->m = p->m;
    switch (solver) {
        case REGOT_SOLVER_BCD:
            solve_bcd(p, tol, max_iter, verbose, o, res.get());
            break;
        case REGOT_SOLVER_SSNS:
            solve_ssns(p, tol, max_iter, verbose, o, res.get());
            break;
        case REGOT_SOLVER_SPARSE_NEWTON:
            solve_newton_family(p, false, tol, max_iter, verbose, o, res.get());
            break;
        case REGOT_SOLVER_SPLR:
            solve_newton_family(p, true, tol, max_iter, verbose, o, res.get());
            break;
        case REGOT_SOLVER_NEWTON:
            solve_newton_dense(p, tol, max_iter, verbose, o, res.get());
            break;
        case REGOT_SOLVER_LBFGS_DUAL:
            solve_lbfgs_dual(p, tol, max_iter, verbose, o, res.get());
            break;
        case REGOT_SOLVER_APDAGD:
            solve_apdagd(p, tol, max_iter, verbose, o, res.get());
            break;
        default:
            fail(E_INVALID, "unknown solver id %d", solver);
    }
    *out = res.release();
    CAPI_END
}

int regot_result_niter(const regot_result* r) { return r ? r->niter : -1; }

int regot_result_hist_len(const regot_result* r)
{
    return r ? static_cast<int>(r->obj_vals.size()) : -1;
}

int regot_result_dens_len(const regot_result* r)
{
    return r ? static_cast<int>(r->densities.size()) : -1;
}

int regot_result_get(const regot_result* r, double* dual, double* obj_vals, double* mar_errs,
                     double* run_times, double* densities)
{
    CAPI_BEGIN
    if (!r) fail(E_INVALID, "null result");
    auto cp = [](double* dst, const std::vector<double>& v) {
        if (dst && !v.empty()) std::copy(v.begin(), v.end(), dst);
    };
    cp(dual, r->dual);
    cp(obj_vals, r->obj_vals);
    cp(mar_errs, r->mar_errs);
    cp(run_times, r->run_times);
    cp(densities, r->densities);
    CAPI_END
}

int regot_result_plan(regot_problem* p, const regot_result* r, double* plan)
{
    if (!r) return set_error(E_INVALID, "null result");
    if (!p || p->n != r->n || p->m != r->m)
        return set_error(E_INVALID, "problem and result sizes differ");
    if (r->qrot && !r->plan.p) {
        CAPI_BEGIN
        if (!plan) fail(E_INVALID, "plan is null");
        HIP_CHECK(hipSetDevice(p->device));
        qrot_plan_host(p, r->dual.data(), plan);
        CAPI_END
    }
    if (!r->plan.p) return regot_get_plan(p, r->dual.data(), plan);
    // a solver that keeps its own primal plan (EROT and QROT APDAGD)
    CAPI_BEGIN
    if (!plan) fail(E_INVALID, "plan is null");
    HIP_CHECK(hipSetDevice(r->plan_device));
    HIP_CHECK(hipMemcpy2D(plan, sizeof(double) * r->m, r->plan.p, sizeof(double) * r->plan_ld,
                          sizeof(double) * r->m, r->n, hipMemcpyDeviceToHost));
    CAPI_END
}

int regot_result_verify(regot_problem* p, const regot_result* r, double* mar_err_fp64,
                        double* obj_fp64)
{
    CAPI_BEGIN
    if (!p || !r) fail(E_INVALID, "null handle");
    if (p->n != r->n || p->m != r->m) fail(E_INVALID, "problem and result sizes differ");
    if (r->qrot) fail(E_INVALID, "verify applies to EROT results");
    HIP_CHECK(hipSetDevice(p->device));
    const int n = p->n, m = p->m;
    hipStream_t s = p->stream;
    HIP_CHECK(hipMemcpyAsync(p->gamma, r->dual.data(), (n + m - 1) * sizeof(double),
                             hipMemcpyHostToDevice, s));
    const bool was = p->precise;
    p->precise = true;
    double gn2 = 0.0;
    double f = 0.0;
    try {
        f = p->dev_obj_grad(p->gamma, p->grad, &gn2);
    } catch (...) {
        p->precise = was;
        throw;
    }
    p->precise = was;
    if (obj_fp64) *obj_fp64 = f;
    if (mar_err_fp64) *mar_err_fp64 = std::sqrt(gn2);
    CAPI_END
}

int regot_result_stats(const regot_result* r, double* out)
{
    if (!r || !out) return set_error(E_INVALID, "null argument");
    for (int q = 0; q < 3; q++) {
        out[2 * q] = r->phase_ms[q];
        out[2 * q + 1] = static_cast<double>(r->phase_n[q]);
    }
    out[6] = static_cast<double>(r->cg_iters);
    out[7] = r->alg_bytes;
    return OK;
}

int regot_result_precise_from(const regot_result* r)
{
    if (!r) return set_error(E_INVALID, "null result");
    return r->precise_from;
}

void regot_result_free(regot_result* r) { delete r; }

}  // extern "C"
