// qrot.hpp -- quadratically regularised OT (QROT) on the device: the dual
// passes (qrot_kernels.hip), the rounded primal value (primal_kernels.hip)
// and the host drivers (qrot.cpp).  The QROT dual variable is
// gamma = [alpha (n); beta (m)]: no elimination of beta_m.
#pragma once

#include <memory>

#include "problem.hpp"

namespace regot {

int qrot_rows(int n, int m);

// The generalised Hessian H = [diag(h1) sigma; sigma' diag(h2)] at one gamma
// (src/qrot_hess.h:11-36): sigma as a bitmask (launch_qrot_mask), h1 = sigma 1,
// h2 = sigma' 1, plus the vectors of its block solve.
struct QrotHess {
    DBuf<unsigned long long> mask;
    DBuf<double> h1, h2, hinv, tn, tn2, tm, tm2, ones, v, r, pv, q, dsc, rowpart4;
    HBuf<double> hs;
    double hinv_shift = -1.0;
    QrotHess(int n, int m)
    {
        mask.alloc(static_cast<size_t>(n) * ceil_div(m, 256) * 4);
        rowpart4.alloc(static_cast<size_t>(n) * ceil_div(m, 256) * 4);
        const int c = n > m ? n : m;
        h1.alloc(n);
        h2.alloc(m);
        hinv.alloc(n);
        tn.alloc(n);
        tn2.alloc(n);
        v.alloc(n);
        tm.alloc(m);
        tm2.alloc(m);
        r.alloc(m);
        pv.alloc(m);
        q.alloc(m);
        ones.alloc(c);
        dsc.alloc(8);
        hs.ensure(8);
    }
};

// Device scratch of the QROT passes and root iterations, sized for a shape;
// owned by the problem handle.
struct QrotScratch {
    DBuf<double> rowpart, colpart, sqpart, rsum, csum, sq, out, spart, fpart, gpart;
    // rounding (qrot_round_plan): row / column scales, errors, |err_a|_1, totals
    DBuf<double> x, y, ea, eb, l1part, l1, scal, tot;
    DBuf<int> kpart, kprev, done, open;
    HBuf<double> hout;
    HBuf<int> hopen;
    std::unique_ptr<QrotHess> hess;
    QrotScratch(int n, int m)
    {
        const int rows = qrot_rows(n, m);
        const long nstrip = ceil_div(m, 256), nblk = ceil_div(n, rows);
        rowpart.alloc(static_cast<size_t>(nstrip) * n);
        colpart.alloc(static_cast<size_t>(nblk) * m);
        sqpart.alloc(static_cast<size_t>(nstrip) * nblk);
        const size_t a = static_cast<size_t>(nblk) * m, b = static_cast<size_t>(nstrip) * n;
        spart.alloc(a > b ? a : b);
        kpart.alloc(a > b ? a : b);
        rsum.alloc(n);
        csum.alloc(m);
        sq.alloc(1);
        out.alloc(2);
        const int nw = ceil_div(n > m ? n : m, 256);
        fpart.alloc(4 * static_cast<size_t>(nw));
        gpart.alloc(nw);
        const int c = n > m ? n : m;
        kprev.alloc(c);
        done.alloc(c);
        open.alloc(1);
        hout.ensure(2);
        hopen.ensure(1);
        x.alloc(n);
        y.alloc(m);
        ea.alloc(n);
        eb.alloc(m);
        l1part.alloc(ceil_div(n, 256));
        l1.alloc(1);
        scal.alloc(2 * static_cast<size_t>(nstrip) * nblk);
        tot.alloc(2);
    }
};
// Row sums (per 256-column strip: rowpart[strip * n + i]), column sums (per
// row block: colpart[blk * m + j]) and sums of squares (sqpart[blk * strips +
// strip]) of ((D / div) x_i) y_j, D = (alpha_i + beta_j - M_ij)_+; x, y and
// any output may be null.
void launch_qrot_tile(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                      const double* beta, double div, const double* x, const double* y,
                      double* rowpart, double* colpart, double* sqpart, hipStream_t s);
// P_ij = ((D_ij / reg) x_i) y_j + ea_i eb_j / l1; scal[2 wg] = sum P M,
// scal[2 wg + 1] = sum P^2 per workgroup; P written when non-null.
void launch_qrot_round(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                       const double* beta, double reg, const double* x, const double* y,
                       const double* ea, const double* eb, const double* l1, double* P,
                       long ldp, double* scal, hipStream_t s);
void launch_qrot_pl_init(double* xc, int* kprev, int* done, int count, int* open, hipStream_t s);
// One Newton step of the per-column (cols) or per-row roots; *open counts the
// entries still moving.
void launch_qrot_pl_step(int bulk, const void* M, long ldm, int n, int m, bool cols,
                         const double* pot, const double* w, double reg, double* xc,
                         int* kprev, int* done, double* spart, int* kpart, int* open,
                         hipStream_t s);
void launch_qrot_plan(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                      const double* beta, double reg, double* plan, long r0, int nr,
                      hipStream_t s);
// f into out[0] and (with grad) grad and ||grad||^2 into out[1]; part holds
// 4 ceil(max(n, m) / 256) doubles, gpart ceil(max(n, m) / 256).
void launch_qrot_finish(const double* alpha, const double* beta, const double* a,
                        const double* b, int n, int m, double reg, double tau, const double* sq,
                        const double* rsum, const double* csum, double* grad, double* part,
                        double* gpart, double* out, hipStream_t s);

void launch_qrot_fold(const double* part, int nparts, long stride, int count, double* out,
                      hipStream_t s);
// APDAGD's / PDAAM's plan pass P = cold P + tau plan(alpha, beta) (init: P = plan),
// then out[0] = <M, P>, out[1] = ||P||^2, out[2] = ||P1 - a||^2 + ||P'1 - b||^2
// (device); scratch as documented at the definition.
void launch_qrot_apd_plan(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                          const double* beta, double reg, double cold, double tau, bool init,
                          double* P,
                          long ldp, const double* a, const double* b, double* rowpart,
                          double* colpart, double* rsum, double* csum, double* wpart,
                          double* out, hipStream_t s);
void launch_qrot_sum1(const double* part, int count, double* out, hipStream_t s);

// sigma (bitmask, see QrotHess) at alpha, beta
void launch_qrot_mask(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                      const double* beta, unsigned long long* mask, hipStream_t s);
// Row sums of sigma_ij y_j (rows_out) into rowpart[(strip * 4 + k) * n + i]
// (4 partials per 256-column strip: fold 4 ceil(m / 256) of them) and column
// sums of sigma_ij x_i (cols_out) into colpart as launch_qrot_tile; null x or
// y count ones.
void launch_qrot_sig(const unsigned long long* mask, int n, int m, bool rows_out, bool cols_out,
                     const double* x, const double* y, double* rowpart, double* colpart,
                     hipStream_t s);
void launch_qrot_ainv(const double* x, const double* h, double shift, double c, double* out,
                      int count, hipStream_t s);
void launch_qrot_delta_fin(const double* x, const double* h, double shift, const double* t,
                           double c, double* out, int count, hipStream_t s);
void launch_qrot_recip(const double* h, double shift, double* out, int count, hipStream_t s);
// out_i = (t_i + h_i x_i) + c
void launch_qrot_hfin(const double* x, const double* h, const double* t, double c, double* out,
                      int count, hipStream_t s);

// The generalised Hessian at device gamma (n + m) into the problem's QrotHess.
QrotHess& qrot_hess_build(regot_problem* p, const double* gamma);
// out (device, n + m) = (H + shift I + tau vv')^-1 rhs, v = [1; -1], by block
// elimination with CG on the Schur complement (Eigen's rules: from x0
// (device, m; may alias out + n) or zero, ||r||^2 < tol^2 ||rhs_beta||^2);
// returns the CG steps.
int qrot_hess_solve(regot_problem* p, QrotHess& H, const double* rhs, double shift, double tau,
                    double tol, int max_iter, double* out, const double* x0 = nullptr);
// out (device, n + m) = (H + tau vv') x (src/qrot_hess.cpp:193-214, shift 0).
void apply_qrot_hess(regot_problem* p, QrotHess& H, const double* x, double tau, double* out);

// f (and grad, ||grad||^2 when grad is non-null) of the QROT dual at device
// gamma (n + m), src/qrot_problem.cpp:226-258.
double qrot_obj_grad(regot_problem* p, const double* gamma, double* grad, double tau,
                     double* gnorm2);
// optimal_beta (cols = true: out = beta (m) given alpha = pot) or
// optimal_alpha (cols = false: out = alpha (n) given beta = pot), device.
void qrot_optimal(regot_problem* p, bool cols, const double* pot, double* out);

// <P, M> + reg/2 ||P||^2 of P = approx_proj((alpha (+) beta - M)_+ / reg)
// for device alpha (n), beta (m); P (device, pitch ldp) written when non-null.
double qrot_round_plan(regot_problem* p, const double* alpha, const double* beta, double* P,
                       long ldp);
// The QROT plan (alpha (+) beta - M)+ / reg of a host dual (n + m) into host memory.
void qrot_plan_host(regot_problem* p, const double* dual, double* plan);

}  // namespace regot
