// grad.hip -- gradients of EROT losses with respect to the cost matrix, by the
// implicit function theorem on the dual optimality conditions.
//
// At the dual optimum gamma* (g(gamma*, M) = 0, g the dual gradient) a loss
// L(gamma, M) of the plan T = exp((alpha (+) beta - M) / reg) has
//
//   dL/dM = dL/dM|_gamma - w' dg/dM,     H w = dL/dgamma,
//
// with H the dual Hessian (the sparsified Hessian with every entry kept) and
// dg_i/dM_ij = dg_{n+j}/dM_ij = -T_ij / reg, so
//
//   G_ij = dL/dM_ij = T_ij (P_i + Q_j - c M_ij / reg)
//
// for the two losses here:
//   entropic  L = <T, M> - reg h(T), h(T) = sum T (1 - log T), which equals
//             -f(gamma) on the optimal plan (src/sinkhorn_problem_obj.cpp:300-319
//             for the entropy convention): dL/dgamma = H gamma,
//             P_i = 1 + (w_i - alpha_i) / reg, Q_j = (w_{n+j} - beta_j) / reg, c = 0;
//   sharp     L = <T, M>: dL/dalpha_i = sum_j M_ij T_ij / reg, dL/dbeta_j =
//             sum_i M_ij T_ij / reg, P_i = 1 + w_i / reg, Q_j = w_{n+j} / reg, c = 1.
//
// For the entropic loss w = gamma solves H w = H gamma, so G = T: the adjoint
// reproduces the envelope theorem, a built-in consistency check.  H w = rhs is
// solved by the device CG of the sparsified Hessian (shift 0).
#include <algorithm>
#include <cmath>
#include <vector>

#include "hessian.hpp"
#include "regot_hip.h"

namespace regot {

int set_error(int code, const char* what);

namespace {

constexpr double kLog2e = 1.4426950408889634074;

// R_i = sum_j M_ij T_ij (one wave per row, FP64).
template <typename TM>
__global__ __launch_bounds__(256) void k_tm_rows(const TM* __restrict__ M, long ldm, int n, int m,
                                                 const double* __restrict__ alpha,
                                                 const double* __restrict__ beta, double k,
                                                 double* __restrict__ R)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const TM* Mi = M + static_cast<long>(i) * ldm;
    const double ai = alpha[i];
    double s = 0.0;
    for (int j = lane; j < m; j += 64) {
        const double mij = static_cast<double>(Mi[j]);
        s += mij * exp2(k * ((ai + beta[j]) - mij));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) R[i] = s;
}

// C_j = sum_i M_ij T_ij (one thread per column, rows in order).
template <typename TM>
__global__ __launch_bounds__(256) void k_tm_cols(const TM* __restrict__ M, long ldm, int n, int m,
                                                 const double* __restrict__ alpha,
                                                 const double* __restrict__ beta, double k,
                                                 double* __restrict__ C)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= m) return;
    const double bj = beta[j];
    double s = 0.0;
    for (int i = 0; i < n; i++) {
        const double mij = static_cast<double>(M[static_cast<long>(i) * ldm + j]);
        s += mij * exp2(k * ((alpha[i] + bj) - mij));
    }
    C[j] = s;
}

// rhs = [R / reg; C[0:m-1] / reg]
__global__ void k_rhs_sharp(const double* R, const double* C, int n, int m, double reg,
                            double* rhs)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) rhs[t] = R[t] / reg;
    if (t < m - 1) rhs[n + t] = C[t] / reg;
}

// P_i = 1 + (w_i - e alpha_i) / reg, Q_j = (w_{n+j} - e beta_j) / reg (Q_{m-1} = 0)
__global__ void k_pq(const double* w, const double* alpha, const double* beta, int n, int m,
                     double reg, double e, double* P, double* Q)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) P[t] = 1.0 + (w[t] - e * alpha[t]) / reg;
    if (t < m) Q[t] = (t < m - 1) ? (w[n + t] - e * beta[t]) / reg : 0.0;
}

// G_ij = T_ij (P_i + Q_j - c M_ij / reg), T in FP64.
template <typename TM, typename TO>
__global__ __launch_bounds__(256) void k_grad_field(const TM* __restrict__ M, long ldm, int n,
                                                    int m, const double* __restrict__ alpha,
                                                    const double* __restrict__ beta, double k,
                                                    double reg, const double* __restrict__ P,
                                                    const double* __restrict__ Q, double c,
                                                    TO* __restrict__ G, long ldg)
{
    for (int i = blockIdx.y; i < n; i += gridDim.y) {
        const TM* Mi = M + static_cast<long>(i) * ldm;
        TO* Gi = G + static_cast<long>(i) * ldg;
        const double ai = alpha[i], pi = P[i];
        for (int j = blockIdx.x * 256 + threadIdx.x; j < m; j += gridDim.x * 256) {
            const double mij = static_cast<double>(Mi[j]);
            const double t = exp2(k * ((ai + beta[j]) - mij));
            Gi[j] = static_cast<TO>(t * (pi + Q[j] - c * mij / reg));
        }
    }
}

template <typename TM>
void launch_grad(regot_problem* p, int loss, const double* alpha, const double* beta,
                 const double* P, const double* Q, void* G, int G_dtype, long ldg)
{
    const TM* M = static_cast<const TM*>(p->dM);
    const dim3 grid(std::min(ceil_div(p->m, 256), 64), std::min(p->n, 65535));
    const double c = loss == 1 ? 1.0 : 0.0;
    const double k = kLog2e / p->reg;
    if (G_dtype == REGOT_F32)
        hipLaunchKernelGGL((k_grad_field<TM, float>), grid, dim3(256), 0, p->stream, M, p->ldm,
                           p->n, p->m, alpha, beta, k, p->reg, P, Q, c, static_cast<float*>(G),
                           ldg);
    else
        hipLaunchKernelGGL((k_grad_field<TM, double>), grid, dim3(256), 0, p->stream, M, p->ldm,
                           p->n, p->m, alpha, beta, k, p->reg, P, Q, c, static_cast<double*>(G),
                           ldg);
    KERNEL_CHECK();
}

template <typename TM>
void launch_tm(regot_problem* p, const double* alpha, const double* beta, double* R, double* C)
{
    const TM* M = static_cast<const TM*>(p->dM);
    const double k = kLog2e / p->reg;
    hipLaunchKernelGGL(k_tm_rows<TM>, dim3(ceil_div(p->n, 4)), dim3(256), 0, p->stream, M, p->ldm,
                       p->n, p->m, alpha, beta, k, R);
    hipLaunchKernelGGL(k_tm_cols<TM>, dim3(ceil_div(p->m, 256)), dim3(256), 0, p->stream, M,
                       p->ldm, p->n, p->m, alpha, beta, k, C);
    KERNEL_CHECK();
}

}  // namespace
}  // namespace regot

using namespace regot;

extern "C" {

int regot_implicit_grad(regot_problem* p, const double* gamma, int loss, double cg_tol,
                        int cg_max_iter, void* G, int G_dtype, long long ldg, double* loss_value,
                        double* w_out, int* cg_iters, double* cg_err)
{
    try {
        if (!p || !gamma || !G) fail(E_INVALID, "null argument");
        if (loss != 0 && loss != 1) fail(E_INVALID, "loss must be 0 (entropic) or 1 (sharp)");
        if (G_dtype != REGOT_F32 && G_dtype != REGOT_F64) fail(E_INVALID, "bad G dtype");
        if (ldg < p->m) fail(E_INVALID, "ldg %lld < m %d", ldg, p->m);
        HIP_CHECK(hipSetDevice(p->device));
        const int n = p->n, m = p->m;
        const long N = static_cast<long>(n) + m - 1;
        hipStream_t s = p->stream;
        HIP_CHECK(hipMemcpyAsync(p->gamma, gamma, N * sizeof(double), hipMemcpyHostToDevice, s));
        // f and grad at gamma (FP64 sums), then the full dual Hessian
        const bool was = p->precise;
        p->precise = true;
        double f = 0.0;
        try {
            f = p->dev_obj_grad(p->gamma, p->grad, nullptr);
        } catch (...) {
            p->precise = was;
            throw;
        }
        p->precise = was;
        hess_topk(p, p->gamma, p->grad, 1.0, true);
        DBuf<double> alpha(n), beta(m), rhs(N), w(N), R(n), C(m), P(n), Q(m);
        launch_split_gamma(p->gamma, n, m, alpha, beta, s);
        double lv = 0.0;
        if (loss == 0) {
            hess_apply_hs(p, p->gamma, rhs);
            lv = -f;
        } else {
            if (p->bulk == REGOT_F32)
                launch_tm<float>(p, alpha, beta, R, C);
            else
                launch_tm<double>(p, alpha, beta, R, C);
            hipLaunchKernelGGL(k_rhs_sharp, dim3(ceil_div(std::max(n, m), 256)), dim3(256), 0, s,
                               R.p, C.p, n, m, p->reg, rhs.p);
            KERNEL_CHECK();
            std::vector<double> hR(n);
            HIP_CHECK(hipMemcpyAsync(hR.data(), R.p, n * sizeof(double), hipMemcpyDeviceToHost,
                                     s));
            p->sync();
            for (int i = 0; i < n; i++) lv += hR[i];
        }
        const double tol = cg_tol > 0.0 ? cg_tol : 1e-10;
        const int maxit = cg_max_iter > 0 ? cg_max_iter : 10 * m + 1000;
        double err = 0.0;
        // For the entropic loss H gamma = rhs: the dual itself is the warm start
        // (the CG then only confirms the residual); the sharp loss starts at 0.
        const int it = hess_cg(p, rhs, 0.0, tol, maxit, w, &err, loss == 0 ? p->gamma.p : nullptr,
                               true);
        hipLaunchKernelGGL(k_pq, dim3(ceil_div(std::max(n, m), 256)), dim3(256), 0, s, w.p,
                           alpha.p, beta.p, n, m, p->reg, loss == 0 ? 1.0 : 0.0, P.p, Q.p);
        KERNEL_CHECK();
        if (p->bulk == REGOT_F32)
            launch_grad<float>(p, loss, alpha, beta, P, Q, G, G_dtype, static_cast<long>(ldg));
        else
            launch_grad<double>(p, loss, alpha, beta, P, Q, G, G_dtype, static_cast<long>(ldg));
        if (w_out)
            HIP_CHECK(hipMemcpyAsync(w_out, w.p, N * sizeof(double), hipMemcpyDeviceToHost, s));
        p->sync();
        if (loss_value) *loss_value = lv;
        if (cg_iters) *cg_iters = it;
        if (cg_err) *cg_err = err;
    } catch (const regot::Error& e) {
        return set_error(e.code, e.what());
    } catch (const std::bad_alloc&) {
        return set_error(E_NOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return set_error(E_INTERNAL, e.what());
    }
    return OK;
}

}  // extern "C"
