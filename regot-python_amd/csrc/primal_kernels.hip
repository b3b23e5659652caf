// primal_kernels.hip -- feasible rounding of T(gamma) and the primal value.
//
// approx_proj (src/approx_proj.h:6-26; Altschuler, Weed and Rigollet 2017,
// Algorithm 2) rounds a plan onto the transport polytope of (a, b):
//   x = min(a ./ rowsum(T), 1);  T' = diag(x) T
//   y = min(b ./ colsum(T'), 1); T'' = T' diag(y)
//   P = T'' + err_a err_b' / |err_a|_1,  err_a = a - rowsum(T''),
//                                        err_b = b - colsum(T'').
// Problem::primal_val (src/sinkhorn_problem_obj.cpp:300-319) is
// <P, M> - reg sum P (1 - log P) for T = T(gamma).  T is never stored: three
// weighted-sum passes over M (row sums of T; column sums of x_i T_ij; row and
// column sums of x_i T_ij y_j) give x, y, err_a and err_b, and a fourth pass
// forms P element by element, accumulating <P, M> and the entropy term (and
// writing P when the caller wants the plan).  FP64 throughout, T by the
// device exp2 of the one-pass objective kernel; every reduction is in a
// fixed order.
#include <algorithm>

#include "dexp2.hpp"
#include "problem.hpp"
#include "qrot.hpp"
#include "regot_hip.h"

namespace regot {

int set_error(int code, const char* what);

namespace {

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ double bsum(double v, double* sm4)
{
    v = wsum(v);
    if ((threadIdx.x & 63) == 0) sm4[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = ((sm4[0] + sm4[1]) + sm4[2]) + sm4[3];
    __syncthreads();
    return r;
}

// out[t] = sum over s of part[s * stride + t] (fixed order), t < count.
__global__ void k_fold(const double* __restrict__ part, int nparts, long stride, int count,
                       double* __restrict__ out)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= count) return;
    double v = 0.0;
    for (int s = 0; s < nparts; s++) v += part[static_cast<long>(s) * stride + t];
    out[t] = v;
}

// w = min(target ./ sums, 1)
__global__ void k_scale(const double* __restrict__ target, const double* __restrict__ sums,
                        int count, double* __restrict__ w)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < count) w[t] = fmin(target[t] / sums[t], 1.0);
}

// err = target - sums; l1part[wg] = this workgroup's sum of |err|.
__global__ __launch_bounds__(256) void k_err(const double* __restrict__ target,
                                             const double* __restrict__ sums, int count,
                                             double* __restrict__ err,
                                             double* __restrict__ l1part)
{
    __shared__ double sm4[4];
    const int t = blockIdx.x * 256 + threadIdx.x;
    double e = 0.0;
    if (t < count) {
        e = target[t] - sums[t];
        err[t] = e;
    }
    const double s = bsum(fabs(e), sm4);
    if (l1part && threadIdx.x == 0) l1part[blockIdx.x] = s;
}

// One workgroup: out[0] = sum of the first count doubles of part.
__global__ __launch_bounds__(256) void k_total(const double* __restrict__ part, int count,
                                               int width, double* __restrict__ out)
{
    __shared__ double sm4[4];
    for (int c = 0; c < width; c++) {
        double v = 0.0;
        for (int i = threadIdx.x; i < count; i += 256) v += part[static_cast<long>(i) * width + c];
        v = bsum(v, sm4);
        if (threadIdx.x == 0) out[c] = v;
    }
}

// P_ij = (T_ij x_i) y_j + err_a_i err_b_j / l1 (the reference's order of
// operations); scal[2 wg] = sum P M, scal[2 wg + 1] = sum P (1 - log P) of the
// workgroup; P stored when non-null (pitch ldp).  Tiles as k_sums_tile.
template <typename TM>
__global__ __launch_bounds__(256) void k_round_tile(const TM* __restrict__ M, long ldm, int n,
                                                    int m, const double* __restrict__ alpha,
                                                    const double* __restrict__ beta, double k,
                                                    int rows, const double* __restrict__ x,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ ea,
                                                    const double* __restrict__ eb,
                                                    const double* __restrict__ l1p,
                                                    double* __restrict__ P, long ldp,
                                                    double* __restrict__ scal)
{
    __shared__ __align__(16) double tab[128];
    __shared__ double ss[2][4];
    dev::exp2_tab_load(tab);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int strip = blockIdx.x, blk = blockIdx.y;
    const int c0 = strip * 256 + lane * 4;
    const int r0 = blk * rows;
    const int r1 = min(n, r0 + rows);
    const double l1 = *l1p;
    double bk[4], yq[4], eq[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const bool in = c0 + q < m;
        bk[q] = in ? beta[c0 + q] : 0.0;
        yq[q] = in ? y[c0 + q] : 0.0;
        eq[q] = in ? eb[c0 + q] : 0.0;
    }
    double mp = 0.0, en = 0.0;
    for (int r = r0 + wave; r < r1; r += 4) {
        const double ai = alpha[r], xi = x[r], ei = ea[r];
        const TM* row = M + static_cast<long>(r) * ldm;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int j = c0 + q;
            if (j < m) {
                const double mij = static_cast<double>(row[j]);
                const double t = dev::dexp2((ai + bk[q] - mij) * k, tab);
                const double v = (t * xi) * yq[q] + ei * eq[q] / l1;
                if (P) P[static_cast<long>(r) * ldp + j] = v;
                mp += v * mij;
                en += v * (1.0 - log(v));
            }
        }
    }
    mp = wsum(mp);
    en = wsum(en);
    if (lane == 0) {
        ss[0][wave] = mp;
        ss[1][wave] = en;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const long wg = static_cast<long>(blk) * gridDim.x + strip;
        scal[2 * wg] = ((ss[0][0] + ss[0][1]) + ss[0][2]) + ss[0][3];
        scal[2 * wg + 1] = ((ss[1][0] + ss[1][1]) + ss[1][2]) + ss[1][3];
    }
}

}  // namespace

// The rounded plan of T(gamma) (device gamma) and its primal value.  P
// (device, pitch ldp) may be null; out[0] = <P, M>, out[1] = sum P (1 - log P)
// (host).
static void round_plan(regot_problem* p, const double* d_gamma, double* P, long ldp,
                       double* out)
{
    const int n = p->n, m = p->m;
    hipStream_t s = p->stream;
    const int nstrip = ceil_div(m, sums_strip_width());   // weighted-sum passes
    const int rows = sums_rows(n, m);
    const int nblk = ceil_div(n, rows);
    const int nstrip4 = ceil_div(m, 256);                  // rounding pass (4 columns per lane)
    DBuf<double> alpha(n), beta(m), rowpart(static_cast<size_t>(nstrip) * n),
        colpart(static_cast<size_t>(nblk) * m), rsum(n), csum(m), x(n), y(m), ea(n), eb(m),
        l1part(ceil_div(n, 256)), l1(1), scal(2L * nstrip4 * nblk), tot(2);
    launch_split_gamma(d_gamma, n, m, alpha, beta, s);
    const dim3 gn(ceil_div(n, 256)), gm(ceil_div(m, 256));
    // x = min(a ./ rowsum(T), 1)
    launch_wsums(p->bulk, p->dM, p->ldm, n, m, alpha, beta, p->reg, nullptr, nullptr, rowpart,
                 nullptr, s);
    hipLaunchKernelGGL(k_fold, gn, dim3(256), 0, s, rowpart.p, nstrip, static_cast<long>(n), n,
                       rsum.p);
    hipLaunchKernelGGL(k_scale, gn, dim3(256), 0, s, p->a.p, rsum.p, n, x.p);
    KERNEL_CHECK();
    // y = min(b ./ colsum(diag(x) T), 1)
    launch_wsums(p->bulk, p->dM, p->ldm, n, m, alpha, beta, p->reg, x, nullptr, rowpart, colpart,
                 s);
    hipLaunchKernelGGL(k_fold, gm, dim3(256), 0, s, colpart.p, nblk, static_cast<long>(m), m,
                       csum.p);
    hipLaunchKernelGGL(k_scale, gm, dim3(256), 0, s, p->b.p, csum.p, m, y.p);
    KERNEL_CHECK();
    // err_a, err_b of diag(x) T diag(y); |err_a|_1
    launch_wsums(p->bulk, p->dM, p->ldm, n, m, alpha, beta, p->reg, x, y, rowpart, colpart, s);
    hipLaunchKernelGGL(k_fold, gn, dim3(256), 0, s, rowpart.p, nstrip, static_cast<long>(n), n,
                       rsum.p);
    hipLaunchKernelGGL(k_fold, gm, dim3(256), 0, s, colpart.p, nblk, static_cast<long>(m), m,
                       csum.p);
    hipLaunchKernelGGL(k_err, gn, dim3(256), 0, s, p->a.p, rsum.p, n, ea.p, l1part.p);
    hipLaunchKernelGGL(k_err, gm, dim3(256), 0, s, p->b.p, csum.p, m, eb.p,
                       static_cast<double*>(nullptr));
    hipLaunchKernelGGL(k_total, dim3(1), dim3(256), 0, s, l1part.p, ceil_div(n, 256), 1, l1.p);
    KERNEL_CHECK();
    // P, <P, M>, entropy
    const double k = 1.4426950408889634074 / p->reg;
    const dim3 grid(nstrip4, nblk);
    if (p->bulk == 0)
        hipLaunchKernelGGL(k_round_tile<float>, grid, dim3(256), 0, s,
                           static_cast<const float*>(p->dM), p->ldm, n, m, alpha.p, beta.p, k,
                           rows, x.p, y.p, ea.p, eb.p, l1.p, P, ldp, scal.p);
    else
        hipLaunchKernelGGL(k_round_tile<double>, grid, dim3(256), 0, s,
                           static_cast<const double*>(p->dM), p->ldm, n, m, alpha.p, beta.p, k,
                           rows, x.p, y.p, ea.p, eb.p, l1.p, P, ldp, scal.p);
    KERNEL_CHECK();
    hipLaunchKernelGGL(k_total, dim3(1), dim3(256), 0, s, scal.p, nstrip4 * nblk, 2, tot.p);
    KERNEL_CHECK();
    HIP_CHECK(hipMemcpyAsync(out, tot.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

// QROT: the same rounding of plan_ij = (alpha_i + beta_j - M_ij)_+ / reg
// (Problem::primal_val, src/qrot_problem.cpp:10-26): three weighted-sum passes
// and the rounding pass, <P, M> + reg/2 ||P||^2.
double qrot_round_plan(regot_problem* p, const double* alpha, const double* beta, double* P,
                       long ldp)
{
    const int n = p->n, m = p->m;
    hipStream_t s = p->stream;
    const int rows = qrot_rows(n, m);
    const int nstrip = ceil_div(m, 256);
    const int nblk = ceil_div(n, rows);
    if (!p->qrot) p->qrot = std::make_unique<QrotScratch>(n, m);
    QrotScratch& w = *p->qrot;
    DBuf<double>&rowpart = w.rowpart, &colpart = w.colpart, &rsum = w.rsum, &csum = w.csum,
    &x = w.x, &y = w.y, &ea = w.ea, &eb = w.eb, &l1part = w.l1part, &l1 = w.l1, &scal = w.scal,
    &tot = w.tot;
    const dim3 gn(ceil_div(n, 256)), gm(ceil_div(m, 256));
    const double reg = p->reg;
    launch_qrot_tile(p->bulk, p->dM, p->ldm, n, m, alpha, beta, reg, nullptr, nullptr, rowpart,
                     nullptr, nullptr, s);
    hipLaunchKernelGGL(k_fold, gn, dim3(256), 0, s, rowpart.p, nstrip, static_cast<long>(n), n,
                       rsum.p);
    hipLaunchKernelGGL(k_scale, gn, dim3(256), 0, s, p->a.p, rsum.p, n, x.p);
    KERNEL_CHECK();
    launch_qrot_tile(p->bulk, p->dM, p->ldm, n, m, alpha, beta, reg, x, nullptr, nullptr,
                     colpart, nullptr, s);
    hipLaunchKernelGGL(k_fold, gm, dim3(256), 0, s, colpart.p, nblk, static_cast<long>(m), m,
                       csum.p);
    hipLaunchKernelGGL(k_scale, gm, dim3(256), 0, s, p->b.p, csum.p, m, y.p);
    KERNEL_CHECK();
    launch_qrot_tile(p->bulk, p->dM, p->ldm, n, m, alpha, beta, reg, x, y, rowpart, colpart,
                     nullptr, s);
    hipLaunchKernelGGL(k_fold, gn, dim3(256), 0, s, rowpart.p, nstrip, static_cast<long>(n), n,
                       rsum.p);
    hipLaunchKernelGGL(k_fold, gm, dim3(256), 0, s, colpart.p, nblk, static_cast<long>(m), m,
                       csum.p);
    hipLaunchKernelGGL(k_err, gn, dim3(256), 0, s, p->a.p, rsum.p, n, ea.p, l1part.p);
    hipLaunchKernelGGL(k_err, gm, dim3(256), 0, s, p->b.p, csum.p, m, eb.p,
                       static_cast<double*>(nullptr));
    hipLaunchKernelGGL(k_total, dim3(1), dim3(256), 0, s, l1part.p, ceil_div(n, 256), 1, l1.p);
    KERNEL_CHECK();
    launch_qrot_round(p->bulk, p->dM, p->ldm, n, m, alpha, beta, reg, x, y, ea, eb, l1, P, ldp,
                      scal, s);
    hipLaunchKernelGGL(k_total, dim3(1), dim3(256), 0, s, scal.p, nstrip * nblk, 2, tot.p);
    KERNEL_CHECK();
    double out[2];
    HIP_CHECK(hipMemcpyAsync(out, tot.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return out[0] + 0.5 * reg * out[1];
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

extern "C" {

int regot_primal_val(regot_problem* p, const double* gamma, double* value)
{
    CAPI_BEGIN
    if (!p || !gamma || !value) fail(E_INVALID, "null argument");
    HIP_CHECK(hipSetDevice(p->device));
    const int N = p->n + p->m - 1;
    HIP_CHECK(hipMemcpyAsync(p->gamma, gamma, N * sizeof(double), hipMemcpyHostToDevice,
                             p->stream));
    double out[2];
    round_plan(p, p->gamma, nullptr, 0, out);
    *value = out[0] - p->reg * out[1];
    CAPI_END
}

int regot_round_plan(regot_problem* p, const double* gamma, double* plan)
{
    CAPI_BEGIN
    if (!p || !gamma || !plan) fail(E_INVALID, "null argument");
    HIP_CHECK(hipSetDevice(p->device));
    const int n = p->n, m = p->m;
    HIP_CHECK(hipMemcpyAsync(p->gamma, gamma, (n + m - 1) * sizeof(double),
                             hipMemcpyHostToDevice, p->stream));
    DBuf<double> P(static_cast<size_t>(n) * m);
    double out[2];
    round_plan(p, p->gamma, P, m, out);
    HIP_CHECK(hipMemcpy(plan, P.p, sizeof(double) * static_cast<size_t>(n) * m,
                        hipMemcpyDeviceToHost));
    CAPI_END
}

}  // extern "C"
