// apdagd_kernels.hip -- the primal-plan pass of APDAGD.
//
// APDAGD keeps an explicit primal plan, a running combination of the plans of
// its dual iterates:  P <- (1 - tau) P + tau T(gamma),
// T_ij = exp((alpha_i + beta_j - M_ij) / reg)
// (src/sinkhorn_apdagd.cpp:138-143), and tests convergence on the plan's
// marginals and primal value  <M, P> - reg sum P (1 - log P)  (:146-152).
// One streaming pass does all of it: read M and P, write P, and reduce row
// sums, column sums, <M, P> and the entropy term, in FP64 throughout (T with
// an FP64 exponential regardless of the bulk precision, as the reference
// forms it: 2^((alpha_i + beta_j - M_ij) log2(e) / reg) by dexp2.hpp).
//
// Tiling: a workgroup owns a 256-column strip of a block of rows (16-256);
// each lane owns 4 adjacent columns (one 16-byte load of FP32 M, two of P per
// row) and its 4 waves take the block's rows round robin, so every wave load
// is a coalesced row segment.  Row sums are reduced across the wave per row
// and written per strip; column sums are accumulated per lane down the block
// and merged through LDS per block.  A finish kernel folds the strips and
// blocks in a fixed order (no atomics: reproducible).
#include <algorithm>

#include "dexp2.hpp"
#include "dlog.hpp"
#include "problem.hpp"

namespace regot {
namespace {

// Rows per workgroup: enough workgroups to fill the chip (>= 2048 at 4 waves
// each), at most 256 rows so the column partials stay small (<= n/16 blocks).
int apd_rows(int n, int m)
{
    const long nstrip = ceil_div(m, 256);
    long r = static_cast<long>(n) * nstrip / 2048;
    r = std::max(16L, std::min(256L, r));
    return static_cast<int>((r + 3) / 4 * 4);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ void ld4(const float* p, int c0, int m, double (&v)[4])
{
    if (c0 + 3 < m) {
        const float4 u = *reinterpret_cast<const float4*>(p + c0);
        v[0] = u.x;
        v[1] = u.y;
        v[2] = u.z;
        v[3] = u.w;
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = (c0 + q < m) ? static_cast<double>(p[c0 + q]) : 0.0;
    }
}

__device__ __forceinline__ void ld4(const double* p, int c0, int m, double (&v)[4])
{
    if (c0 + 3 < m) {
        const double2 u0 = *reinterpret_cast<const double2*>(p + c0);
        const double2 u1 = *reinterpret_cast<const double2*>(p + c0 + 2);
        v[0] = u0.x;
        v[1] = u0.y;
        v[2] = u1.x;
        v[3] = u1.y;
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = (c0 + q < m) ? p[c0 + q] : 0.0;
    }
}

__device__ __forceinline__ void st4(double* p, int c0, int m, const double (&v)[4])
{
    if (c0 + 3 < m) {
        *reinterpret_cast<double2*>(p + c0) = make_double2(v[0], v[1]);
        *reinterpret_cast<double2*>(p + c0 + 2) = make_double2(v[2], v[3]);
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (c0 + q < m) p[c0 + q] = v[q];
    }
}

// grid (strips = ceil(m / 256), row blocks = ceil(n / rows)), 256 threads.
// beta has m entries (beta_{m-1} = 0).  init = 1: P = T (P is not read).
// rowpart[strip * n + i], colpart[block * m + j], scal[2 * wg + {0, 1}] =
// this workgroup's sum of M P and of P (1 - log P).
template <typename TM>
__global__ __launch_bounds__(256) void k_apd_plan(const TM* __restrict__ M, long ldm, int n, int m,
                                                  const double* __restrict__ alpha,
                                                  const double* __restrict__ beta, int rows,
                                                  double k, double tau, int init,
                                                  double* __restrict__ P,
                                                  long ldp, double* __restrict__ rowpart,
                                                  double* __restrict__ colpart,
                                                  double* __restrict__ scal)
{
    __shared__ double sm[4][256];
    __shared__ double ss[2][4];
    __shared__ __align__(16) double tab[128];
    __shared__ double ltab[384];
    dev::exp2_tab_load(tab);
    dev::log_tab_load(ltab);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int strip = blockIdx.x, blk = blockIdx.y;
    const int c0 = strip * 256 + lane * 4;
    const int r0 = blk * rows;
    const int r1 = min(n, r0 + rows);
    double b4[4];
    ld4(beta, c0, m, b4);
    double cs[4] = {0.0, 0.0, 0.0, 0.0};
    double mp = 0.0, en = 0.0;
    const double keep = 1.0 - tau;
    for (int r = r0 + wave; r < r1; r += 4) {
        const double ai = alpha[r];
        double mv[4], pv[4] = {0.0, 0.0, 0.0, 0.0};
        ld4(M + static_cast<long>(r) * ldm, c0, m, mv);
        double* prow = P + static_cast<long>(r) * ldp;
        if (!init) ld4(prow, c0, m, pv);
        double rs = 0.0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (c0 + q < m) {
                const double t = dev::dexp2((ai + b4[q] - mv[q]) * k, tab);
                const double v = init ? t : keep * pv[q] + tau * t;
                pv[q] = v;
                rs += v;
                cs[q] += v;
                mp += mv[q] * v;
                en += v * (1.0 - dev::dlog(v, ltab));
            }
        }
        if (c0 < m) st4(prow, c0, m, pv);
        rs = wave_sum(rs);
        if (lane == 0) rowpart[static_cast<long>(strip) * n + r] = rs;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) sm[wave][lane * 4 + q] = cs[q];
    mp = wave_sum(mp);
    en = wave_sum(en);
    if (lane == 0) {
        ss[0][wave] = mp;
        ss[1][wave] = en;
    }
    __syncthreads();
    const int j = strip * 256 + threadIdx.x;
    if (j < m)
        colpart[static_cast<long>(blk) * m + j] =
            ((sm[0][threadIdx.x] + sm[1][threadIdx.x]) + sm[2][threadIdx.x]) + sm[3][threadIdx.x];
    if (threadIdx.x == 0) {
        const long wg = static_cast<long>(blk) * gridDim.x + strip;
        scal[2 * wg] = ((ss[0][0] + ss[0][1]) + ss[0][2]) + ss[0][3];
        scal[2 * wg + 1] = ((ss[1][0] + ss[1][1]) + ss[1][2]) + ss[1][3];
    }
}

__device__ __forceinline__ double block_sum256(double v, double* sm4)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) sm4[threadIdx.x >> 6] = v;
    __syncthreads();
    const double r = ((sm4[0] + sm4[1]) + sm4[2]) + sm4[3];
    __syncthreads();
    return r;
}

// Thread t < n: row sum of row t over the strips; t < m: column sum of column
// t over the blocks.  part[wg] = sum over the workgroup's threads of
// (r_t - a_t)^2 + (c_t - b_t)^2.
__global__ __launch_bounds__(256) void k_apd_marg(const double* __restrict__ rowpart, int nstrip,
                                                  const double* __restrict__ colpart, int nblk,
                                                  const double* __restrict__ a,
                                                  const double* __restrict__ b, int n, int m,
                                                  double* __restrict__ rowsum,
                                                  double* __restrict__ colsum,
                                                  double* __restrict__ part)
{
    __shared__ double sm4[4];
    const int t = blockIdx.x * 256 + threadIdx.x;
    double e = 0.0;
    if (t < n) {
        double r = 0.0;
        // eight partials in flight, added in order
        int s = 0;
        for (; s + 8 <= nstrip; s += 8) {
            double x[8];
#pragma unroll
            for (int u = 0; u < 8; u++) x[u] = rowpart[static_cast<long>(s + u) * n + t];
#pragma unroll
            for (int u = 0; u < 8; u++) r += x[u];
        }
        for (; s < nstrip; s++) r += rowpart[static_cast<long>(s) * n + t];
        if (rowsum) rowsum[t] = r;
        e += (r - a[t]) * (r - a[t]);
    }
    if (t < m) {
        double c = 0.0;
        int q = 0;
        for (; q + 8 <= nblk; q += 8) {
            double x[8];
#pragma unroll
            for (int u = 0; u < 8; u++) x[u] = colpart[static_cast<long>(q + u) * m + t];
#pragma unroll
            for (int u = 0; u < 8; u++) c += x[u];
        }
        for (; q < nblk; q++) c += colpart[static_cast<long>(q) * m + t];
        if (colsum) colsum[t] = c;
        e += (c - b[t]) * (c - b[t]);
    }
    e = block_sum256(e, sm4);
    if (threadIdx.x == 0) part[blockIdx.x] = e;
}

// One workgroup: out[0] = squared marginal error, out[1] = <M, P>,
// out[2] = sum P (1 - log P).
__global__ __launch_bounds__(256) void k_apd_final(const double* __restrict__ part, int npart,
                                                   const double* __restrict__ scal, long nwg,
                                                   double* __restrict__ out)
{
    __shared__ double sm4[4];
    double e = 0.0, mp = 0.0, en = 0.0;
    for (int i = threadIdx.x; i < npart; i += 256) e += part[i];
    for (long i = threadIdx.x; i < nwg; i += 256) {
        mp += scal[2 * i];
        en += scal[2 * i + 1];
    }
    e = block_sum256(e, sm4);
    mp = block_sum256(mp, sm4);
    en = block_sum256(en, sm4);
    if (threadIdx.x == 0) {
        out[0] = e;
        out[1] = mp;
        out[2] = en;
    }
}

__global__ __launch_bounds__(256) void k_log_check(const double* __restrict__ x,
                                                   double* __restrict__ y, long count)
{
    __shared__ double ltab[384];
    dev::log_tab_load(ltab);
    __syncthreads();
    for (long i = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; i < count;
         i += static_cast<long>(gridDim.x) * 256)
        y[i] = dev::dlog(x[i], ltab);
}

}  // namespace

void launch_log_check(const double* x, double* y, long count, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_log_check, dim3(std::min(4096L, (count + 255) / 256)), dim3(256), 0, s,
                       x, y, count);
    KERNEL_CHECK();
}

// Scratch sizes (doubles) for a plan pass on an n x m problem.
void apd_scratch_sizes(int n, int m, long* rowpart, long* colpart, long* scal, long* part)
{
    const long nstrip = ceil_div(m, 256), nblk = ceil_div(n, apd_rows(n, m));
    *rowpart = nstrip * n;
    *colpart = nblk * m;
    *scal = 2 * nstrip * nblk;
    *part = ceil_div(std::max(n, m), 256);
}

void launch_apd_plan(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                     const double* beta, const double* a, const double* b, double reg,
                     double tau, bool init, double* P, long ldp, double* rowpart,
                     double* colpart, double* scal, double* part, double* out, hipStream_t s)
{
    if (n <= 0 || m <= 0) fail(E_INVALID, "plan pass on an empty problem");
    if (ldp % 2 != 0 || ldp < m) fail(E_INTERNAL, "plan pitch %ld for m = %d", ldp, m);
    const int nstrip = ceil_div(m, 256);
    const int rows = apd_rows(n, m);
    const int nblk = ceil_div(n, rows);
    const dim3 grid(nstrip, nblk);
    const double k = 1.4426950408889634074 / reg;
    if (bulk == 0)
        hipLaunchKernelGGL(k_apd_plan<float>, grid, dim3(256), 0, s,
                           static_cast<const float*>(M), ldm, n, m, alpha, beta, rows, k, tau,
                           init ? 1 : 0, P, ldp, rowpart, colpart, scal);
    else
        hipLaunchKernelGGL(k_apd_plan<double>, grid, dim3(256), 0, s,
                           static_cast<const double*>(M), ldm, n, m, alpha, beta, rows, k, tau,
                           init ? 1 : 0, P, ldp, rowpart, colpart, scal);
    KERNEL_CHECK();
    const int npart = ceil_div(std::max(n, m), 256);
    hipLaunchKernelGGL(k_apd_marg, dim3(npart), dim3(256), 0, s, rowpart, nstrip, colpart, nblk,
                       a, b, n, m, nullptr, nullptr, part);
    KERNEL_CHECK();
    hipLaunchKernelGGL(k_apd_final, dim3(1), dim3(256), 0, s, part, npart, scal,
                       static_cast<long>(nstrip) * nblk, out);
    KERNEL_CHECK();
}

}  // namespace regot
