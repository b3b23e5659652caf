// objgrad_kernels.hip -- objective and gradient of the dual in one pass over M.
//
// f(gamma) = reg sum_ij T_ij - alpha'a - betat'bt and
// g = [T 1 - a ; (T' 1 - b)_{0..m-2}],  T_ij = 2^((alpha_i + beta_j - M_ij) k),
// k = log2(e) / reg, in FP64 throughout (src/sinkhorn_problem_obj.cpp:139-297,
// 324-393 form the same T with FP64 exp and sum it in FP64; no shift is
// applied there either, so overflow and underflow behave alike).  2^x is the
// table-driven FP64 exp2 of dexp2.hpp (about 0.5 ulp), several times cheaper
// than the library routine, which bounded this pass by FP64 issue rate.
//
// The row and column sums come out of the same read of M, so each T_ij is
// formed once (the FP64 exponential is the expensive part of this pass).
// Tiling: a workgroup owns a 256-column strip of a block of rows; each lane
// owns 4 adjacent columns and the 4 waves take the block's rows round robin
// (coalesced row segments).  Row sums are reduced
// across the wave per row and written per strip, column sums accumulate per
// lane down the block and are merged through LDS per block; two small
// kernels fold strips, blocks and the scalars in a fixed order.
#include <algorithm>

#include "dense_device.hpp"
#include "dexp2.hpp"
#include "problem.hpp"

namespace regot {
namespace {

// Wave sum by DPP and shuffle steps (the same bits on every lane).
__device__ __forceinline__ double wsum(double v) { return dev::group_sum<64>(v); }

__device__ __forceinline__ void ldm4(const float* p, int c0, int m, double (&v)[4])
{
    if (c0 + 3 < m) {
        typedef float f4 __attribute__((ext_vector_type(4)));
        const f4 u = __builtin_nontemporal_load(reinterpret_cast<const f4*>(p + c0));
        v[0] = u.x;
        v[1] = u.y;
        v[2] = u.z;
        v[3] = u.w;
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = (c0 + q < m) ? static_cast<double>(p[c0 + q]) : 0.0;
    }
}

__device__ __forceinline__ void ldm4(const double* p, int c0, int m, double (&v)[4])
{
    if (c0 + 3 < m) {
        typedef double d2 __attribute__((ext_vector_type(2)));
        const d2 u0 = __builtin_nontemporal_load(reinterpret_cast<const d2*>(p + c0));
        const d2 u1 = __builtin_nontemporal_load(reinterpret_cast<const d2*>(p + c0 + 2));
        v[0] = u0.x;
        v[1] = u0.y;
        v[2] = u1.x;
        v[3] = u1.y;
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = (c0 + q < m) ? p[c0 + q] : 0.0;
    }
}

// Columns per lane.  8 would spread each row's wave sum over twice the
// elements (27.5 instead of 30 VALU instructions per element) but needs 88
// VGPRs, 5 waves per SIMD instead of 8: measured slower (objective call at
// 16384^2 316 against 297 us, at 4096^2 61.5 against 51.0 us).
constexpr int kCPL = 4;
constexpr int kStrip = 64 * kCPL;     // columns per workgroup strip

// The rows of one tile: FULL = every column of the strip is inside the
// matrix (no per-element bounds), WT = weights x, y present.
template <typename TM, bool FULL, bool WT>
__device__ __forceinline__ void sums_rows_loop(const TM* __restrict__ M, long ldm, int n, int m,
                                               const double* __restrict__ alpha, double k64,
                                               const double* __restrict__ xw, int c0, int r, int r1,
                                               int strip, const double (&bk)[kCPL],
                                               const double (&yq)[kCPL], const double* tab,
                                               double (&cs)[kCPL], double* __restrict__ rowpart)
{
    const int lane = threadIdx.x & 63;
    for (; r < r1; r += 4) {
        double mv[kCPL];
#pragma unroll
        for (int h = 0; h < kCPL; h += 4) {
            double v4[4];
            if (FULL && sizeof(TM) == 4) {
                typedef float f4 __attribute__((ext_vector_type(4)));
                const f4 u = __builtin_nontemporal_load(reinterpret_cast<const f4*>(
                    reinterpret_cast<const float*>(M) + r * ldm + c0 + h));
                v4[0] = u.x;
                v4[1] = u.y;
                v4[2] = u.z;
                v4[3] = u.w;
            } else {
                ldm4(M + static_cast<long>(r) * ldm, c0 + h, m, v4);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) mv[h + q] = v4[q];
        }
        const double ak = alpha[r] * k64;
        double rs = 0.0;
#pragma unroll
        for (int q = 0; q < kCPL; q++) {
            if (FULL || c0 + q < m) {
                // 64 (alpha_i + beta_j - M_ij) / reg in log2 units
                double t = dev::dexp2_64(fma(-k64, mv[q], ak + bk[q]), tab);
                if (WT) t *= (xw ? xw[r] : 1.0) * yq[q];
                rs += t;
                cs[q] += t;
            }
        }
        rs = wsum(rs);
        if (lane == 0) rowpart[static_cast<long>(strip) * n + r] = rs;
    }
}

// grid (strips = ceil(m / kStrip), row blocks = ceil(n / rows)), 256 threads.
// beta has m entries (beta_{m-1} = 0).  Sums of w_ij = x_i T_ij y_j (WT:
// weights given, else 1): rowpart[strip * n + i] = sum of row i over the
// strip; colpart[block * m + j] (when non-null) = sum of column j over the
// block.  Each lane owns kCPL adjacent columns.  (Loading the next row's M
// ahead measured 5% slower: more registers, lower occupancy.)
template <typename TM, bool WT>
__global__ __launch_bounds__(256) void k_sums_tile(const TM* __restrict__ M, long ldm, int n, int m,
                                                   const double* __restrict__ alpha,
                                                   const double* __restrict__ beta, double k,
                                                   int rows, const double* __restrict__ xw,
                                                   const double* __restrict__ yw,
                                                   double* __restrict__ rowpart,
                                                   double* __restrict__ colpart)
{
    __shared__ double sm[4][kStrip];
    __shared__ __align__(16) double tab[128];
    dev::exp2_tab_load(tab);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int strip = blockIdx.x, blk = blockIdx.y;
    const int c0 = strip * kStrip + lane * kCPL;
    const int r0 = blk * rows;
    const int r1 = min(n, r0 + rows);
    // exponents in units of 1/64 (dexp2_64): exactly 64 times the log2 ones
    const double k64 = 64.0 * k;
    double bk[kCPL], yq[kCPL], cs[kCPL];
#pragma unroll
    for (int q = 0; q < kCPL; q++) {
        bk[q] = (c0 + q < m) ? beta[c0 + q] * k64 : 0.0;
        yq[q] = (WT && yw && c0 + q < m) ? yw[c0 + q] : 1.0;
        cs[q] = 0.0;
    }
    if ((strip + 1) * kStrip <= m)   // workgroup-uniform
        sums_rows_loop<TM, true, WT>(M, ldm, n, m, alpha, k64, xw, c0, r0 + wave, r1, strip, bk,
                                     yq, tab, cs, rowpart);
    else
        sums_rows_loop<TM, false, WT>(M, ldm, n, m, alpha, k64, xw, c0, r0 + wave, r1, strip,
                                      bk, yq, tab, cs, rowpart);
    if (!colpart) return;   // grid-uniform
#pragma unroll
    for (int q = 0; q < kCPL; q++) sm[wave][lane * kCPL + q] = cs[q];
    __syncthreads();
    for (int c = threadIdx.x; c < kStrip; c += 256) {
        const int j = strip * kStrip + c;
        if (j < m)
            colpart[static_cast<long>(blk) * m + j] =
                ((sm[0][c] + sm[1][c]) + sm[2][c]) + sm[3][c];
    }
}

// Four workgroup sums at once, each in bsum's order (wave sums, then the
// four wave totals left to right): one barrier instead of eight.
__device__ __forceinline__ void bsum4(double (&v)[4], double (*sm)[4])
{
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = wsum(v[q]);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < 4; q++) sm[q][threadIdx.x >> 6] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = ((sm[q][0] + sm[q][1]) + sm[q][2]) + sm[q][3];
}

// Thread t < n: r_t = sum over strips, grad_t = r_t - a_t; thread t < m - 1:
// c_t = sum over blocks (when colpart is set), grad_{n+t} = c_t - b_t.
// part[4 wg + {0..3}] = this workgroup's sum r, sum alpha a, sum betat bt,
// sum g^2.
__global__ __launch_bounds__(256) void k_sums_fold(const double* __restrict__ rowpart, int nstrip,
                                                   const double* __restrict__ colpart, int nblk,
                                                   const double* __restrict__ alpha,
                                                   const double* __restrict__ beta,
                                                   const double* __restrict__ a,
                                                   const double* __restrict__ b, int n, int m,
                                                   double* __restrict__ grad,
                                                   double* __restrict__ part)
{
    __shared__ double sm44[4][4];
    const int t = blockIdx.x * 256 + threadIdx.x;
    double sr = 0.0, sa = 0.0, sb = 0.0, g2 = 0.0;
    if (t < n) {
        // eight partials in flight, added in order
        double r = 0.0;
        int s = 0;
        for (; s + 8 <= nstrip; s += 8) {
            double x[8];
#pragma unroll
            for (int u = 0; u < 8; u++) x[u] = rowpart[static_cast<long>(s + u) * n + t];
#pragma unroll
            for (int u = 0; u < 8; u++) r += x[u];
        }
        for (; s < nstrip; s++) r += rowpart[static_cast<long>(s) * n + t];
        const double g = r - a[t];
        if (grad) grad[t] = g;
        sr = r;
        sa = alpha[t] * a[t];
        g2 = g * g;
    }
    if (t < m - 1) {
        sb = beta[t] * b[t];
        if (colpart) {
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
            const double g = c - b[t];
            if (grad) grad[n + t] = g;
            g2 += g * g;
        }
    }
    double v[4] = {sr, sa, sb, g2};
    bsum4(v, sm44);
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x] = v[0];
        part[4 * blockIdx.x + 1] = v[1];
        part[4 * blockIdx.x + 2] = v[2];
        part[4 * blockIdx.x + 3] = v[3];
    }
}

// out[0] = reg sum r - alpha'a - betat'bt, out[1] = ||g||^2.
__global__ __launch_bounds__(256) void k_sums_final(const double* __restrict__ part, int npart,
                                                    double reg, double* __restrict__ out)
{
    __shared__ double sm44[4][4];
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < npart; i += 256) {
        v[0] += part[4 * i];
        v[1] += part[4 * i + 1];
        v[2] += part[4 * i + 2];
        v[3] += part[4 * i + 3];
    }
    bsum4(v, sm44);
    if (threadIdx.x == 0) {
        out[0] = reg * v[0] - v[1] - v[2];
        out[1] = v[3];
    }
}

__global__ __launch_bounds__(256) void k_exp2_check(const double* __restrict__ x,
                                                    double* __restrict__ y, long count)
{
    __shared__ __align__(16) double tab[128];
    dev::exp2_tab_load(tab);
    __syncthreads();
    for (long i = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; i < count;
         i += static_cast<long>(gridDim.x) * 256)
        y[i] = dev::dexp2(x[i], tab);
}

}  // namespace

void launch_exp2_check(const double* x, double* y, long count, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_exp2_check, dim3(std::min(4096L, (count + 255) / 256)), dim3(256), 0, s,
                       x, y, count);
    KERNEL_CHECK();
}

// Rows per workgroup of the one-pass kernels: >= 1024 workgroups where the
// problem allows, 32 to 256 rows (the column partials stay <= n/32 rows;
// at 16384^2, 128- and 256-row blocks measured 7% faster than 64).
int sums_rows(int n, int m)
{
    const long nstrip = ceil_div(m, kStrip);
    long r = static_cast<long>(n) * nstrip / 1024;
    r = std::max(32L, std::min(256L, r));
    return static_cast<int>((r + 3) / 4 * 4);
}

void sums_scratch_sizes(int n, int m, long* rowpart, long* colpart, long* part)
{
    const long nstrip = ceil_div(m, kStrip), nblk = ceil_div(n, sums_rows(n, m));
    *rowpart = nstrip * n;
    *colpart = nblk * m;
    *part = 4L * ceil_div(std::max(n, m), 256);
}

void launch_objgrad_fused(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                          const double* beta, const double* a, const double* b, double reg,
                          bool with_grad, double* grad, double* rowpart, double* colpart,
                          double* part, double* out, hipStream_t s)
{
    if (n <= 0 || m <= 0) fail(E_INVALID, "objective pass on an empty problem");
    const int nstrip = ceil_div(m, kStrip);
    const int rows = sums_rows(n, m);
    const int nblk = ceil_div(n, rows);
    const dim3 grid(nstrip, nblk);
    const double k = 1.4426950408889634074 / reg;
    double* cp = with_grad ? colpart : nullptr;
    if (bulk == 0)
        hipLaunchKernelGGL((k_sums_tile<float, false>), grid, dim3(256), 0, s,
                           static_cast<const float*>(M), ldm, n, m, alpha, beta, k, rows, nullptr,
                           nullptr, rowpart, cp);
    else
        hipLaunchKernelGGL((k_sums_tile<double, false>), grid, dim3(256), 0, s,
                           static_cast<const double*>(M), ldm, n, m, alpha, beta, k, rows,
                           nullptr, nullptr, rowpart, cp);
    KERNEL_CHECK();
    const int nf = ceil_div(std::max(n, m), 256);
    hipLaunchKernelGGL(k_sums_fold, dim3(nf), dim3(256), 0, s, rowpart, nstrip, cp, nblk, alpha,
                       beta, a, b, n, m, with_grad ? grad : nullptr, part);
    KERNEL_CHECK();
    hipLaunchKernelGGL(k_sums_final, dim3(1), dim3(256), 0, s, part, nf, reg, out);
    KERNEL_CHECK();
}

void launch_wsums(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                  const double* beta, double reg, const double* x, const double* y,
                  double* rowpart, double* colpart, hipStream_t s)
{
    if (n <= 0 || m <= 0) fail(E_INVALID, "weighted sums on an empty problem");
    const int rows = sums_rows(n, m);
    const dim3 grid(ceil_div(m, kStrip), ceil_div(n, rows));
    const double k = 1.4426950408889634074 / reg;
    const bool wt = x || y;
    if (bulk == 0 && wt)
        hipLaunchKernelGGL((k_sums_tile<float, true>), grid, dim3(256), 0, s,
                           static_cast<const float*>(M), ldm, n, m, alpha, beta, k, rows, x, y,
                           rowpart, colpart);
    else if (bulk == 0)
        hipLaunchKernelGGL((k_sums_tile<float, false>), grid, dim3(256), 0, s,
                           static_cast<const float*>(M), ldm, n, m, alpha, beta, k, rows, x, y,
                           rowpart, colpart);
    else if (wt)
        hipLaunchKernelGGL((k_sums_tile<double, true>), grid, dim3(256), 0, s,
                           static_cast<const double*>(M), ldm, n, m, alpha, beta, k, rows, x, y,
                           rowpart, colpart);
    else
        hipLaunchKernelGGL((k_sums_tile<double, false>), grid, dim3(256), 0, s,
                           static_cast<const double*>(M), ldm, n, m, alpha, beta, k, rows, x, y,
                           rowpart, colpart);
    KERNEL_CHECK();
}

int sums_strip_width() { return kStrip; }

}  // namespace regot
