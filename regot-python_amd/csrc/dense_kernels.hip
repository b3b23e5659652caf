// dense_kernels.hip -- single-problem streaming passes over the cost matrix M
// (row pass, column pass, FP64 plan) and the on-device BCD bookkeeping.  The
// sweeps themselves are in dense_device.hpp.
#include "dense.hpp"
#include "dense_device.hpp"

#include <cmath>

namespace regot {
namespace {

using namespace dev;

// ---------------------------------------------------------------------------
// Row pass
// ---------------------------------------------------------------------------

// Each wave owns R consecutive rows (row_lse; R = 1 for matrices too small to
// fill the chip with 4-row waves).  The epilogue writes, per row, the log2 sum,
// alpha = reg (log a - ln2 L) and/or the row sum r_i = 2^(k alpha_i + L).
// With FIN, per-workgroup sums of (g^2, r, alpha a) are published for the
// combine kernel of the same BCD iteration.  nt: nontemporal loads of M.
template <typename TM, typename TA, int R, bool FIN>
__global__ __launch_bounds__(256) void k_row_pass(const TM* __restrict__ M, long ldm, int n,
                                                  int m, const double* __restrict__ beta,
                                                  double k, RowEpi epi, BcdState fin,
                                                  const int* __restrict__ ctl, int nt)
{
    if (ctl != nullptr && ctl[0] != 0) return;
    const int lane = threadIdx.x & 63;
    // wave-uniform: row tests below become scalar branches, not exec masks
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long row0 = (static_cast<long>(blockIdx.x) * 4 + wave) * R;

    const double* rowpot = epi.rowpot;
    double* aout = epi.alpha_out;
    double wg2 = 0.0, wr = 0.0, waa = 0.0;

    if (row0 < n) {
        double L[R];
        row_lse<TM, TA, R>(M, ldm, n, 0, m, beta, k, epi.loga, rowpot, row0, lane, L, nt != 0);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const long i = row0 + r;
            if (i >= n) break;
            if (lane == 0) {
                if (epi.lse2) epi.lse2[i] = L[r];
                if (aout) aout[i] = epi.reg * (epi.loga[i] - kLn2 * L[r]);
                if (epi.rowsum || FIN) {
                    const double ri = exp2(rowpot[i] * k + L[r]);
                    if (epi.rowsum) epi.rowsum[i] = ri;
                    if (FIN) {
                        const double g = ri - fin.a[i];
                        wg2 += g * g;
                        wr += ri;
                        waa += rowpot[i] * fin.a[i];
                    }
                }
            }
        }
    }

    if (FIN) {
        // Per-workgroup partial sums; the combine kernel of this iteration
        // reduces them after the kernel boundary (no cross-XCD fences here).
        __shared__ double red[4][3];
        if (lane == 0) {
            red[wave][0] = wg2;
            red[wave][1] = wr;
            red[wave][2] = waa;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
#pragma unroll
            for (int q = 0; q < 3; q++)
                fin.rpart[3 * blockIdx.x + q] = ((red[0][q] + red[1][q]) + red[2][q]) + red[3][q];
            // beta_k[m-1] for the objective: the combine kernel of this
            // iteration overwrites beta while other workgroups still decide
            if (blockIdx.x == 0) fin.bm[0] = beta[m - 1];
        }
    }
}

// ---------------------------------------------------------------------------
// Column pass
// ---------------------------------------------------------------------------

// Workgroup (x, y) = column tile x (64V columns) over row block y (rpb rows):
// one FP64 log2 partial sum per (row block, column) into part[y * m + j].
template <typename TM, typename TA>
__global__ __launch_bounds__(256) void k_col_partial(
    const TM* __restrict__ M, long ldm, int n, int m, const double* __restrict__ alpha,
    const double* __restrict__ colpot, const double* __restrict__ logb, double k, int rpb,
    double* __restrict__ part, const int* __restrict__ ctl, int nt)
{
    constexpr int TW = 64 * VecOf<TM>::N;
    if (ctl != nullptr && ctl[0] != 0) return;
    const long rb = static_cast<long>(blockIdx.y) * rpb;
    const long re = (rb + rpb < n) ? rb + rpb : n;
    col_partial_tile<TM, TA>(M, ldm, m, alpha, colpot, logb, k, blockIdx.x * TW, rb, re,
                             part + static_cast<long>(blockIdx.y) * m, nt != 0);
}

constexpr int kCombineCols = 64;   // columns per combine workgroup (4 threads each)

// Stop test of the on-device BCD loop for iterate k = st.k, computed
// identically (same data, same order) by every combine workgroup:
// g_k = [r - a ; column residual of the update that produced beta_k],
// f_k = reg sum r - (alpha_k + beta_{k,m-1})'a - (beta_k - beta_{k,m-1})'b over
// j < m-1.  The test order is that of src/sinkhorn_bcd.cpp:65-79.
__device__ void bcd_decide(const BcdState& st, int kk, double* f_out, double* g_out,
                           bool* done_out, bool* stall_out)
{
    const double* cp = st.cpart + 2L * st.ntiles * (kk & 1);
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < st.nrb; q += 256) {
        v[0] += st.rpart[3 * q];
        v[1] += st.rpart[3 * q + 1];
        v[2] += st.rpart[3 * q + 2];
    }
    for (int q = threadIdx.x; q < st.ntiles; q += 256) {
        v[3] += cp[2 * q];
        v[4] += cp[2 * q + 1];
    }
    block_sums_256<5>(v);
    const double bm = st.bm[0];
    const double gnorm = sqrt(v[0] + v[3]);
    const double da = v[2] + bm * st.sum_a;
    const double db = v[4] - bm * st.sum_b1;
    *f_out = st.reg * v[1] - da - db;
    *g_out = gnorm;
    *done_out = (gnorm < st.tol) || (kk >= st.max_iter);
    *stall_out = false;
    if (!*done_out && st.stall_window > 0) {
        const double* sv = st.stall + 2 * (kk & 1);
        *stall_out = !(gnorm < sv[0]) && (kk - static_cast<int>(sv[1]) >= st.stall_window);
    }
}

// Combine the nblk partials of each column (4 threads per column, fixed
// order), then the column epilogue.  With BCD, first take the stop decision
// for iterate st.k; when it stops, beta is left untouched (beta_k is the
// result), otherwise beta <- beta_{k+1} and the tile sums of the new column
// residual go to the other parity half of cpart.  Workgroup 0 records the
// history entry and advances or stops the loop.
// SUMS: the partials are plain FP64 column sums of exp((alpha_i + colpot_j -
// M_ij) / reg) (the single-read BCD pass), so log2 sum_i 2^((alpha_i - M_ij) k)
// = log2(sum) - colpot_j k.
template <bool BCD, bool SUMS>
__global__ __launch_bounds__(256) void k_col_combine(const double* __restrict__ part, int nblk,
                                                     int m, double k, ColEpi epi, BcdState st,
                                                     const int* __restrict__ ctl)
{
    int kk = 0;
    double f = 0.0, gnorm = 0.0;
    bool done = false, stall = false;
    if (BCD) {
        if (st.ctl[0] != 0) return;
        kk = st.k;
        bcd_decide(st, kk, &f, &gnorm, &done, &stall);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            const int h = kk % st.hist_cap;
            st.hist_f[h] = f;
            st.hist_g[h] = gnorm;
            st.hist_ts[h] = static_cast<long long>(wall_clock64());
            if (st.stall_window > 0) {
                // running minimum for the next iteration (other parity slot)
                const double* sv = st.stall + 2 * (kk & 1);
                double* nx = st.stall + 2 * ((kk + 1) & 1);
                const bool better = gnorm < sv[0];
                nx[0] = better ? gnorm : sv[0];
                nx[1] = better ? static_cast<double>(kk) : sv[1];
            }
        }
        done = done || stall;
    } else if (ctl != nullptr && ctl[0] != 0) {
        return;
    }

    const int sub = threadIdx.x & 3;
    const int j = blockIdx.x * kCombineCols + (threadIdx.x >> 2);
    double r2 = 0.0, bb = 0.0;
    if (!done) {
        double L;
        if (SUMS) {
            // 32 loads in flight per thread, then added in order p = sub, sub + 4, ...
            double cs = 0.0;
            if (j < m) {
                int p = sub;
                for (; p + 124 < nblk; p += 128) {
                    double x[32];
#pragma unroll
                    for (int u = 0; u < 32; u++) x[u] = part[static_cast<long>(p + 4 * u) * m + j];
#pragma unroll
                    for (int u = 0; u < 32; u++) cs += x[u];
                }
                for (; p < nblk; p += 4) cs += part[static_cast<long>(p) * m + j];
            }
            cs += __shfl_xor(cs, 1);
            cs += __shfl_xor(cs, 2);
            // colpot_j is read before sub 0 of the same column writes beta_out_j
            L = (j < m && cs > 0.0) ? log2(cs) - epi.colpot[j] * k : -INFINITY;
        } else {
            L = combine_partials(part, nblk, m, j, j < m, sub);
        }
        if (sub == 0 && j < m) {
            if (epi.lse2) epi.lse2[j] = L;
            double cs = 0.0;
            if (epi.colsum || (epi.resid && !epi.beta_out)) {
                cs = exp2(epi.colpot[j] * k + L);
                if (epi.colsum) epi.colsum[j] = cs;
            }
            double bj = epi.colpot ? epi.colpot[j] : 0.0;
            double g;
            if (epi.beta_out) {
                bj = epi.reg * (epi.logb[j] - kLn2 * L);
                epi.beta_out[j] = bj;
                g = epi.b[j] * expm1(bj / epi.reg + kLn2 * L - epi.logb[j]);
            } else {
                g = cs - epi.b[j];
            }
            if (epi.resid) epi.resid[j] = g;
            if (j < m - 1) {
                r2 = g * g;
                bb = bj * epi.b[j];
            }
        }
        if (epi.tilepart) {
            double v[2] = {r2, bb};
            block_sums_256<2>(v);
            r2 = v[0];
            bb = v[1];
            if (threadIdx.x == 0) {
                double* tp = epi.tilepart + (BCD ? 2L * gridDim.x * ((kk + 1) & 1) : 0);
                tp[2 * blockIdx.x] = r2;
                tp[2 * blockIdx.x + 1] = bb;
            }
        }
    }
    if (BCD && blockIdx.x == 0 && threadIdx.x == 0) {
        if (done) {
            st.ctl[2] = kk;
            st.ctl[0] = stall ? 2 : 1;
        } else {
            st.ctl[1] = kk + 1;
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 plan, BCD bookkeeping and small helpers
// ---------------------------------------------------------------------------

template <typename TM>
__global__ __launch_bounds__(256) void k_plan(const TM* __restrict__ M, long ldm, int m,
                                              const double* __restrict__ alpha,
                                              const double* __restrict__ beta, double reg,
                                              double* __restrict__ plan, long r0)
{
    const long i = r0 + blockIdx.y;
    const double ai = alpha[i];
    const TM* Mi = M + i * ldm;
    double* Pi = plan + static_cast<long>(blockIdx.y) * m;
    for (int j = blockIdx.x * 256 + threadIdx.x; j < m; j += gridDim.x * 256)
        Pi[j] = exp(((ai + beta[j]) - static_cast<double>(Mi[j])) / reg);
}

__global__ void k_fill(double* p, double v, long count)
{
    for (long i = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; i < count;
         i += static_cast<long>(gridDim.x) * 256)
        p[i] = v;
}

__global__ void k_fill_int(int* p, int v, long count)
{
    for (long i = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; i < count;
         i += static_cast<long>(gridDim.x) * 256)
        p[i] = v;
}

__global__ void k_stamp(long long* dst)
{
    if (threadIdx.x == 0) dst[0] = static_cast<long long>(wall_clock64());
}

__global__ void k_split_gamma(const double* gamma, int n, int m, double* alpha, double* beta)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) alpha[t] = gamma[t];
    if (t < m) beta[t] = (t < m - 1) ? gamma[n + t] : 0.0;
}

int grid_for(long count)
{
    long g = (count + 255) / 256;
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    return static_cast<int>(g);
}

}  // namespace

// Rows per wave of the row pass.
template <typename TM>
constexpr int kRowsPerWave = sizeof(TM) == 4 ? 4 : 2;

// R-row waves while they give >= 1024 workgroups (4 per CU), else 1-row waves.
constexpr int kWideBelow = 1024;

int rows_per_wave(int bulk, int n)
{
    const int R = bulk == 0 ? kRowsPerWave<float> : kRowsPerWave<double>;
    return ceil_div(n, 4 * R) >= kWideBelow ? R : 1;
}

int row_pass_blocks(int bulk, int n) { return ceil_div(n, 4 * rows_per_wave(bulk, n)); }

// Nontemporal loads above the size where M could stay in the 256 MB
// Infinity Cache between passes (measured: cached loads are 10-15% faster
// at 4096^2 and 8192^2 FP32, nontemporal ones 8% faster at 16384^2).
bool stream_nt(int bulk, int n, long ldm)
{
    return static_cast<long>(n) * ldm * (bulk == 0 ? 4 : 8) > (256L << 20);
}

template <typename TM, typename TA, bool FIN>
void row_pass_launch(const void* M, long ldm, int n, int m, const double* beta, double k,
                     const RowEpi& epi, const BcdState& st, const int* ctl, hipStream_t s,
                     int bulk)
{
    constexpr int R = kRowsPerWave<TM>;
    const dim3 grid(row_pass_blocks(bulk, n));
    const int nt = stream_nt(bulk, n, ldm) ? 1 : 0;
    if (rows_per_wave(bulk, n) == R)
        hipLaunchKernelGGL((k_row_pass<TM, TA, R, FIN>), grid, dim3(256), 0, s,
                           static_cast<const TM*>(M), ldm, n, m, beta, k, epi, st, ctl, nt);
    else
        hipLaunchKernelGGL((k_row_pass<TM, TA, 1, FIN>), grid, dim3(256), 0, s,
                           static_cast<const TM*>(M), ldm, n, m, beta, k, epi, st, ctl, nt);
    KERNEL_CHECK();
}

void launch_row_pass(int bulk, const void* M, long ldm, int n, int m, const double* beta,
                     double k, const RowEpi& epi, const int* ctl, hipStream_t s, bool precise)
{
    if (n <= 0) return;
    BcdState none{};
    if (bulk == 0 && !precise)
        row_pass_launch<float, float, false>(M, ldm, n, m, beta, k, epi, none, ctl, s, bulk);
    else if (bulk == 0)
        row_pass_launch<float, double, false>(M, ldm, n, m, beta, k, epi, none, ctl, s, bulk);
    else
        row_pass_launch<double, double, false>(M, ldm, n, m, beta, k, epi, none, ctl, s, bulk);
}

void launch_row_pass_bcd(int bulk, const void* M, long ldm, int n, int m, const double* beta,
                         double k, const RowEpi& epi, const BcdState& st, hipStream_t s,
                         bool precise)
{
    if (bulk == 0 && !precise)
        row_pass_launch<float, float, true>(M, ldm, n, m, beta, k, epi, st, st.ctl, s, bulk);
    else if (bulk == 0)
        row_pass_launch<float, double, true>(M, ldm, n, m, beta, k, epi, st, st.ctl, s, bulk);
    else
        row_pass_launch<double, double, true>(M, ldm, n, m, beta, k, epi, st, st.ctl, s, bulk);
}

int col_pass_blocks(int n, int m)
{
    // Aim for >= 512 workgroups (2 per CU on 256 CUs) with >= 16 rows each:
    // long column streams per workgroup measured fastest (16384^2 FP32:
    // 163 us at 512 workgroups, 166 us at 1024, 175 us at 2048; 4096^2: 35 us
    // at 512, 40 us at 2048).
    const int tiles = ceil_div(m, 256);
    int nblk = ceil_div(512, tiles);
    const int maxblk = ceil_div(n, 16);
    if (nblk > maxblk) nblk = maxblk;
    if (nblk < 1) nblk = 1;
    return nblk;
}

int col_combine_blocks(int m) { return ceil_div(m, kCombineCols); }

void launch_col_pass(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                     const double* colpot_guess, const double* logb, double k, double* part,
                     int nblk, const ColEpi& epi, const BcdState* st, const int* ctl,
                     hipStream_t s, bool precise)
{
    if (m <= 0 || n <= 0) return;
    const int rpb = ceil_div(n, nblk);
    const dim3 grid(ceil_div(m, bulk == 0 ? 256 : 128), nblk);
    // the column sweep measured fastest with nontemporal loads at every size
    // (4096^2: 24.5 us against 26.9 us cached)
    const int nt = 1;
    if (bulk == 0 && !precise)
        hipLaunchKernelGGL((k_col_partial<float, float>), grid, dim3(256), 0, s,
                           static_cast<const float*>(M), ldm, n, m, alpha, colpot_guess, logb,
                           k, rpb, part, ctl, nt);
    else if (bulk == 0)
        hipLaunchKernelGGL((k_col_partial<float, double>), grid, dim3(256), 0, s,
                           static_cast<const float*>(M), ldm, n, m, alpha, colpot_guess, logb,
                           k, rpb, part, ctl, nt);
    else
        hipLaunchKernelGGL((k_col_partial<double, double>), grid, dim3(256), 0, s,
                           static_cast<const double*>(M), ldm, n, m, alpha, colpot_guess, logb,
                           k, rpb, part, ctl, nt);
    KERNEL_CHECK();
    if (st != nullptr)
        hipLaunchKernelGGL((k_col_combine<true, false>), dim3(col_combine_blocks(m)), dim3(256),
                           0, s, part, nblk, m, k, epi, *st, ctl);
    else
        hipLaunchKernelGGL((k_col_combine<false, false>), dim3(col_combine_blocks(m)),
                           dim3(256), 0, s, part, nblk, m, k, epi, BcdState{}, ctl);
    KERNEL_CHECK();
}

void launch_col_combine_sums(const double* part, int nblk, int m, double k, const ColEpi& epi,
                             const BcdState& st, hipStream_t s)
{
    if (epi.colpot == nullptr) fail(E_INTERNAL, "column-sum combine needs the column potential");
    hipLaunchKernelGGL((k_col_combine<true, true>), dim3(col_combine_blocks(m)), dim3(256), 0, s,
                       part, nblk, m, k, epi, st, st.ctl);
    KERNEL_CHECK();
}

void launch_plan(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                 const double* beta, double reg, double* plan, long r0, int nr, hipStream_t s)
{
    (void)n;
    if (nr <= 0 || m <= 0) return;
    const int gx = ceil_div(m, 256) < 8 ? ceil_div(m, 256) : 8;
    if (bulk == 0)
        hipLaunchKernelGGL(k_plan<float>, dim3(gx, nr), dim3(256), 0, s,
                           static_cast<const float*>(M), ldm, m, alpha, beta, reg, plan, r0);
    else
        hipLaunchKernelGGL(k_plan<double>, dim3(gx, nr), dim3(256), 0, s,
                           static_cast<const double*>(M), ldm, m, alpha, beta, reg, plan, r0);
    KERNEL_CHECK();
}

void launch_fill(double* p, double v, long count, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fill, dim3(grid_for(count)), dim3(256), 0, s, p, v, count);
    KERNEL_CHECK();
}

void launch_fill_int(int* p, int v, long count, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fill_int, dim3(grid_for(count)), dim3(256), 0, s, p, v, count);
    KERNEL_CHECK();
}

void launch_stamp(long long* dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s, dst);
    KERNEL_CHECK();
}

void launch_split_gamma(const double* gamma, int n, int m, double* alpha, double* beta,
                        hipStream_t s)
{
    const int c = n > m ? n : m;
    hipLaunchKernelGGL(k_split_gamma, dim3(ceil_div(c, 256)), dim3(256), 0, s, gamma, n, m,
                       alpha, beta);
    KERNEL_CHECK();
}

}  // namespace regot
