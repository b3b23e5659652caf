// sparse_kernels.hip -- sparsification of T and sparse Hessian products for gfx950.
//
// Selection is radix-based on the IEEE bit pattern of T (non-negative values
// order like their bit patterns): 11-bit digits from the top, 3 passes for
// FP32 keys, 6 for FP64 keys.  Mass-weighted selection (threshold rule)
// histograms FP64 masses in LDS; count selection (top-k) histograms counts.
// Compaction is count -> exclusive scan -> ordered write; the write position
// of each kept entry comes from wave ballots, so no global atomics decide
// placement and the output is identical run to run.
#include <cmath>

#include "sparse.hpp"

namespace regot {
namespace {

constexpr int RD = 11;
constexpr int RB = 1 << RD;   // buckets per radix pass

template <typename TM>
struct Elem;

// FP32 bulk: log2 T formed in FP64, exp2 in FP32 (same arithmetic as the
// streaming passes).  "Zero in FP64" is exp underflow: log2 T < -1075.
template <>
struct Elem<float> {
    using K = unsigned int;
    static constexpr int KB = 32;
    __device__ static float val(double ab, float mij, double k, double reg, bool& z64)
    {
        (void)reg;
        const double x = fma(-k, static_cast<double>(mij), ab * k);
        z64 = x < -1075.0;
        return exp2f(static_cast<float>(x));
    }
    __device__ static K key(float t) { return __float_as_uint(t); }
    __device__ static double from_key(K kk) { return static_cast<double>(__uint_as_float(kk)); }
};

// FP64 bulk: T = exp((alpha_i + beta_j - M_ij) / reg), as the reference forms it.
template <>
struct Elem<double> {
    using K = unsigned long long;
    static constexpr int KB = 64;
    __device__ static double val(double ab, double mij, double k, double reg, bool& z64)
    {
        (void)k;
        const double t = exp((ab - mij) / reg);
        z64 = (t == 0.0);
        return t;
    }
    __device__ static K key(double t) { return static_cast<K>(__double_as_longlong(t)); }
    __device__ static double from_key(K kk) { return __longlong_as_double(static_cast<long long>(kk)); }
};

__device__ __forceinline__ int digit(int KB, int p, int* width)
{
    const int hi = KB - p * RD;
    const int w = hi < RD ? hi : RD;
    *width = w;
    return hi - w;
}

// ---------------------------------------------------------------------------
// Workgroup (256 threads) helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wave_incl_scan(double v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double u = __shfl_up(v, o);
        if (lane >= o) v += u;
    }
    return v;
}

// Exclusive scan of one double per thread over the workgroup; *total gets the sum.
__device__ double block_excl_scan(double v, double* total)
{
    __shared__ double wt[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double inc = wave_incl_scan(v);
    if (lane == 63) wt[wave] = inc;
    __syncthreads();
    double off = 0.0;
    for (int w = 0; w < wave; w++) off += wt[w];
    *total = ((wt[0] + wt[1]) + wt[2]) + wt[3];
    __syncthreads();
    return off + inc - v;
}

__device__ double block_sum(double v)
{
    double t;
    (void)block_excl_scan(v, &t);
    return t;
}

// Sum of an int flag/count over the workgroup (exact).
__device__ long long block_sum_ll(long long v)
{
    __shared__ long long wt[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) wt[wave] = v;
    __syncthreads();
    return wt[0] + wt[1] + wt[2] + wt[3];
}

// Exclusive prefix of a 0/1 flag over the workgroup in thread order, plus the total.
__device__ int block_flag_prefix(bool flag, int* total)
{
    __shared__ int wt[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long mask = __ballot(flag);
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int in_wave = __popcll(mask & lt);
    __syncthreads();
    if (lane == 0) wt[wave] = __popcll(mask);
    __syncthreads();
    int off = 0;
    for (int w = 0; w < wave; w++) off += wt[w];
    *total = wt[0] + wt[1] + wt[2] + wt[3];
    return off + in_wave;
}

// Bucket search over hist[0..nb) (FP64 masses in LDS), thread t owning 8 buckets.
// ASC: smallest b with base + sum(hist[0..b]) > thr; *before = base + sum(hist[0..b-1]).
// !ASC: largest b with base + sum(hist[b..nb-1]) >= thr; *before = base + sum(hist[b+1..]).
// Returns -1 when no bucket qualifies.  Ties in the claim go to the first bucket
// in scan order (integer LDS atomics), so the result is well defined.
template <bool ASC>
__device__ int select_bucket(const double* hist, int nb, double base, double thr, double* before)
{
    __shared__ int claim;
    __shared__ double claim_before[RB / 8];
    const int t = threadIdx.x;
    if (t == 0) claim = ASC ? 0x7fffffff : -1;
    // thread t owns buckets [8u, 8u+8) with u = t (ASC) or u = 255 - t (descending order)
    const int u = ASC ? t : (255 - t);
    double own = 0.0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int b = 8 * u + q;
        if (b < nb) own += hist[b];
    }
    double total;
    const double pre = block_excl_scan(own, &total);   // __syncthreads inside
    double cum = base + pre;
    int found = ASC ? 0x7fffffff : -1;
    double fb = 0.0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int b = ASC ? 8 * u + q : 8 * u + 7 - q;
        if (b >= nb) continue;
        const double nxt = cum + hist[b];
        const bool hit = ASC ? (nxt > thr) : (nxt >= thr);
        if (hit && found == (ASC ? 0x7fffffff : -1)) {
            found = b;
            fb = cum;
        }
        cum = nxt;
    }
    if (ASC ? (found != 0x7fffffff) : (found >= 0)) {
        claim_before[u] = fb;
        if (ASC)
            atomicMin(&claim, found);
        else
            atomicMax(&claim, found);
    }
    __syncthreads();
    const int b = claim;
    int res = -1;
    if (ASC ? (b != 0x7fffffff) : (b >= 0)) {
        res = b;
        *before = claim_before[b / 8];
    }
    __syncthreads();
    return res;
}

// ---------------------------------------------------------------------------
// Transpose
// ---------------------------------------------------------------------------

template <typename TM>
__global__ __launch_bounds__(256) void k_transpose(const TM* __restrict__ M, long ldm, int n,
                                                   int m, TM* __restrict__ MT, long ldt)
{
    __shared__ TM tile[64][65];
    const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        const int i = i0 + r, j = j0 + tx;
        tile[r][tx] = (i < n && j < m) ? M[static_cast<long>(i) * ldm + j] : TM(0);
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const int j = j0 + r, i = i0 + tx;
        if (j < m && i < n) MT[static_cast<long>(j) * ldt + i] = tile[tx][r];
    }
}

// ---------------------------------------------------------------------------
// Threshold rule
// ---------------------------------------------------------------------------

// tau_j = smallest value v of column j with sum_{T_ij <= v} T_ij > delta
// (the entry where the ascending cumulative mass first passes delta).
template <typename TM>
__global__ __launch_bounds__(256) void k_col_threshold(const TM* __restrict__ MT, long ldt,
                                                       ElemCtx c, double delta, int find_small,
                                                       double* __restrict__ tau)
{
    using E = Elem<TM>;
    using K = typename E::K;
    __shared__ double hist[RB];
    __shared__ K kmax_s[4];
    const int j = blockIdx.x;
    const TM* col = MT + static_cast<long>(j) * ldt;
    const double bj = c.beta[j];
    const int npass = (E::KB + RD - 1) / RD;
    K prefix = 0;
    double base = 0.0;
    bool none = false;
    K kmax = 0;
    for (int p = 0; p < npass; p++) {
        int width;
        const int shift = digit(E::KB, p, &width);
        for (int q = threadIdx.x; q < RB; q += 256) hist[q] = 0.0;
        __syncthreads();
        for (int i = threadIdx.x; i < c.n; i += 256) {
            bool z;
            const TM t = E::val(c.alpha[i] + bj, col[i], c.k, c.reg, z);
            const K key = E::key(t);
            if (p == 0) kmax = key > kmax ? key : kmax;
            if (p == 0 || (key >> (shift + width)) == prefix)
                atomicAdd(&hist[(key >> shift) & ((K(1) << width) - 1)], static_cast<double>(t));
        }
        __syncthreads();
        double before = 0.0;
        const int b = select_bucket<true>(hist, 1 << width, base, delta, &before);
        if (b < 0) {
            none = true;
            break;
        }
        base = before;
        prefix = (prefix << width) | static_cast<K>(b);
    }
    if (none) {
        // wave max, then across the 4 waves
        K v = kmax;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const K w = static_cast<K>(__shfl_xor(static_cast<unsigned long long>(v), o));
            v = w > v ? w : v;
        }
        if ((threadIdx.x & 63) == 0) kmax_s[threadIdx.x >> 6] = v;
        __syncthreads();
        K mx = kmax_s[0];
        for (int w = 1; w < 4; w++) mx = kmax_s[w] > mx ? kmax_s[w] : mx;
        if (threadIdx.x == 0) tau[j] = find_small ? INFINITY : E::from_key(mx);
    } else if (threadIdx.x == 0) {
        tau[j] = E::from_key(prefix);
    }
}

// Per row: dropped mass D_i (entries below their column threshold); if
// D_i > 2 delta, rho_i = largest value v of the dropped entries with
// sum_{d_ij >= v} d_ij >= D_i - 2 delta (else +inf); then the stored and
// FP64-zero kept counts under kept <=> T >= min(tau_j, rho_i) or FP64 zero.
template <typename TM>
__global__ __launch_bounds__(256) void k_row_threshold(const TM* __restrict__ M, long ldm,
                                                       ElemCtx c, const double* __restrict__ tau,
                                                       double delta2, double* __restrict__ rho,
                                                       int* __restrict__ cnt,
                                                       int* __restrict__ zcnt)
{
    using E = Elem<TM>;
    using K = typename E::K;
    __shared__ double hist[RB];
    const int i = blockIdx.x;
    const TM* row = M + static_cast<long>(i) * ldm;
    const double ai = c.alpha[i];
    const int m1 = c.m - 1;
    double dsum = 0.0;
    for (int j = threadIdx.x; j < m1; j += 256) {
        bool z;
        const TM t = E::val(ai + c.beta[j], row[j], c.k, c.reg, z);
        if (static_cast<double>(t) < tau[j]) dsum += static_cast<double>(t);
    }
    const double D = block_sum(dsum);
    double r_i = INFINITY;
    if (D > delta2) {
        const double target = D - delta2;
        const int npass = (E::KB + RD - 1) / RD;
        K prefix = 0;
        double above = 0.0;
        for (int p = 0; p < npass; p++) {
            int width;
            const int shift = digit(E::KB, p, &width);
            for (int q = threadIdx.x; q < RB; q += 256) hist[q] = 0.0;
            __syncthreads();
            for (int j = threadIdx.x; j < m1; j += 256) {
                bool z;
                const TM t = E::val(ai + c.beta[j], row[j], c.k, c.reg, z);
                const TM d = (static_cast<double>(t) < tau[j]) ? t : TM(0);
                const K key = E::key(d);
                if (p == 0 || (key >> (shift + width)) == prefix)
                    atomicAdd(&hist[(key >> shift) & ((K(1) << width) - 1)],
                              static_cast<double>(d));
            }
            __syncthreads();
            double before = 0.0;
            const int b = select_bucket<false>(hist, 1 << width, above, target, &before);
            if (b < 0) break;   // cannot happen: the total dropped mass exceeds target
            above = before;
            prefix = (prefix << width) | static_cast<K>(b);
        }
        r_i = E::from_key(prefix);
    }
    long long ns = 0, nz = 0;
    for (int j = threadIdx.x; j < m1; j += 256) {
        bool z;
        const TM t = E::val(ai + c.beta[j], row[j], c.k, c.reg, z);
        const double th = fmin(tau[j], r_i);
        const bool kept = (static_cast<double>(t) >= th) || z;
        ns += (kept && t != TM(0)) ? 1 : 0;
        nz += (kept && z) ? 1 : 0;
    }
    ns = block_sum_ll(ns);
    nz = block_sum_ll(nz);
    if (threadIdx.x == 0) {
        rho[i] = r_i;
        cnt[i] = static_cast<int>(ns);
        zcnt[i] = static_cast<int>(nz);
    }
}

// ---------------------------------------------------------------------------
// Top-k rule
// ---------------------------------------------------------------------------

// Count histogram of the current digit over entries whose higher digits match
// prefix; workgroups cover row ranges, LDS histogram, then added to the global
// one (integer adds: the totals do not depend on order).
template <typename TM>
__global__ __launch_bounds__(256) void k_topk_hist(const TM* __restrict__ M, long ldm, ElemCtx c,
                                                   int rows_per_block, int shift, int width,
                                                   unsigned long long prefix, int first,
                                                   unsigned long long* __restrict__ ghist)
{
    using E = Elem<TM>;
    using K = typename E::K;
    __shared__ unsigned int hist[RB];
    for (int q = threadIdx.x; q < RB; q += 256) hist[q] = 0u;
    __syncthreads();
    const int m1 = c.m - 1;
    const long r0 = static_cast<long>(blockIdx.x) * rows_per_block;
    const long r1 = (r0 + rows_per_block < c.n) ? r0 + rows_per_block : c.n;
    for (long i = r0; i < r1; i++) {
        const TM* row = M + i * ldm;
        const double ai = c.alpha[i];
        for (int j = threadIdx.x; j < m1; j += 256) {
            bool z;
            const TM t = E::val(ai + c.beta[j], row[j], c.k, c.reg, z);
            const K key = E::key(t);
            if (first || (static_cast<unsigned long long>(key >> (shift + width)) == prefix))
                atomicAdd(&hist[(key >> shift) & ((K(1) << width) - 1)], 1u);
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < (1 << width); q += 256)
        if (hist[q]) atomicAdd(ghist + q, static_cast<unsigned long long>(hist[q]));
}

template <typename TM>
__global__ __launch_bounds__(256) void k_row_ties(const TM* __restrict__ M, long ldm, ElemCtx c,
                                                  unsigned long long kstar, int* __restrict__ cnt)
{
    using E = Elem<TM>;
    const int i = blockIdx.x;
    const TM* row = M + static_cast<long>(i) * ldm;
    const double ai = c.alpha[i];
    long long nt = 0;
    for (int j = threadIdx.x; j < c.m - 1; j += 256) {
        bool z;
        const TM t = E::val(ai + c.beta[j], row[j], c.k, c.reg, z);
        nt += (static_cast<unsigned long long>(E::key(t)) == kstar) ? 1 : 0;
    }
    nt = block_sum_ll(nt);
    if (threadIdx.x == 0) cnt[i] = static_cast<int>(nt);
}

// ---------------------------------------------------------------------------
// Kept-set count and write
// ---------------------------------------------------------------------------

// Row kernels walk row i in 256-column chunks; for the top-k rule the rank of
// a tie is tie_off[i] + ties earlier in the row, so ties are taken in
// row-major order.
template <typename TM, bool WRITE>
__global__ __launch_bounds__(256) void k_row_kept(const TM* __restrict__ M, long ldm, ElemCtx c,
                                                  KeepRule rule, int* __restrict__ cnt,
                                                  const int* __restrict__ rowptr,
                                                  int* __restrict__ colind,
                                                  double* __restrict__ val)
{
    using E = Elem<TM>;
    const int i = blockIdx.x;
    const TM* row = M + static_cast<long>(i) * ldm;
    const double ai = c.alpha[i];
    const double inv_reg = 1.0 / c.reg;
    const int m1 = c.m - 1;
    long long tie_seen = 0;
    long long written = 0;
    const double r_i = (rule.mode == 0) ? rule.rho[i] : 0.0;
    const long long toff = (rule.mode == 1) ? rule.tie_off[i] : 0;
    for (int j0 = 0; j0 < m1; j0 += 256) {
        const int j = j0 + threadIdx.x;
        bool kept = false;
        TM t = TM(0);
        bool tie = false;
        if (j < m1) {
            bool z;
            t = E::val(ai + c.beta[j], row[j], c.k, c.reg, z);
            if (rule.mode == 0) {
                kept = (static_cast<double>(t) >= fmin(rule.tau[j], r_i)) || z;
            } else {
                const unsigned long long key = static_cast<unsigned long long>(E::key(t));
                kept = key > rule.kstar || (rule.preserve && (i == 0 || j == 0));
                tie = (key == rule.kstar);
            }
        }
        if (rule.mode == 1) {
            int ntie;
            const int tp = block_flag_prefix(tie, &ntie);
            if (tie && toff + tie_seen + tp < rule.tie_take) kept = true;
            tie_seen += ntie;
        }
        const bool stored = kept && t != TM(0);
        int nst;
        const int pos = block_flag_prefix(stored, &nst);
        if (WRITE && stored) {
            const long q = static_cast<long>(rowptr[i]) + written + pos;
            colind[q] = j;
            val[q] = static_cast<double>(t) * inv_reg;
        }
        written += nst;
    }
    if (!WRITE && threadIdx.x == 0) cnt[i] = static_cast<int>(written);
}

// ---------------------------------------------------------------------------
// CSC from CSR: a stable counting sort of the stored entries by column
// ---------------------------------------------------------------------------

// Workgroup b owns rows [b R, b R + R); cnt[b][j] = its entries in column j.
// Columns within a row are distinct, so the adds of one row never collide;
// the counts are integers, so they do not depend on the order of the adds.
__global__ __launch_bounds__(256) void k_csc_count(const int* __restrict__ rowptr,
                                                   const int* __restrict__ colind, int n,
                                                   int rows_per_block, int m1,
                                                   int* __restrict__ cnt)
{
    int* mine = cnt + static_cast<long>(blockIdx.x) * m1;
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = (r0 + rows_per_block < n) ? r0 + rows_per_block : n;
    if (r0 >= r1) return;
    for (int k = rowptr[r0] + threadIdx.x; k < rowptr[r1]; k += 256) atomicAdd(&mine[colind[k]], 1);
}

// Per column j: off[b][j] = entries of column j in workgroups before b
// (exclusive prefix over b, in place), total[j] = all of them.
__global__ __launch_bounds__(256) void k_csc_prefix(int* __restrict__ cnt, int nb, int m1,
                                                    int* __restrict__ total)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= m1) return;
    int run = 0;
    // 16 loads in flight per pass: the counts are independent, only the
    // running sum is serial
    for (int b0 = 0; b0 < nb; b0 += 16) {
        int c[16];
#pragma unroll
        for (int q = 0; q < 16; q++)
            c[q] = (b0 + q < nb) ? cnt[static_cast<long>(b0 + q) * m1 + j] : 0;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            if (b0 + q < nb) cnt[static_cast<long>(b0 + q) * m1 + j] = run;
            run += c[q];
        }
    }
    total[j] = run;
}

// Workgroup b scatters its rows in order: entry (i, j) goes to
// colptr[j] + off[b][j], then off[b][j] += 1; rows are separated by a barrier,
// so within a column the entries land in row order.
__global__ __launch_bounds__(256) void k_csc_scatter(const int* __restrict__ rowptr,
                                                     const int* __restrict__ colind,
                                                     const double* __restrict__ rval, int n,
                                                     int rows_per_block, int m1,
                                                     const int* __restrict__ colptr,
                                                     int* __restrict__ off,
                                                     int* __restrict__ rowind,
                                                     double* __restrict__ cval)
{
    int* mine = off + static_cast<long>(blockIdx.x) * m1;
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = (r0 + rows_per_block < n) ? r0 + rows_per_block : n;
    for (int i = r0; i < r1; i++) {
        for (int k = rowptr[i] + threadIdx.x; k < rowptr[i + 1]; k += 256) {
            const int j = colind[k];
            const int q = colptr[j] + mine[j];
            mine[j] += 1;
            rowind[q] = i;
            cval[q] = rval[k];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Scans
// ---------------------------------------------------------------------------

// One 1024-thread workgroup; thread t scans a contiguous chunk.
template <typename TO>
__global__ __launch_bounds__(1024) void k_scan(const int* __restrict__ cnt, int count,
                                               TO* __restrict__ out, long long* total)
{
    __shared__ long long wt[16];
    const int t = threadIdx.x;
    const int per = (count + 1023) / 1024;
    const int b = t * per, e = (b + per < count) ? b + per : count;
    long long own = 0;
    for (int q = b; q < e; q++) own += cnt[q];
    // inclusive wave scan
    long long inc = own;
    const int lane = t & 63, wave = t >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long u = __shfl_up(inc, o);
        if (lane >= o) inc += u;
    }
    if (lane == 63) wt[wave] = inc;
    __syncthreads();
    long long off = 0;
    for (int w = 0; w < wave; w++) off += wt[w];
    long long run = off + inc - own;
    for (int q = b; q < e; q++) {
        out[q] = static_cast<TO>(run);
        run += cnt[q];
    }
    if (t == 1023) {
        long long tot = 0;
        for (int w = 0; w < 16; w++) tot += wt[w];
        if (total) *total = tot;
        if (sizeof(TO) == sizeof(int)) out[count] = static_cast<TO>(tot);
    }
}

// ---------------------------------------------------------------------------
// Sparse products (one wave per row / column, fixed summation order)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_spmv_rows(const int* __restrict__ rowptr,
                                                   const int* __restrict__ colind,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ x, int n,
                                                   double* __restrict__ y,
                                                   const double* __restrict__ h1, double shift,
                                                   int div)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    double s = 0.0;
    for (int q = rowptr[i] + lane; q < rowptr[i + 1]; q += 64) s += val[q] * x[colind[q]];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) y[i] = div ? s / (h1[i] + shift) : s;
}

__global__ __launch_bounds__(256) void k_spmv_cols(const int* __restrict__ colptr,
                                                   const int* __restrict__ rowind,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ y, int m1,
                                                   const double* __restrict__ z,
                                                   const double* __restrict__ h2, double shift,
                                                   int mode, double* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= m1) return;
    double s = 0.0;
    for (int q = colptr[j] + lane; q < colptr[j + 1]; q += 64) s += val[q] * y[rowind[q]];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0)
        out[j] = mode == 0 ? s : mode == 1 ? (h2[j] * z[j] + shift * z[j] - s) : (h2[j] * z[j] + s);
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

template <typename TM>
static const TM* cM(const void* p)
{
    return static_cast<const TM*>(p);
}

void launch_transpose(int bulk, const void* M, long ldm, int n, int m, void* MT, long ldt,
                      hipStream_t s)
{
    const dim3 grid(ceil_div(m, 64), ceil_div(n, 64));
    if (bulk == 0)
        hipLaunchKernelGGL(k_transpose<float>, grid, dim3(256), 0, s, cM<float>(M), ldm, n, m,
                           static_cast<float*>(MT), ldt);
    else
        hipLaunchKernelGGL(k_transpose<double>, grid, dim3(256), 0, s, cM<double>(M), ldm, n, m,
                           static_cast<double*>(MT), ldt);
    KERNEL_CHECK();
}

void launch_col_threshold(int bulk, const void* MT, long ldt, const ElemCtx& c, double delta,
                          int find_small, double* tau, hipStream_t s)
{
    if (c.m <= 1) return;
    if (bulk == 0)
        hipLaunchKernelGGL(k_col_threshold<float>, dim3(c.m - 1), dim3(256), 0, s,
                           cM<float>(MT), ldt, c, delta, find_small, tau);
    else
        hipLaunchKernelGGL(k_col_threshold<double>, dim3(c.m - 1), dim3(256), 0, s,
                           cM<double>(MT), ldt, c, delta, find_small, tau);
    KERNEL_CHECK();
}

void launch_row_threshold(int bulk, const void* M, long ldm, const ElemCtx& c,
                          const double* tau, double delta2, double* rho, int* cnt, int* zcnt,
                          hipStream_t s)
{
    if (bulk == 0)
        hipLaunchKernelGGL(k_row_threshold<float>, dim3(c.n), dim3(256), 0, s, cM<float>(M), ldm,
                           c, tau, delta2, rho, cnt, zcnt);
    else
        hipLaunchKernelGGL(k_row_threshold<double>, dim3(c.n), dim3(256), 0, s, cM<double>(M),
                           ldm, c, tau, delta2, rho, cnt, zcnt);
    KERNEL_CHECK();
}

void topk_select(int bulk, const void* M, long ldm, const ElemCtx& c, long long k,
                 SparseHessian& H, unsigned long long* kstar, long long* n_greater,
                 hipStream_t s)
{
    const int KB = bulk == 0 ? 32 : 64;
    const int npass = (KB + RD - 1) / RD;
    const int rows_per_block = std::max(1, ceil_div(c.n, 2048));
    const int nblocks = ceil_div(c.n, rows_per_block);
    H.hist.ensure(RB);
    H.hll.ensure(RB);
    unsigned long long prefix = 0;
    long long above = 0;
    for (int p = 0; p < npass; p++) {
        const int hi = KB - p * RD;
        const int width = hi < RD ? hi : RD;
        const int shift = hi - width;
        HIP_CHECK(hipMemsetAsync(H.hist.p, 0, RB * sizeof(unsigned long long), s));
        if (bulk == 0)
            hipLaunchKernelGGL(k_topk_hist<float>, dim3(nblocks), dim3(256), 0, s, cM<float>(M),
                               ldm, c, rows_per_block, shift, width, prefix, p == 0 ? 1 : 0,
                               H.hist.p);
        else
            hipLaunchKernelGGL(k_topk_hist<double>, dim3(nblocks), dim3(256), 0, s,
                               cM<double>(M), ldm, c, rows_per_block, shift, width, prefix,
                               p == 0 ? 1 : 0, H.hist.p);
        KERNEL_CHECK();
        HIP_CHECK(hipMemcpyAsync(H.hll.p, H.hist.p, RB * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        int b = (1 << width) - 1;
        for (; b >= 0; b--) {
            const long long h = H.hll.p[b];
            if (above + h >= k) break;
            above += h;
        }
        if (b < 0) fail(E_INTERNAL, "top-k radix select lost the k-th entry");
        prefix = (prefix << width) | static_cast<unsigned long long>(b);
    }
    *kstar = prefix;
    *n_greater = above;
}

void launch_row_ties(int bulk, const void* M, long ldm, const ElemCtx& c,
                     unsigned long long kstar, int* cnt, hipStream_t s)
{
    if (bulk == 0)
        hipLaunchKernelGGL(k_row_ties<float>, dim3(c.n), dim3(256), 0, s, cM<float>(M), ldm, c,
                           kstar, cnt);
    else
        hipLaunchKernelGGL(k_row_ties<double>, dim3(c.n), dim3(256), 0, s, cM<double>(M), ldm, c,
                           kstar, cnt);
    KERNEL_CHECK();
}

void launch_row_count(int bulk, const void* M, long ldm, const ElemCtx& c, const KeepRule& r,
                      int* cnt, hipStream_t s)
{
    if (bulk == 0)
        hipLaunchKernelGGL((k_row_kept<float, false>), dim3(c.n), dim3(256), 0, s, cM<float>(M),
                           ldm, c, r, cnt, nullptr, nullptr, nullptr);
    else
        hipLaunchKernelGGL((k_row_kept<double, false>), dim3(c.n), dim3(256), 0, s,
                           cM<double>(M), ldm, c, r, cnt, nullptr, nullptr, nullptr);
    KERNEL_CHECK();
}

void launch_row_write(int bulk, const void* M, long ldm, const ElemCtx& c, const KeepRule& r,
                      const int* rowptr, int* colind, double* val, hipStream_t s)
{
    if (bulk == 0)
        hipLaunchKernelGGL((k_row_kept<float, true>), dim3(c.n), dim3(256), 0, s, cM<float>(M),
                           ldm, c, r, nullptr, rowptr, colind, val);
    else
        hipLaunchKernelGGL((k_row_kept<double, true>), dim3(c.n), dim3(256), 0, s,
                           cM<double>(M), ldm, c, r, nullptr, rowptr, colind, val);
    KERNEL_CHECK();
}

void launch_csr_to_csc(const SparseHessian& H, int* work, int rows_per_block, int* colcnt,
                       int* colptr, long long* total, int* rowind, double* cval, hipStream_t s)
{
    const int n = H.n, m1 = H.m - 1;
    if (m1 <= 0) return;
    const int nb = ceil_div(n, rows_per_block);
    HIP_CHECK(hipMemsetAsync(work, 0, sizeof(int) * static_cast<size_t>(nb) * m1, s));
    hipLaunchKernelGGL(k_csc_count, dim3(nb), dim3(256), 0, s, H.rowptr.p, H.colind.p, n,
                       rows_per_block, m1, work);
    hipLaunchKernelGGL(k_csc_prefix, dim3(ceil_div(m1, 256)), dim3(256), 0, s, work, nb, m1,
                       colcnt);
    KERNEL_CHECK();
    launch_scan(colcnt, m1, colptr, total, s);
    hipLaunchKernelGGL(k_csc_scatter, dim3(nb), dim3(256), 0, s, H.rowptr.p, H.colind.p,
                       H.rval.p, n, rows_per_block, m1, colptr, work, rowind, cval);
    KERNEL_CHECK();
}

void launch_scan(const int* cnt, int count, int* out, long long* total, hipStream_t s)
{
    hipLaunchKernelGGL(k_scan<int>, dim3(1), dim3(1024), 0, s, cnt, count, out, total);
    KERNEL_CHECK();
}

void launch_scan64(const int* cnt, int count, long long* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_scan<long long>, dim3(1), dim3(1024), 0, s, cnt, count, out, nullptr);
    KERNEL_CHECK();
}

void launch_spmv_rows(const SparseHessian& H, const double* x, double* y, bool div,
                      double shift, hipStream_t s)
{
    if (H.n <= 0) return;
    hipLaunchKernelGGL(k_spmv_rows, dim3(ceil_div(H.n, 4)), dim3(256), 0, s, H.rowptr.p,
                       H.colind.p, H.rval.p, x, H.n, y, H.h1.p, shift, div ? 1 : 0);
    KERNEL_CHECK();
}

void launch_spmv_cols(const SparseHessian& H, const double* y, const double* z, double shift,
                      int mode, double* out, hipStream_t s)
{
    if (H.m <= 1) return;
    hipLaunchKernelGGL(k_spmv_cols, dim3(ceil_div(H.m - 1, 4)), dim3(256), 0, s, H.colptr.p,
                       H.rowind.p, H.cval.p, y, H.m - 1, z, H.h2.p, shift, mode, out);
    KERNEL_CHECK();
}

}  // namespace regot
