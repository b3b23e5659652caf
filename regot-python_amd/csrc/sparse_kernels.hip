// sparse_kernels.hip -- sparsification of T and sparse Hessian products for gfx950.
//
// Selection is radix-based on the IEEE bit pattern of T (non-negative values
// order like their bit patterns): 11-bit digits from the top, 3 passes for
// FP32 keys, 6 for FP64 keys.  Mass-weighted selection (threshold rule)
// histograms FP64 masses in LDS; count selection (top-k) histograms counts.
// Compaction is count -> exclusive scan -> ordered write; the write position
// of each kept entry comes from wave ballots, so no global atomics decide
// placement and the output is identical run to run.
#include <algorithm>
#include <cmath>
#include <cstring>

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
// Threshold rule, tau_j from row-major M (FP32 bulk; no transpose)
//
// tau_j is the smallest key K of column j with sum_{key <= K} T_ij > delta.
// Level 1 (k_tau_hist): per column, FP64 mass and count of T over 64 buckets
// by binade (tau_bucket: every tau_j exceeds delta / n, since n entries of
// at most delta / n cannot pass delta, so bucket 0 takes T below the binade
// of delta / n, buckets 1..62 one binade each upwards, bucket 63 the rest)
// and the largest key, by strips of 64 columns (lane = column) over row
// blocks.  k_tau_fold sums the row blocks in a fixed order and finds each
// column's boundary bucket (or settles tau_j when the column's mass never
// passes delta); k_tau_gather collects the entries of the boundary bucket;
// k_tau_select ends the selection on those candidates with the radix passes
// of k_col_threshold over their full keys.  Any bucketing monotone in the key
// gives the same tau_j; this one keeps the boundary buckets to one binade.
// ---------------------------------------------------------------------------

constexpr int TB = 64;    // level-1 buckets (tau_bucket)
constexpr int TC = 64;    // columns per strip workgroup
constexpr int TU = 16;    // rows per wave step (loads in flight)

// Level-1 bucket of a key (T >= 0): its binade relative to elo (the biased
// exponent of delta / n), clamped to [0, 63]; monotone in the key.
__device__ __forceinline__ int tau_bucket(unsigned int key, int elo)
{
    const int b = static_cast<int>((key >> 23) & 255u) - elo + 1;
    return b < 0 ? 0 : (b > TB - 1 ? TB - 1 : b);
}

__global__ __launch_bounds__(256) void k_tau_hist(const float* __restrict__ M, long ldm, ElemCtx c,
                                                  int elo, int rows_per_blk,
                                                  double* __restrict__ pmass,
                                                  unsigned int* __restrict__ pcnt,
                                                  unsigned int* __restrict__ pmax)
{
    constexpr int P = TC + 1;   // padded rows: the bucket-fastest copy-out is conflict-free
    __shared__ double hm[TB * P];
    __shared__ unsigned int hc[TB * P];
    __shared__ unsigned int hx[TC];
    // rows are interleaved over the four waves (wave w takes i0 + w + 4u);
    // alpha_i comes through vector loads (a scalar load would share its wait
    // counter with the LDS atomics)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m1 = c.m - 1;
    const int j = blockIdx.x * TC + lane;
    for (int q = threadIdx.x; q < TB * P; q += 256) {
        hm[q] = 0.0;
        hc[q] = 0u;
    }
    if (threadIdx.x < TC) hx[threadIdx.x] = 0u;
    __syncthreads();
    const bool ok = j < m1;
    const double bj = ok ? c.beta[j] : 0.0;
    const long i0 = static_cast<long>(blockIdx.y) * rows_per_blk;
    const long i1 = (i0 + rows_per_blk < c.n) ? i0 + rows_per_blk : c.n;
    unsigned int kmax = 0u;
    if (ok) {
        const float* Mj = M + j;
        for (long ib = i0 + wave; ib < i1; ib += 4 * TU) {
            float mv[TU];
            double av[TU];
#pragma unroll
            for (int u = 0; u < TU; u++) {
                const long i = ib + 4L * u;
                mv[u] = i < i1 ? Mj[i * ldm] : 0.f;
                av[u] = i < i1 ? c.alpha[i] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < TU; u++) {
                if (ib + 4L * u >= i1) break;
                bool z;
                const float t = Elem<float>::val(av[u] + bj, mv[u], c.k, c.reg, z);
                const unsigned int key = __float_as_uint(t);
                const int b = tau_bucket(key, elo);
                atomicAdd(&hm[b * P + lane], static_cast<double>(t));
                atomicAdd(&hc[b * P + lane], 1u);
                kmax = key > kmax ? key : kmax;
            }
        }
        atomicMax(&hx[lane], kmax);
    }
    __syncthreads();
    // partials [row block][column][bucket]
    const long rb = blockIdx.y;
    for (int q = threadIdx.x; q < TB * TC; q += 256) {
        const int b = q % TB, l = q / TB;
        const int jj = blockIdx.x * TC + l;
        if (jj < m1) {
            pmass[(rb * m1 + jj) * TB + b] = hm[b * P + l];
            pcnt[(rb * m1 + jj) * TB + b] = hc[b * P + l];
        }
    }
    if (threadIdx.x < TC && blockIdx.x * TC + static_cast<int>(threadIdx.x) < m1)
        pmax[rb * m1 + blockIdx.x * TC + threadIdx.x] = hx[threadIdx.x];
}

// Per column (one wave, lane = bucket): the boundary bucket (smallest b with
// sum_{b' <= b} mass_b' > delta), the mass below it and its entry count; or,
// when the column's whole mass stays <= delta, tau_j itself (+inf with
// find_small, else the largest entry).  The row blocks are summed in order,
// the buckets by a fixed-order wave scan.
__global__ __launch_bounds__(256) void k_tau_fold(const double* __restrict__ pmass,
                                                  const unsigned int* __restrict__ pcnt,
                                                  const unsigned int* __restrict__ pmax, int nrb,
                                                  int m1, double delta, int find_small,
                                                  int* __restrict__ bucket,
                                                  double* __restrict__ base,
                                                  int* __restrict__ count,
                                                  double* __restrict__ tau)
{
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= m1) return;
    double mb = 0.0;
    unsigned int cb = 0u, mx = 0u;
    for (int r = 0; r < nrb; r++) {
        const long o = (static_cast<long>(r) * m1 + j) * TB + lane;
        mb += pmass[o];
        cb += pcnt[o];
        const unsigned int v = pmax[static_cast<long>(r) * m1 + j];
        mx = v > mx ? v : mx;
    }
    const double inc = wave_incl_scan(mb);
    const unsigned long long hit = __ballot(inc > delta);
    if (hit == 0ull) {
        if (lane == 0) {
            bucket[j] = -1;
            count[j] = 0;
            tau[j] = find_small ? INFINITY : static_cast<double>(__uint_as_float(mx));
        }
        return;
    }
    const int B = __ffsll(static_cast<long long>(hit)) - 1;
    if (lane == B) {
        bucket[j] = B;
        base[j] = inc - mb;
        count[j] = static_cast<int>(cb);
    }
}

// The entries of each column's boundary bucket, into cand[off_j ..): row
// block rb writes after the blocks before it (their level-1 counts of the
// bucket), the four waves of a workgroup claim slots by LDS counters.
__global__ __launch_bounds__(256) void k_tau_gather(const float* __restrict__ M, long ldm,
                                                    ElemCtx c, int elo, int rows_per_blk,
                                                    const int* __restrict__ bucket,
                                                    const int* __restrict__ off,
                                                    const unsigned int* __restrict__ pcnt,
                                                    float* __restrict__ cand)
{
    __shared__ int any;
    __shared__ int lfill[TC];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m1 = c.m - 1;
    const int j = blockIdx.x * TC + lane;
    const int B = j < m1 ? bucket[j] : -1;
    if (threadIdx.x == 0) any = 0;
    if (threadIdx.x < TC) lfill[threadIdx.x] = 0;
    __syncthreads();
    if (B >= 0) any = 1;
    __syncthreads();
    if (!any || B < 0) return;
    const long rb = blockIdx.y;
    int o = off[j];
    for (long r = 0; r < rb; r++) o += static_cast<int>(pcnt[(r * m1 + j) * TB + B]);
    const int cap = static_cast<int>(pcnt[(rb * m1 + j) * TB + B]);
    if (cap == 0) return;
    const double bj = c.beta[j];
    const long i0 = rb * rows_per_blk;
    const long i1 = (i0 + rows_per_blk < c.n) ? i0 + rows_per_blk : c.n;
    const float* Mj = M + j;
    for (long ib = i0 + wave; ib < i1; ib += 4 * TU) {
        float mv[TU];
        double av[TU];
#pragma unroll
        for (int u = 0; u < TU; u++) {
            const long i = ib + 4L * u;
            mv[u] = i < i1 ? Mj[i * ldm] : 0.f;
            av[u] = i < i1 ? c.alpha[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < TU; u++) {
            if (ib + 4L * u >= i1) break;
            bool z;
            const float t = Elem<float>::val(av[u] + bj, mv[u], c.k, c.reg, z);
            if (tau_bucket(__float_as_uint(t), elo) == B) {
                const int pos = atomicAdd(&lfill[lane], 1);
                if (pos < cap) cand[o + pos] = t;
            }
        }
    }
}

// Radix passes of k_col_threshold over the candidates of a column's boundary
// bucket (their full keys: 11-, 11- and 10-bit digits), from the bucket's
// base mass.  If rounding in the re-summed masses leaves no bucket passing
// delta, the largest candidate key is taken (the entry whose mass makes the
// sum pass).
__global__ __launch_bounds__(256) void k_tau_select(const int* __restrict__ bucket,
                                                    const double* __restrict__ base_in,
                                                    const int* __restrict__ count,
                                                    const int* __restrict__ off,
                                                    const float* __restrict__ cand, double delta,
                                                    double* __restrict__ tau)
{
    __shared__ double hist[RB];
    __shared__ unsigned int kmax_s[4];
    const int j = blockIdx.x;
    const int B = bucket[j];
    if (B < 0) return;
    const float* cj = cand + off[j];
    const int cn = count[j];
    unsigned int prefix = 0u;
    double base = base_in[j];
    bool none = false;
    unsigned int kmax = 0u;
#pragma unroll 1
    for (int p = 0; p < 3; p++) {
        const int shift = p == 0 ? 21 : (p == 1 ? 10 : 0);
        const int width = p == 2 ? 10 : 11;
        for (int q = threadIdx.x; q < RB; q += 256) hist[q] = 0.0;
        __syncthreads();
        for (int e = threadIdx.x; e < cn; e += 256) {
            const float t = cj[e];
            const unsigned int key = __float_as_uint(t) & 0x7fffffffu;
            if (p == 0) kmax = key > kmax ? key : kmax;
            if (p == 0 || (key >> (shift + width)) == prefix)
                atomicAdd(&hist[(key >> shift) & ((1u << width) - 1u)], static_cast<double>(t));
        }
        __syncthreads();
        double before = 0.0;
        const int b = select_bucket<true>(hist, 1 << width, base, delta, &before);
        if (b < 0) {
            none = true;
            break;
        }
        base = before;
        prefix = (prefix << width) | static_cast<unsigned int>(b);
    }
    if (none) {
        unsigned int v = kmax;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned int w = static_cast<unsigned int>(__shfl_xor(static_cast<int>(v), o));
            v = w > v ? w : v;
        }
        if ((threadIdx.x & 63) == 0) kmax_s[threadIdx.x >> 6] = v;
        __syncthreads();
        unsigned int mx = kmax_s[0];
        for (int w = 1; w < 4; w++) mx = kmax_s[w] > mx ? kmax_s[w] : mx;
        if (threadIdx.x == 0) tau[j] = static_cast<double>(__uint_as_float(mx));
    } else if (threadIdx.x == 0) {
        tau[j] = static_cast<double>(__uint_as_float(prefix));
    }
}

// ---------------------------------------------------------------------------
// Top-k rule
// ---------------------------------------------------------------------------

// Four consecutive elements from column c0 (c0 % 4 == 0) of a 16-byte aligned
// row of `lim` valid entries; past lim the values are 0 and nothing is read.
__device__ __forceinline__ void load4(const float* __restrict__ row, int c0, int lim, float (&v)[4])
{
    if (c0 + 3 < lim) {
        const float4 u = *reinterpret_cast<const float4*>(row + c0);
        v[0] = u.x;
        v[1] = u.y;
        v[2] = u.z;
        v[3] = u.w;
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = (c0 + q < lim) ? row[c0 + q] : 0.0f;
    }
}

__device__ __forceinline__ void load4(const double* __restrict__ row, int c0, int lim,
                                      double (&v)[4])
{
    if (c0 + 3 < lim) {
        const double2 u0 = *reinterpret_cast<const double2*>(row + c0);
        const double2 u1 = *reinterpret_cast<const double2*>(row + c0 + 2);
        v[0] = u0.x;
        v[1] = u0.y;
        v[2] = u1.x;
        v[3] = u1.y;
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++) v[q] = (c0 + q < lim) ? row[c0 + q] : 0.0;
    }
}

// Exclusive scan of one int per thread over the 256-thread workgroup, in
// thread order; *total gets the sum.
__device__ int block_excl_scan_int(int v, int* total)
{
    __shared__ int wt[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(inc, o);
        if (lane >= o) inc += u;
    }
    __syncthreads();
    if (lane == 63) wt[wave] = inc;
    __syncthreads();
    int off = 0;
    for (int w = 0; w < wave; w++) off += wt[w];
    *total = wt[0] + wt[1] + wt[2] + wt[3];
    return off + inc - v;
}

// Exclusive prefix, over the lanes of a wave, of per-lane flag counts (four
// flags per lane, lane-major order) by ballots and mbcnt; *total gets the
// wave's count.
__device__ __forceinline__ int wave_prefix4(const bool (&f)[4], int* total)
{
    int pre = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const unsigned long long b = __ballot(f[q]);
        pre = __builtin_amdgcn_mbcnt_hi(static_cast<unsigned int>(b >> 32),
                                        __builtin_amdgcn_mbcnt_lo(static_cast<unsigned int>(b),
                                                                  pre));
        tot += __popcll(b);
    }
    *total = tot;
    return pre;
}

// Count histogram of the current digit over entries whose higher digits match
// prefix.  Workgroup b covers rows [b R, b R + R); thread t owns the column
// quads 4t + 1024 s, keeps their beta in registers and walks the rows four at
// a time (four vector loads in flight).  LDS histogram, then added to the
// global one (integer adds: the totals do not depend on order).
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
    const K dmask = (K(1) << width) - 1;
    // run-length aggregation: T is smooth, so a thread's consecutive keys
    // mostly share a bucket (all of them in the first pass, whose digit is
    // nearly the exponent); one LDS add per run instead of one per entry
    unsigned int run_b = 0u, run_c = 0u;
    for (int c0 = threadIdx.x * 4; c0 < m1; c0 += 1024) {
        double b4[4];
        load4(c.beta, c0, c.m, b4);
        for (long i0 = r0; i0 < r1; i0 += 4) {
            TM v[4][4];
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (i0 + u < r1) load4(M + (i0 + u) * ldm, c0, c.m, v[u]);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (i0 + u >= r1) break;
                const double ai = c.alpha[i0 + u];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (c0 + q >= m1) break;
                    bool z;
                    const TM t = E::val(ai + b4[q], v[u][q], c.k, c.reg, z);
                    const K key = E::key(t);
                    if (first || (static_cast<unsigned long long>(key >> (shift + width)) == prefix)) {
                        const unsigned int bk = static_cast<unsigned int>((key >> shift) & dmask);
                        if (bk != run_b) {
                            if (run_c) atomicAdd(&hist[run_b], run_c);
                            run_b = bk;
                            run_c = 0u;
                        }
                        run_c++;
                    }
                }
            }
        }
    }
    if (run_c) atomicAdd(&hist[run_b], run_c);
    __syncthreads();
    for (int q = threadIdx.x; q < (1 << width); q += 256)
        if (hist[q]) atomicAdd(ghist + q, static_cast<unsigned long long>(hist[q]));
}

// ---- Top-k with one radix pass over M --------------------------------------
// After the first (top-digit) histogram pass, one more pass over M does the
// rest of the selection's reading: every row counts its entries kept without
// a tie slot among those outside the boundary bucket b1 (top digit above b1,
// or preserved; nonzero), and the entries inside it -- the only ones the
// k-th key can be among -- are gathered with their positions.  The remaining
// radix passes, the ties (key == k-th key) and the per-row counts then run
// on the candidates alone; the CSR write is the third and last read of M.

constexpr int kCandStage = 1024;   // one 1024-column sub-step of a row

// Row i (one workgroup, 4096 columns per step): gcnt[i] = 2 x the stored
// entries kept without a tie slot outside bucket b1, tcnt[i] = 0 (the
// candidates' ties are added later); entries of bucket b1 are staged in LDS
// and appended to cand (one global reservation per flush, at most cap
// entries in all).
template <typename TM>
__global__ __launch_bounds__(256) void k_topk_gather(const TM* __restrict__ M, long ldm, ElemCtx c,
                                                     int shift1, unsigned long long b1,
                                                     int preserve, int* __restrict__ tcnt,
                                                     int* __restrict__ gcnt,
                                                     TopkCand* __restrict__ cand,
                                                     unsigned long long* __restrict__ ncand,
                                                     long long cap)
{
    using E = Elem<TM>;
    __shared__ unsigned long long skey[kCandStage];
    __shared__ int scol[kCandStage];
    __shared__ int scnt;
    __shared__ long long sbase;
    const int i = blockIdx.x;
    const TM* row = M + static_cast<long>(i) * ldm;
    const double ai = c.alpha[i];
    const int m1 = c.m - 1;
    if (threadIdx.x == 0) scnt = 0;
    __syncthreads();
    int ng = 0;
    // 4096 columns per step (four 16-byte loads per thread in flight,
    // unconditional on full steps); each 1024-column sub-step forms its
    // values and flags branch-free, appends its candidates with one prefix
    // (ballots and mbcnt) and one LDS reservation per wave, and stages at
    // most 1024 of them, flushed after it
    for (int s0 = 0; s0 < m1; s0 += 4096) {
        TM v[4][4];
        double b4[4][4];
        if (s0 + 4096 <= m1) {   // workgroup-uniform
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int cu = s0 + 1024 * u + threadIdx.x * 4;
                load4(row, cu, cu + 4, v[u]);
                load4(c.beta, cu, cu + 4, b4[u]);
            }
        } else {
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int cu = s0 + 1024 * u + threadIdx.x * 4;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    v[u][q] = TM(0);
                    b4[u][q] = 0.0;
                }
                if (cu < m1) {
                    load4(row, cu, c.m, v[u]);
                    load4(c.beta, cu, c.m, b4[u]);
                }
            }
        }
        const int lane = threadIdx.x & 63;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (s0 + 1024 * u >= m1) break;   // uniform
            unsigned long long key[4];
            bool isc[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int j = s0 + 1024 * u + threadIdx.x * 4 + q;
                const bool in = j < m1;
                bool z;
                const TM t = E::val(ai + b4[u][q], v[u][q], c.k, c.reg, z);
                key[q] = static_cast<unsigned long long>(E::key(t));
                const unsigned long long top = key[q] >> shift1;
                isc[q] = in && top == b1;
                const bool base = top > b1 || (preserve && (i == 0 || j == 0));
                ng += (in && !isc[q] && base && t != TM(0)) ? 1 : 0;
            }
            int tot;
            int pos = wave_prefix4(isc, &tot);
            if (tot > 0) {   // wave-uniform
                int wb = 0;
                if (lane == 0) wb = atomicAdd(&scnt, tot);
                pos += __shfl(wb, 0);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (isc[q]) {
                        skey[pos] = key[q];
                        scol[pos] = s0 + 1024 * u + threadIdx.x * 4 + q;
                        pos++;
                    }
                }
            }
            // a sub-step stages at most 1024 entries: flush them now
            __syncthreads();
            const int cnt = scnt;
            if (cnt > 0) {
                if (threadIdx.x == 0)
                    sbase = static_cast<long long>(
                        atomicAdd(ncand, static_cast<unsigned long long>(cnt)));
                __syncthreads();
                const long long base = sbase;
                for (int e = threadIdx.x; e < cnt; e += 256)
                    if (base + e < cap) {
                        TopkCand cd;
                        cd.key = skey[e];
                        cd.i = i;
                        cd.j = scol[e];
                        cand[base + e] = cd;
                    }
                __syncthreads();
                if (threadIdx.x == 0) scnt = 0;
                __syncthreads();
            }
        }
    }
    int tot;
    (void)block_excl_scan_int(ng, &tot);
    if (threadIdx.x == 0) {
        gcnt[i] = 2 * tot;
        tcnt[i] = 0;
    }
}

// Count histogram of one digit of the candidates' keys whose higher digits
// equal prefix (run-merged LDS adds, then global adds).
__global__ __launch_bounds__(256) void k_cand_hist(const TopkCand* __restrict__ cand, long long cap,
                                                   const unsigned long long* __restrict__ ncand,
                                                   int shift, int width,
                                                   unsigned long long prefix,
                                                   unsigned long long* __restrict__ ghist)
{
    __shared__ unsigned int hist[RB];
    for (int q = threadIdx.x; q < RB; q += 256) hist[q] = 0u;
    __syncthreads();
    // the gather's own count, never more than the slots it could write
    const long long nc = static_cast<long long>(*ncand) < cap ? static_cast<long long>(*ncand) : cap;
    const unsigned long long dmask = (1ull << width) - 1ull;
    unsigned int run_b = 0u, run_c = 0u;
    for (long long e = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; e < nc;
         e += static_cast<long long>(gridDim.x) * 256) {
        const unsigned long long key = cand[e].key;
        if ((key >> (shift + width)) != prefix) continue;
        const unsigned int bk = static_cast<unsigned int>((key >> shift) & dmask);
        if (bk != run_b) {
            if (run_c) atomicAdd(&hist[run_b], run_c);
            run_b = bk;
            run_c = 0u;
        }
        run_c++;
    }
    if (run_c) atomicAdd(&hist[run_b], run_c);
    __syncthreads();
    for (int q = threadIdx.x; q < (1 << width); q += 256)
        if (hist[q]) atomicAdd(ghist + q, static_cast<unsigned long long>(hist[q]));
}

// The candidates' share of the row tallies under the k-th key kstar: ties
// into tcnt; stored entries kept without a tie slot (key > kstar, or
// preserved; nonzero) add 2 to gcnt; a preserved tie in column 0 of a row
// i > 0 sets gcnt's low bit (k_row_tally's encoding).
__global__ __launch_bounds__(256) void k_cand_classify(const TopkCand* __restrict__ cand,
                                                       long long cap,
                                                       const unsigned long long* __restrict__ ncand,
                                                       unsigned long long kstar, int preserve,
                                                       int n, int* __restrict__ tcnt,
                                                       int* __restrict__ gcnt)
{
    const long long nc = static_cast<long long>(*ncand) < cap ? static_cast<long long>(*ncand) : cap;
    const long long e = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (e >= nc) return;
    const TopkCand cd = cand[e];
    if (cd.i < 0 || cd.i >= n) return;
    const bool pres = preserve && (cd.i == 0 || cd.j == 0);
    const bool base = cd.key > kstar || pres;
    if (base && cd.key != 0ull) atomicAdd(gcnt + cd.i, 2);
    if (cd.key == kstar) {
        atomicAdd(tcnt + cd.i, 1);
        if (cd.j == 0 && pres && cd.i > 0) atomicOr(gcnt + cd.i, 1);
    }
}

// Top-k tally of row i (one workgroup, 4096 columns per step): ties
// (key == kstar) into tcnt[i]; stored entries kept without a tie slot
// (key > kstar, or preserved; nonzero) times 2, plus 1 if column 0 is a tie of
// a row i > 0 under preserve, into gcnt[i].  k_topk_count turns these and the
// tie offsets into the stored count of the row.
template <typename TM>
__global__ __launch_bounds__(256) void k_row_tally(const TM* __restrict__ M, long ldm, ElemCtx c,
                                                   unsigned long long kstar, int preserve,
                                                   int* __restrict__ tcnt, int* __restrict__ gcnt)
{
    using E = Elem<TM>;
    const int i = blockIdx.x;
    const TM* row = M + static_cast<long>(i) * ldm;
    const double ai = c.alpha[i];
    const int m1 = c.m - 1;
    int nt = 0, ng = 0, tie0 = 0;
    for (int s0 = threadIdx.x * 4; s0 < m1; s0 += 4096) {
        TM v[4][4];
        double b4[4][4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int cu = s0 + 1024 * u;
            if (cu < m1) {
                load4(row, cu, c.m, v[u]);
                load4(c.beta, cu, c.m, b4[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int j = s0 + 1024 * u + q;
                if (j >= m1) break;
                bool z;
                const TM t = E::val(ai + b4[u][q], v[u][q], c.k, c.reg, z);
                const unsigned long long key = static_cast<unsigned long long>(E::key(t));
                const bool base = key > kstar || (preserve && (i == 0 || j == 0));
                const bool tie = key == kstar;
                nt += tie ? 1 : 0;
                ng += (base && t != TM(0)) ? 1 : 0;
                if (j == 0 && tie && preserve && i > 0) tie0 = 1;
            }
        }
    }
    int tot_t, tot_g;
    (void)block_excl_scan_int(nt, &tot_t);
    (void)block_excl_scan_int(ng, &tot_g);
    if (threadIdx.x == 0) {
        tcnt[i] = tot_t;
        gcnt[i] = 2 * tot_g + tie0;
    }
}

// Stored entries of row i under the top-k rule: the base count plus the
// nonzero ties it takes, the ties of the row having ranks tie_off[i] ..
// tie_off[i] + ties - 1 (row-major) and the first tie_take ranks being taken.
// A tie in column 0 of a row i > 0 is preserved, so it is counted once; every
// tie of row 0 is preserved; ties are nonzero unless kstar is the key of 0.
__global__ void k_topk_count(const int* tcnt, const int* __restrict__ gcnt,
                             const long long* __restrict__ tie_off, long long tie_take,
                             unsigned long long kstar, int preserve, int n, int* cnt)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const long long rem = tie_take - tie_off[i];
    const int nt = tcnt[i];
    const long long taken = rem < 0 ? 0 : (rem < nt ? rem : nt);
    long long extra = 0;
    if (kstar != 0ull && !(preserve && i == 0))
        extra = taken - (((gcnt[i] & 1) && taken > 0) ? 1 : 0);
    cnt[i] = (gcnt[i] >> 1) + static_cast<int>(extra);
}


// T of one FP32-bulk element with the raw v_exp_f32: bit for bit
// Elem<float>::val wherever that result is at least 2^-126; below, 0.
__device__ __forceinline__ float val_raw(double ab, float mij, double k)
{
    return __builtin_amdgcn_exp2f(static_cast<float>(fma(-k, static_cast<double>(mij), ab * k)));
}

// One 256-column step of a row (4 columns per lane, from cj = s0 + 4 lane):
// full steps (s0 + 256 <= m1) load unconditionally, the last partial step
// reads only columns < m1.
template <typename TM, bool TOPK>
__device__ __forceinline__ void rw_load(const TM* __restrict__ row, const ElemCtx& c,
                                        const KeepRule& rule, int s0, int m1, TM (&v)[4],
                                        double (&b)[4], double (&t)[4])
{
    const int cj = s0 + (threadIdx.x & 63) * 4;
    if (s0 + 256 <= m1) {   // wave-uniform
        load4(row, cj, cj + 4, v);
        load4(c.beta, cj, cj + 4, b);
        if (!TOPK) load4(rule.tau, cj, cj + 4, t);
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            v[q] = TM(0);
            b[q] = 0.0;
            t[q] = 0.0;
        }
        if (cj < m1) {
            load4(row, cj, m1, v);
            load4(c.beta, cj, m1, b);
            if (!TOPK) load4(rule.tau, cj, m1, t);
        }
    }
}

// The kept flags, values and the ordered write of one step (see k_row_write).
template <typename TM, bool TOPK, bool RAW>
__device__ __forceinline__ void rw_step(const ElemCtx& c, const KeepRule& rule, int i, int cj,
                                        int m1, double ai, double r_i, long long rem0,
                                        double inv_reg, long base_pos, long end_pos,
                                        const TM (&v)[4], const double (&b)[4],
                                        const double (&tt)[4], long long* tie_seen,
                                        long* written, int* __restrict__ colind,
                                        double* __restrict__ val)
{
    using E = Elem<TM>;
    TM t[4];
    bool base[4], tie[4];
    bool any_tie = false;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int j = cj + q;
        const bool in = j < m1;
        bool z = false;
        if constexpr (RAW)
            t[q] = val_raw(ai + b[q], v[q], c.k);
        else
            t[q] = E::val(ai + b[q], v[q], c.k, c.reg, z);
        if (TOPK) {
            const unsigned long long key = static_cast<unsigned long long>(E::key(t[q]));
            base[q] = in && (key > rule.kstar || (rule.preserve && (i == 0 || j == 0)));
            tie[q] = in && key == rule.kstar;
            any_tie = any_tie || tie[q];
        } else {
            base[q] = in && ((static_cast<double>(t[q]) >= fmin(tt[q], r_i)) || z);
            tie[q] = false;
        }
    }
    if (TOPK && __ballot(any_tie) != 0ull) {
        int tot;
        int r = wave_prefix4(tie, &tot);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (tie[q]) {
                if (*tie_seen + r < rem0) base[q] = true;
                r++;
            }
        }
        *tie_seen += tot;
    }
    bool st[4];
#pragma unroll
    for (int q = 0; q < 4; q++) st[q] = base[q] && t[q] != TM(0);
    int tot;
    long q0 = base_pos + *written + wave_prefix4(st, &tot);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (st[q] && q0 < end_pos) {
            colind[q0] = cj + q;
            val[q0] = static_cast<double>(t[q]) * inv_reg;
        }
        q0 += st[q] ? 1 : 0;
    }
    *written += tot;
}

// Ordered CSR write of row i under either rule: one wave per row, 256
// columns (4 per lane) per step, positions from ballots (no LDS, no
// barriers), the next step's loads in flight during the current step (full
// steps load unconditionally).  Top-k ties are taken in row-major order: a
// tie of chunk rank r is kept when tie_off[i] + ties of earlier chunks + r <
// tie_take.  RAW (FP32 top-k, k-th value at least 1.125 * 2^-126): T by the
// raw v_exp_f32, which equals the exact value on every entry the rule can
// keep (key >= the k-th key) and leaves the others below it; row 0 and the
// first step (column 0), where preserved entries are kept at any value,
// use the exact function.  Only columns < m - 1 are read (of M, beta and
// tau), so the kernel also runs on M' with the roles of alpha / beta and
// tau / rho exchanged, writing the CSC.
template <typename TM, bool TOPK, bool RAW>
__global__ __launch_bounds__(256) void k_row_write(const TM* __restrict__ M, long ldm, ElemCtx c,
                                                   KeepRule rule, const int* __restrict__ rowptr,
                                                   int* __restrict__ colind,
                                                   double* __restrict__ val)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= c.n) return;   // wave-uniform
    const int lane = threadIdx.x & 63;
    const TM* row = M + static_cast<long>(i) * ldm;
    const double ai = c.alpha[i];
    const double inv_reg = 1.0 / c.reg;
    const int m1 = c.m - 1;
    const double r_i = TOPK ? 0.0 : rule.rho[i];
    const long long rem0 = TOPK ? rule.tie_take - rule.tie_off[i] : 0;
    long long tie_seen = 0;
    long written = 0;
    const long base_pos = rowptr[i];
    const long end_pos = rowptr[i + 1];   // writes never pass the counted end
    const int last = m1 > 0 ? ((m1 - 1) / 256) * 256 : 0;
    TM v[4];
    double b[4], tt[4];
    rw_load<TM, TOPK>(row, c, rule, 0, m1, v, b, tt);
    for (int s0 = 0; s0 < m1; s0 += 256) {
        TM vn[4];
        double bn[4], tn[4];
        rw_load<TM, TOPK>(row, c, rule, s0 + 256 <= last ? s0 + 256 : last, m1, vn, bn, tn);
        const int cj = s0 + lane * 4;
        if (RAW && i != 0 && s0 != 0)   // wave-uniform
            rw_step<TM, TOPK, RAW>(c, rule, i, cj, m1, ai, r_i, rem0, inv_reg, base_pos, end_pos,
                                   v, b, tt, &tie_seen, &written, colind, val);
        else
            rw_step<TM, TOPK, false>(c, rule, i, cj, m1, ai, r_i, rem0, inv_reg, base_pos,
                                     end_pos, v, b, tt, &tie_seen, &written, colind, val);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            v[q] = vn[q];
            b[q] = bn[q];
            tt[q] = tn[q];
        }
    }
}

// The row writer for a bulk type and rule (RAW for FP32 top-k when the k-th
// value is at least 1.125 * 2^-126: the raw exp2 of an argument below -126,
// whose exact value is subnormal, stays well below it).
template <typename TM>
void launch_row_write_t(const TM* M, long ldm, const ElemCtx& c, const KeepRule& r, int rows,
                        const int* rowptr, int* colind, double* val, hipStream_t s)
{
    const dim3 grid(ceil_div(rows, 4));
    if (r.mode != 1)
        hipLaunchKernelGGL((k_row_write<TM, false, false>), grid, dim3(256), 0, s, M, ldm, c, r,
                           rowptr, colind, val);
    else if (sizeof(TM) == 4 && r.kstar >= 0x00900000ull)
        hipLaunchKernelGGL((k_row_write<TM, true, sizeof(TM) == 4>), grid, dim3(256), 0, s, M,
                           ldm, c, r, rowptr, colind, val);
    else
        hipLaunchKernelGGL((k_row_write<TM, true, false>), grid, dim3(256), 0, s, M, ldm, c, r,
                           rowptr, colind, val);
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
    for (int k = rowptr[r0] + threadIdx.x; k < rowptr[r1]; k += 256) {
        const int j = colind[k];
        if (static_cast<unsigned>(j) < static_cast<unsigned>(m1)) atomicAdd(&mine[j], 1);
    }
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

// k_csc_count with the workgroup's column counts in LDS (m - 1 <= 16384):
// LDS atomics, eight entries per thread in flight, then every count of the
// workgroup is written, zeros included, so cnt needs no clearing first.
__global__ __launch_bounds__(256) void k_csc_count_lds(const int* __restrict__ rowptr,
                                                       const int* __restrict__ colind, int n,
                                                       int rows_per_block, int m1,
                                                       int* __restrict__ cnt)
{
    __shared__ int sc[16384];
    const int t = threadIdx.x;
    int* mine = cnt + static_cast<long>(blockIdx.x) * m1;
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = (r0 + rows_per_block < n) ? r0 + rows_per_block : n;
    for (int j = t; j < m1; j += 256) sc[j] = 0;
    __syncthreads();
    if (r0 < r1) {
        const int k1 = rowptr[r1];
        for (int k = rowptr[r0] + t; k < k1; k += 256 * 8) {
            int jj[8];
#pragma unroll
            for (int u = 0; u < 8; u++) jj[u] = (k + 256 * u < k1) ? colind[k + 256 * u] : -1;
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (static_cast<unsigned>(jj[u]) < static_cast<unsigned>(m1)) atomicAdd(&sc[jj[u]], 1);
        }
    }
    __syncthreads();
    for (int j = t; j < m1; j += 256) mine[j] = sc[j];
}

// k_csc_prefix for nb <= 512 with 64 columns per 1024-thread workgroup: wave w
// holds counts b = w L .. w L + L - 1 (L = ceil(nb / 16) <= 32) of its lane's
// column in registers, the sixteen segment totals are exchanged through LDS,
// and each wave writes its exclusive prefixes.
__global__ __launch_bounds__(1024) void k_csc_prefix16(int* __restrict__ cnt, int nb, int m1,
                                                       int* __restrict__ total)
{
    __shared__ int st[16][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + lane;
    const int len = (nb + 15) / 16;
    const int b0 = w * len;
    int c[32];
    int own = 0;
#pragma unroll
    for (int q = 0; q < 32; q++) {
        const int b = b0 + q;
        c[q] = (q < len && b < nb && j < m1) ? cnt[static_cast<long>(b) * m1 + j] : 0;
    }
#pragma unroll
    for (int q = 0; q < 32; q++) own += c[q];
    st[w][lane] = own;
    __syncthreads();
    int run = 0, tot = 0;
    for (int q = 0; q < 16; q++) {
        const int v = st[q][lane];
        if (q < w) run += v;
        tot += v;
    }
#pragma unroll
    for (int q = 0; q < 32; q++) {
        const int b = b0 + q;
        if (q < len && b < nb && j < m1) cnt[static_cast<long>(b) * m1 + j] = run;
        run += c[q];
    }
    if (w == 0 && j < m1) total[j] = tot;
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
            if (static_cast<unsigned>(j) >= static_cast<unsigned>(m1)) continue;
            const int q = colptr[j] + mine[j];
            mine[j] += 1;
            rowind[q] = i;
            cval[q] = rval[k];
        }
        __syncthreads();
    }
}

// k_csc_scatter with the workgroup's column offsets, row pointers and a
// chunk of its entries in LDS (m - 1 <= kScatterCols, at most kScatterRows
// rows per workgroup): the global loads go out in batches, sixteen column
// offsets or four entries per thread at a time, and the rows, still taken in
// order, depend on each other only through LDS.  Same placement as
// k_csc_scatter: colptr[j] + off[b][j] + the entries of column j in earlier
// rows of the workgroup.  off is not updated.
constexpr int kScatterCols = 16384;
constexpr int kScatterRows = 511;
constexpr int kScatterChunk = 1024;

__global__ __launch_bounds__(256) void k_csc_scatter_lds(const int* __restrict__ rowptr,
                                                         const int* __restrict__ colind,
                                                         const double* __restrict__ rval, int n,
                                                         int rows_per_block, int m1,
                                                         const int* __restrict__ colptr,
                                                         const int* __restrict__ off,
                                                         int* __restrict__ rowind,
                                                         double* __restrict__ cval)
{
    __shared__ int soff[kScatterCols];
    __shared__ int srp[kScatterRows + 1];
    __shared__ int scol[kScatterChunk];
    __shared__ double sval[kScatterChunk];
    const int t = threadIdx.x;
    const int* mine = off + static_cast<long>(blockIdx.x) * m1;
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = (r0 + rows_per_block < n) ? r0 + rows_per_block : n;
    if (r0 >= r1) return;
    const int nr = r1 - r0;
    for (int q = t; q <= nr; q += 256) srp[q] = rowptr[r0 + q];
    for (int j0 = 0; j0 < m1; j0 += 256 * 16) {
        int c[16], o[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int j = j0 + 256 * u + t;
            c[u] = (j < m1) ? colptr[j] : 0;
            o[u] = (j < m1) ? mine[j] : 0;
        }
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int j = j0 + 256 * u + t;
            if (j < m1) soff[j] = c[u] + o[u];
        }
    }
    __syncthreads();
    const int k_end = srp[nr];
    int i = 0;   // the next row to place (the same on every thread)
    for (int c0 = srp[0]; c0 < k_end; c0 += kScatterChunk) {
        const int c1 = (c0 + kScatterChunk < k_end) ? c0 + kScatterChunk : k_end;
        int cc[kScatterChunk / 256];
        double vv[kScatterChunk / 256];
#pragma unroll
        for (int u = 0; u < kScatterChunk / 256; u++) {
            const int k = c0 + 256 * u + t;
            if (k < c1) {
                cc[u] = colind[k];
                vv[u] = rval[k];
            }
        }
#pragma unroll
        for (int u = 0; u < kScatterChunk / 256; u++) {
            const int k = c0 + 256 * u + t;
            if (k < c1) {
                scol[k - c0] = cc[u];
                sval[k - c0] = vv[u];
            }
        }
        __syncthreads();
        // the rows with entries in [c0, c1), in order; a row that runs past c1
        // is finished in the next chunk
        while (i < nr) {
            const int rb = srp[i], re = srp[i + 1];
            if (rb >= c1) break;
            const int b = (rb > c0) ? rb : c0;
            const int e = (re < c1) ? re : c1;
            for (int k = b + t; k < e; k += 256) {
                const int j = scol[k - c0];
                if (static_cast<unsigned>(j) < static_cast<unsigned>(m1)) {
                    const int q = soff[j];
                    soff[j] = q + 1;
                    rowind[q] = r0 + i;
                    cval[q] = sval[k - c0];
                }
            }
            __syncthreads();
            if (re > c1) break;
            i++;
        }
        __syncthreads();   // scol and sval are rewritten by the next chunk
    }
}

// True when row i of the CSR stores column j (binary search; the columns of
// a row are ascending).
__device__ __forceinline__ bool csr_has(const int* __restrict__ rowptr,
                                        const int* __restrict__ colind, int i, int j)
{
    int lo = rowptr[i], hi = rowptr[i + 1];
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const int cm = colind[mid];
        if (cm == j) return true;
        if (cm < j)
            lo = mid + 1;
        else
            hi = mid;
    }
    return false;
}

// CSC straight from M for dense keep sets.  Workgroup (x, y) covers the
// column strip [256 x, 256 x + 256) and the count blocks [y B, y B + B) of
// rows, in tiles of 32 rows: thread t evaluates column 256 x + t on the tile
// (coalesced row segments), keeps a 32-bit mask of its stored rows and their
// values in LDS; a workgroup scan of the per-column counts then lets
// consecutive threads write consecutive entries of each column's run, so the
// stores are coalesced.  Column j's run starts at colptr[j] + off[y B][j].
// The keep rule is the CSR's; a top-k tie is kept exactly when the CSR
// stores it.
template <typename TM>
__global__ __launch_bounds__(256) void k_csc_strip(const TM* __restrict__ M, long ldm, ElemCtx c,
                                                   KeepRule rule, const int* __restrict__ rowptr,
                                                   const int* __restrict__ colind,
                                                   int rows_per_block, int blocks_per_wg,
                                                   const int* __restrict__ off,
                                                   const int* __restrict__ colptr,
                                                   int* __restrict__ rowind,
                                                   double* __restrict__ cval)
{
    using E = Elem<TM>;
    constexpr int RT = 32;
    __shared__ TM tv[RT][256];
    __shared__ unsigned int msk[256];
    __shared__ int base[256];
    __shared__ int posb[256];
    __shared__ int endb[256];
    const int m1 = c.m - 1;
    const int tid = threadIdx.x;
    const int j = blockIdx.x * 256 + tid;
    const int b0 = blockIdx.y * blocks_per_wg;
    const long r0 = static_cast<long>(b0) * rows_per_block;
    const long re = r0 + static_cast<long>(blocks_per_wg) * rows_per_block;
    const int r1 = static_cast<int>(re < c.n ? re : c.n);
    const bool topk = rule.mode == 1;
    const bool col_ok = j < m1;
    int pos = col_ok ? colptr[j] + off[static_cast<long>(b0) * m1 + j] : 0;
    endb[tid] = col_ok ? colptr[j + 1] : 0;
    const double bj = col_ok ? c.beta[j] : 0.0;
    const double tj = (col_ok && !topk) ? rule.tau[j] : 0.0;
    const double inv_reg = 1.0 / c.reg;
    // every thread runs every tile: the barriers below are workgroup-wide
    for (int i0 = static_cast<int>(r0); i0 < r1; i0 += RT) {
        unsigned int mask = 0u;
        if (col_ok) {
#pragma unroll
            for (int u0 = 0; u0 < RT; u0 += 8) {
                TM v[8];
#pragma unroll
                for (int u = 0; u < 8; u++)
                    v[u] = (i0 + u0 + u < r1) ? M[static_cast<long>(i0 + u0 + u) * ldm + j]
                                              : TM(0);
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int i = i0 + u0 + u;
                    if (i >= r1) break;
                    bool z;
                    const TM t = E::val(c.alpha[i] + bj, v[u], c.k, c.reg, z);
                    if (t == TM(0)) continue;
                    bool kept;
                    if (topk) {
                        const unsigned long long key = static_cast<unsigned long long>(E::key(t));
                        kept = key > rule.kstar || (rule.preserve && (i == 0 || j == 0));
                        if (!kept && key == rule.kstar) kept = csr_has(rowptr, colind, i, j);
                    } else {
                        kept = static_cast<double>(t) >= fmin(tj, rule.rho[i]);
                    }
                    if (kept) {
                        mask |= 1u << (u0 + u);
                        tv[u0 + u][tid] = t;
                    }
                }
            }
        }
        const int cnt = __popc(mask);
        int tot;
        const int ex = block_excl_scan_int(cnt, &tot);
        msk[tid] = mask;
        base[tid] = ex;
        posb[tid] = pos;
        __syncthreads();
        for (int e = tid; e < tot; e += 256) {
            // the column: the last one whose run starts at or before e
            int lo = 0, hi = 255;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (base[mid] <= e)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const int k = e - base[lo];
            unsigned int mm = msk[lo];
            for (int q = 0; q < k; q++) mm &= mm - 1u;
            const int rho = __ffs(mm) - 1;
            const int q = posb[lo] + k;
            if (q < endb[lo]) {
                rowind[q] = i0 + rho;
                cval[q] = static_cast<double>(tv[rho][lo]) * inv_reg;
            }
        }
        pos += cnt;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Scans
// ---------------------------------------------------------------------------

// One 1024-thread workgroup, in passes of 16384 counts: thread t holds counts
// 16t .. 16t + 15 of the pass in registers (sixteen loads in flight), the
// thread totals are scanned across the workgroup, and each thread writes its
// sixteen exclusive prefixes from the registers.  out[count] = the total for
// the int form (a CSR pointer array).
template <typename TO>
__global__ __launch_bounds__(1024) void k_scan(const int* __restrict__ cnt, int count,
                                               TO* __restrict__ out, long long* total)
{
    constexpr int kPer = 16;
    __shared__ long long wt[16];
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    long long carry = 0;
    for (int base = 0; base < count; base += 1024 * kPer) {
        const int b = base + kPer * t;
        int c[kPer];
#pragma unroll
        for (int q = 0; q < kPer; q++) c[q] = (b + q < count) ? cnt[b + q] : 0;
        long long own = 0;
#pragma unroll
        for (int q = 0; q < kPer; q++) own += c[q];
        // inclusive wave scan of the thread totals
        long long inc = own;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const long long u = __shfl_up(inc, o);
            if (lane >= o) inc += u;
        }
        if (lane == 63) wt[wave] = inc;
        __syncthreads();
        long long off = carry, pass = 0;
        for (int w = 0; w < 16; w++) {
            if (w < wave) off += wt[w];
            pass += wt[w];
        }
        long long run = off + inc - own;
#pragma unroll
        for (int q = 0; q < kPer; q++) {
            if (b + q < count) out[b + q] = static_cast<TO>(run);
            run += c[q];
        }
        carry += pass;
        __syncthreads();   // wt is rewritten by the next pass
    }
    if (t == 0) {
        if (total) *total = carry;
        if (sizeof(TO) == sizeof(int)) out[count] = static_cast<TO>(carry);
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

void threshold_tau_rows(const void* M, long ldm, const ElemCtx& c, double delta, int find_small,
                        SparseHessian& H, double* tau, hipStream_t s)
{
    const int m1 = c.m - 1, n = c.n;
    if (m1 <= 0) return;
    const int strips = ceil_div(m1, TC);
    // about 2048 workgroups, row blocks of at least 256 rows
    int nrb = std::max(1, std::min(ceil_div(2048, strips), ceil_div(n, 256)));
    const int rpb = ceil_div(n, nrb);
    nrb = ceil_div(n, rpb);
    const size_t ph = static_cast<size_t>(nrb) * TB * m1;
    H.tau_pm.ensure(ph);
    H.tau_pc.ensure(ph);
    H.tau_px.ensure(static_cast<size_t>(nrb) * m1);
    H.tau_bkt.ensure(m1);
    H.tau_base.ensure(m1);
    H.tau_cnt.ensure(m1 + 1);
    H.tau_off.ensure(m1 + 1);
    H.ltmp.ensure(4);
    H.hll.ensure(4);
    const float* Mf = static_cast<const float*>(M);
    // the biased FP32 exponent of delta / n: the binade below which no tau_j lies
    int elo = 0;
    {
        const float lo = static_cast<float>(delta / n);
        unsigned int bits;
        std::memcpy(&bits, &lo, sizeof bits);
        elo = (std::isfinite(lo) && lo > 0.f) ? static_cast<int>((bits >> 23) & 255u) : 0;
    }
    hipLaunchKernelGGL(k_tau_hist, dim3(strips, nrb), dim3(256), 0, s, Mf, ldm, c, elo, rpb,
                       H.tau_pm.p, H.tau_pc.p, H.tau_px.p);
    hipLaunchKernelGGL(k_tau_fold, dim3(ceil_div(m1, 4)), dim3(256), 0, s, H.tau_pm.p,
                       H.tau_pc.p, H.tau_px.p, nrb, m1, delta, find_small, H.tau_bkt.p,
                       H.tau_base.p, H.tau_cnt.p, tau);
    KERNEL_CHECK();
    launch_scan(H.tau_cnt, m1, H.tau_off, H.ltmp.p + 3, s);
    HIP_CHECK(hipMemcpyAsync(H.hll.p + 3, H.ltmp.p + 3, sizeof(long long), hipMemcpyDeviceToHost,
                             s));
    HIP_CHECK(hipStreamSynchronize(s));
    const long long total = H.hll.p[3];
    if (total > 0x7fffffffLL)
        fail(E_INVALID, "threshold rule: %lld boundary candidates (> 2^31)", total);
    if (total > 0) {
        H.tau_cand.ensure(static_cast<size_t>(total));
        hipLaunchKernelGGL(k_tau_gather, dim3(strips, nrb), dim3(256), 0, s, Mf, ldm, c, elo,
                           rpb,
                           H.tau_bkt.p, H.tau_off.p, H.tau_pc.p, H.tau_cand.p);
        hipLaunchKernelGGL(k_tau_select, dim3(m1), dim3(256), 0, s, H.tau_bkt.p, H.tau_base.p,
                           H.tau_cnt.p, H.tau_off.p, H.tau_cand.p, delta, tau);
        KERNEL_CHECK();
    }
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

void topk_select_rows(int bulk, const void* M, long ldm, const ElemCtx& c, long long k,
                      int preserve, SparseHessian& H, unsigned long long* kstar,
                      long long* n_greater, int* tcnt, int* gcnt, hipStream_t s)
{
    const int KB = bulk == 0 ? 32 : 64;
    const int npass = (KB + RD - 1) / RD;
    const int rows_per_block = std::max(1, ceil_div(c.n, 2048));
    const int nblocks = ceil_div(c.n, rows_per_block);
    H.hist.ensure(RB);
    H.hll.ensure(RB);
    // pass over M 1: the top digit
    // digit p covers key bits [KB - p RD - w, KB - p RD), w = min(RD, KB - p RD)
    auto host_digit = [KB](int p, int* w) {
        const int hi = KB - p * RD;
        *w = hi < RD ? hi : RD;
        return hi - *w;
    };
    int width;
    const int shift1 = host_digit(0, &width);
    HIP_CHECK(hipMemsetAsync(H.hist.p, 0, RB * sizeof(unsigned long long), s));
    if (bulk == 0)
        hipLaunchKernelGGL(k_topk_hist<float>, dim3(nblocks), dim3(256), 0, s, cM<float>(M), ldm,
                           c, rows_per_block, shift1, width, 0ull, 1, H.hist.p);
    else
        hipLaunchKernelGGL(k_topk_hist<double>, dim3(nblocks), dim3(256), 0, s, cM<double>(M),
                           ldm, c, rows_per_block, shift1, width, 0ull, 1, H.hist.p);
    KERNEL_CHECK();
    HIP_CHECK(hipMemcpyAsync(H.hll.p, H.hist.p, RB * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    long long above = 0;
    int b = (1 << width) - 1;
    for (; b >= 0; b--) {
        const long long h = H.hll.p[b];
        if (above + h >= k) break;
        above += h;
    }
    if (b < 0) fail(E_INTERNAL, "top-k radix select lost the k-th entry");
    const long long nc = H.hll.p[b];   // entries of the boundary bucket
    // pass over M 2: the row tallies outside the bucket and its entries
    H.topk_cand.ensure(static_cast<size_t>(nc));
    H.topk_ncand.ensure(1);
    HIP_CHECK(hipMemsetAsync(H.topk_ncand.p, 0, sizeof(unsigned long long), s));
    if (bulk == 0)
        hipLaunchKernelGGL(k_topk_gather<float>, dim3(c.n), dim3(256), 0, s, cM<float>(M), ldm, c,
                           shift1, static_cast<unsigned long long>(b), preserve, tcnt, gcnt,
                           H.topk_cand.p, H.topk_ncand.p, nc);
    else
        hipLaunchKernelGGL(k_topk_gather<double>, dim3(c.n), dim3(256), 0, s, cM<double>(M), ldm,
                           c, shift1, static_cast<unsigned long long>(b), preserve, tcnt, gcnt,
                           H.topk_cand.p, H.topk_ncand.p, nc);
    KERNEL_CHECK();
    // the remaining digits over the candidates
    unsigned long long prefix = static_cast<unsigned long long>(b);
    const int nb = std::max(1, std::min(1024, ceil_div(nc, 2048)));
    for (int p = 1; p < npass; p++) {
        int w;
        const int shift = host_digit(p, &w);
        HIP_CHECK(hipMemsetAsync(H.hist.p, 0, RB * sizeof(unsigned long long), s));
        hipLaunchKernelGGL(k_cand_hist, dim3(nb), dim3(256), 0, s, H.topk_cand.p, nc,
                           H.topk_ncand.p, shift, w, prefix, H.hist.p);
        KERNEL_CHECK();
        HIP_CHECK(hipMemcpyAsync(H.hll.p, H.hist.p, RB * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        int bb = (1 << w) - 1;
        for (; bb >= 0; bb--) {
            const long long h = H.hll.p[bb];
            if (above + h >= k) break;
            above += h;
        }
        if (bb < 0) fail(E_INTERNAL, "top-k radix select lost the k-th entry");
        prefix = (prefix << w) | static_cast<unsigned long long>(bb);
    }
    // the candidates' share of the tallies
    if (nc > 0)
        hipLaunchKernelGGL(k_cand_classify, dim3(ceil_div(nc, 256)), dim3(256), 0, s,
                           H.topk_cand.p, nc, H.topk_ncand.p, prefix, preserve, c.n, tcnt, gcnt);
    KERNEL_CHECK();
    *kstar = prefix;
    *n_greater = above;
}

void launch_row_tally(int bulk, const void* M, long ldm, const ElemCtx& c,
                      unsigned long long kstar, int preserve, int* tcnt, int* gcnt,
                      hipStream_t s)
{
    if (bulk == 0)
        hipLaunchKernelGGL(k_row_tally<float>, dim3(c.n), dim3(256), 0, s, cM<float>(M), ldm, c,
                           kstar, preserve, tcnt, gcnt);
    else
        hipLaunchKernelGGL(k_row_tally<double>, dim3(c.n), dim3(256), 0, s, cM<double>(M), ldm,
                           c, kstar, preserve, tcnt, gcnt);
    KERNEL_CHECK();
}

void launch_topk_count(const int* tcnt, const int* gcnt, const KeepRule& r, int n, int* cnt,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_topk_count, dim3(ceil_div(n, 256)), dim3(256), 0, s, tcnt, gcnt,
                       r.tie_off, r.tie_take, r.kstar, r.preserve, n, cnt);
    KERNEL_CHECK();
}

void launch_row_write(int bulk, const void* M, long ldm, const ElemCtx& c, const KeepRule& r,
                      const int* rowptr, int* colind, double* val, hipStream_t s)
{
    if (bulk == 0)
        launch_row_write_t(cM<float>(M), ldm, c, r, c.n, rowptr, colind, val, s);
    else
        launch_row_write_t(cM<double>(M), ldm, c, r, c.n, rowptr, colind, val, s);
    KERNEL_CHECK();
}

void launch_build_csc(int bulk, const void* M, long ldm, const void* MT, long ldt,
                      const ElemCtx& c, const KeepRule& r, const SparseHessian& H, int how,
                      int* work, int rows_per_block, int* colcnt, int* colptr, long long* total,
                      int* rowind, double* cval, hipStream_t s)
{
    const int n = H.n, m1 = H.m - 1;
    if (m1 <= 0) return;
    const int nb = ceil_div(n, rows_per_block);
    if (m1 <= 16384) {
        hipLaunchKernelGGL(k_csc_count_lds, dim3(nb), dim3(256), 0, s, H.rowptr.p, H.colind.p, n,
                           rows_per_block, m1, work);
    } else {
        HIP_CHECK(hipMemsetAsync(work, 0, sizeof(int) * static_cast<size_t>(nb) * m1, s));
        hipLaunchKernelGGL(k_csc_count, dim3(nb), dim3(256), 0, s, H.rowptr.p, H.colind.p, n,
                           rows_per_block, m1, work);
    }
    if (nb <= 512)
        hipLaunchKernelGGL(k_csc_prefix16, dim3(ceil_div(m1, 64)), dim3(1024), 0, s, work, nb,
                           m1, colcnt);
    else
        hipLaunchKernelGGL(k_csc_prefix, dim3(ceil_div(m1, 256)), dim3(256), 0, s, work, nb, m1,
                           colcnt);
    KERNEL_CHECK();
    launch_scan(colcnt, m1, colptr, total, s);
    if (how == 0) {
        if (m1 <= kScatterCols && rows_per_block <= kScatterRows)
            hipLaunchKernelGGL(k_csc_scatter_lds, dim3(nb), dim3(256), 0, s, H.rowptr.p,
                               H.colind.p, H.rval.p, n, rows_per_block, m1, colptr, work, rowind,
                               cval);
        else
            hipLaunchKernelGGL(k_csc_scatter, dim3(nb), dim3(256), 0, s, H.rowptr.p, H.colind.p,
                               H.rval.p, n, rows_per_block, m1, colptr, work, rowind, cval);
    } else if (how == 2) {
        // threshold rule on M': row j of M' is column j of M; the rule is
        // symmetric in (tau_j, rho_i) and alpha_i + beta_j is exact either way
        ElemCtx ct = c;
        ct.alpha = c.beta;
        ct.beta = c.alpha;
        ct.n = m1;
        ct.m = n + 1;
        KeepRule rt = r;
        rt.tau = r.rho;
        rt.rho = r.tau;
        if (bulk == 0)
            launch_row_write_t(cM<float>(MT), ldt, ct, rt, m1, colptr, rowind, cval, s);
        else
            launch_row_write_t(cM<double>(MT), ldt, ct, rt, m1, colptr, rowind, cval, s);
    } else {
        // about 512 rows per workgroup: long front-to-back runs per column
        const int bpw = std::max(1, 512 / rows_per_block);
        const dim3 grid(ceil_div(m1, 256), ceil_div(nb, bpw));
        if (bulk == 0)
            hipLaunchKernelGGL(k_csc_strip<float>, grid, dim3(256), 0, s, cM<float>(M), ldm, c, r,
                               H.rowptr.p, H.colind.p, rows_per_block, bpw, work, colptr, rowind,
                               cval);
        else
            hipLaunchKernelGGL(k_csc_strip<double>, grid, dim3(256), 0, s, cM<double>(M), ldm, c,
                               r, H.rowptr.p, H.colind.p, rows_per_block, bpw, work, colptr,
                               rowind, cval);
    }
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
