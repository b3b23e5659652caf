// dense_device.hpp -- device building blocks of the streaming passes over the
// cost matrix M (gfx950), shared by the single-problem kernels and the
// batched kernels.
//
// Every pass evaluates x_ij = (p_i + q_j - M_ij) / reg in FP64 (M is FP32 data
// upcast exactly, or FP64 in FP64 mode), converts the shifted value to the
// bulk type, and accumulates 2^(x - shift) with a per-lane adaptive shift:
//
//   * the shift starts at a caller-supplied guess (log2 of the expected sum),
//   * it moves down once if the first values all sit more than 8 below it,
//   * it moves up whenever a value exceeds it by more than 8,
//
// so the dominant terms are always formed within a few units of zero, where
// the FP32 rounding of the exponent argument is ~1e-7 relative, and nothing
// overflows or underflows whatever the guess.  In steady state (good guess)
// the shift never moves and the inner loop is: 4 FP64-rate ops (cvt, add,
// fma, cvt), one max, one exp2 and one add per element; every 16 terms the
// FP32 lane sum is moved into an FP64 total.  M is read with nontemporal
// loads (it is streamed once per pass).
//
// Partial results combine as FP64 log2-sums: log2(sum) = L_max + log2(sum
// 2^(L - L_max)).  All reductions run in a fixed order, so results are
// bitwise reproducible run to run.
#pragma once

#include <cmath>

#include "common.hpp"

namespace regot {
namespace dev {

constexpr double kLog2e = 1.4426950408889634074;
constexpr double kLn2 = 0.69314718055994530942;

template <typename T>
struct VecOf;
template <>
struct VecOf<float> {
    using type = float4;
    static constexpr int N = 4;
};
template <>
struct VecOf<double> {
    using type = double2;
    static constexpr int N = 2;
};

// v_exp_f32 directly: exp2f adds a denormal-range rescale (compare, two
// selects, add, ldexp) around it.  Results below 2^-126 flush to 0; with the
// adaptive shift every summed term is relative to a running maximum >= 2^-8,
// so a flushed term is below 2^-118 of the sum.
__device__ __forceinline__ float ex2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ double ex2(double x) { return exp2(x); }
__device__ __forceinline__ float vmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double vmax(double a, double b) { return fmax(a, b); }

template <typename T>
__device__ __forceinline__ T ninf()
{
    return -static_cast<T>(INFINITY);
}

template <typename VT, typename T>
__device__ __forceinline__ T vget(const VT& v, int i);
template <>
__device__ __forceinline__ float vget<float4, float>(const float4& v, int i)
{
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}
template <>
__device__ __forceinline__ double vget<double2, double>(const double2& v, int i)
{
    return i == 0 ? v.x : v.y;
}

// Streaming load of one vector of M.  Nontemporal: M is read once per pass
// and far larger than L2 and MALL, so it should not displace the potentials;
// measured on a 1 GiB row stream this is ~6% faster than a default load.
__device__ __forceinline__ float4 ld_stream(const float* p)
{
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 v = __builtin_nontemporal_load(reinterpret_cast<const f4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ double2 ld_stream(const double* p)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 v = __builtin_nontemporal_load(reinterpret_cast<const d2*>(p));
    return make_double2(v.x, v.y);
}

// Load of one vector of M: nontemporal (nt) for matrices far larger than the
// last-level cache, which are streamed once per pass; cached otherwise, so a
// matrix that fits the 256 MB Infinity Cache stays there between passes.
template <typename T>
__device__ __forceinline__ typename VecOf<T>::type ld_m(const T* p, bool nt)
{
    using VT = typename VecOf<T>::type;
    return nt ? ld_stream(p) : *reinterpret_cast<const VT*>(p);
}

// One cross-lane step of an FP64 value by a DPP move of each half.  Every
// control used here (quad permutes, half-row and row mirrors) gives every
// lane a valid source, so with bound_ctrl and full masks the old value is
// dead and the move needs no zeroed destination first.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

// R wave sums together: lane pairs xor 1 and xor 2 (quad permutes), then the
// mirrored halves of each 8 and each 16 lanes (DPP), then xor 16 and xor 32
// (shuffles issued for all R values before their adds).  Each step adds two
// values that are equal on every lane of a group in either order, so every
// lane ends with the same bits.
template <int R>
__device__ __forceinline__ void wave_sums(double (&x)[R])
{
#pragma unroll
    for (int r = 0; r < R; r++) x[r] += dpp_f64<0xB1>(x[r]);    // quad_perm [1,0,3,2]
#pragma unroll
    for (int r = 0; r < R; r++) x[r] += dpp_f64<0x4E>(x[r]);    // quad_perm [2,3,0,1]
#pragma unroll
    for (int r = 0; r < R; r++) x[r] += dpp_f64<0x141>(x[r]);   // row_half_mirror
#pragma unroll
    for (int r = 0; r < R; r++) x[r] += dpp_f64<0x140>(x[r]);   // row_mirror
    double y[R];
#pragma unroll
    for (int r = 0; r < R; r++) y[r] = __shfl_xor(x[r], 16);
#pragma unroll
    for (int r = 0; r < R; r++) x[r] += y[r];
#pragma unroll
    for (int r = 0; r < R; r++) y[r] = __shfl_xor(x[r], 32);
#pragma unroll
    for (int r = 0; r < R; r++) x[r] += y[r];
}

// Sum over each group of L consecutive lanes (L a power of two <= 64), the
// same bits on every lane of the group: DPP steps inside rows of 16 lanes,
// shuffles across rows.
template <int L>
__device__ __forceinline__ double group_sum(double x)
{
    if (L >= 2) x += dpp_f64<0xB1>(x);    // quad_perm [1,0,3,2]
    if (L >= 4) x += dpp_f64<0x4E>(x);    // quad_perm [2,3,0,1]
    if (L >= 8) x += dpp_f64<0x141>(x);   // row_half_mirror
    if (L >= 16) x += dpp_f64<0x140>(x);  // row_mirror
    if (L >= 32) x += __shfl_xor(x, 16);
    if (L >= 64) x += __shfl_xor(x, 32);
    return x;
}

// Wave maximum (any order gives the same value) and wave sum (fixed order;
// the same bits on every lane) by the same steps.
__device__ __forceinline__ double wave_max(double v)
{
    v = fmax(v, dpp_f64<0xB1>(v));
    v = fmax(v, dpp_f64<0x4E>(v));
    v = fmax(v, dpp_f64<0x141>(v));
    v = fmax(v, dpp_f64<0x140>(v));
    v = fmax(v, __shfl_xor(v, 16));
    return fmax(v, __shfl_xor(v, 32));
}

__device__ __forceinline__ double wave_sum0(double v) { return group_sum<64>(v); }

// Integer sum over the wave, the same on every lane.
__device__ __forceinline__ int wave_isum(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);
    v += __shfl_xor(v, 16);
    return v + __shfl_xor(v, 32);
}

__device__ __forceinline__ double acc_log2(double sh, double s)
{
    return (s > 0.0) ? sh + log2(s) : -INFINITY;
}

// Each lane sums 2^t in the bulk type TA into s and, in FP32, moves s into an
// FP64 total S every kFlush terms, so FP32 rounding accumulates over 16 terms
// at most (<= 1e-6 relative, ~2e-7 typical) whatever the row or column length.
constexpr int kFlushSteps = 4;   // steps of 4 terms between FP32 -> FP64 flushes

template <typename TA>
__device__ __forceinline__ void flush(TA& s, double& S)
{
    if (sizeof(TA) == 4) {
        S += static_cast<double>(s);
        s = TA(0);
    }
}

// Move the shift when the running max leaves [-8, 8]: values up to 2^8 above
// the shift keep the bulk-type rounding of the exponent argument below
// 2^-16 ulp-units (~2e-7 relative), and a lane moves at most once downwards
// (first values) and rarely upwards.  The partial sums are folded into S and
// rescaled by 2^-d in FP64.  Returns true if moved.
template <typename TA>
__device__ __forceinline__ bool reshift(double& sh, TA& s, double& S, TA& rm)
{
    if (rm > TA(8) || (rm < TA(-8) && rm > ninf<TA>())) {
        const TA d = rm;
        sh += static_cast<double>(d);
        const double tot = S + static_cast<double>(s);
        // tot holds terms <= 2^d, so tot * 2^-d is representable; the factor
        // alone needs FP64 when d < -126.  FP32 bulk takes the FP32 exp2 of the
        // factor (one rounding, like one more term) unless it would overflow.
        double f;
        if (sizeof(TA) == 4 && d > TA(-120))
            f = static_cast<double>(ex2(static_cast<float>(-d)));
        else
            f = exp2(-static_cast<double>(d));
        S = (tot != 0.0) ? tot * f : 0.0;
        s = TA(0);
        rm = TA(0);
        return true;
    }
    return false;
}

// Wave-uniform variant for a row shared by the 64 lanes of a wave: the shift
// sh is the same on every lane, and moves (by the wave maximum of rm) when
// any lane's running maximum exceeds it by more than 8, or every lane's sits
// more than 8 below it.  The lanes' totals then merge by a plain sum.  Lanes
// far below the wave maximum lose nothing that matters: a term 2^-126 below
// the maximum is below the FP32 resolution of the sum.  Every lane of the
// wave must call it (wave-wide ballot and shuffles).
template <typename TA>
__device__ __forceinline__ bool reshift_wave(double& sh, TA& s, double& S, TA& rm)
{
    const bool up = __ballot(rm > TA(8)) != 0;
    const bool down = __ballot(rm >= TA(-8)) == 0;
    if (!(up || down)) return false;
    const double d = wave_max(static_cast<double>(rm));
    if (!(d > 8.0 || (d < -8.0 && d > -INFINITY))) return false;
    sh += d;
    const double tot = S + static_cast<double>(s);
    S = (tot != 0.0) ? tot * exp2(-d) : 0.0;
    s = TA(0);
    rm = static_cast<TA>(static_cast<double>(rm) - d);
    return true;
}

// ---------------------------------------------------------------------------
// Row and column accumulation
// ---------------------------------------------------------------------------

// One row x V columns: t_v = (bk_v - sh) - k M_v, masked entries = -inf, with
// the wave-uniform shift of reshift_wave (every lane of the wave calls it).
// TM is the storage type of M, TA the type t, exp2 and the sums run in.
template <typename TM, typename TA, int V, bool MASK>
__device__ __forceinline__ void row_accumulate(double& sh, TA& s, double& S, TA& rm,
                                               const double* bk,
                                               const typename VecOf<TM>::type& mv, double k,
                                               int valid_cols)
{
    TA t[V];
    TA cm = ninf<TA>();
#pragma unroll
    for (int v = 0; v < V; v++) {
        const double mvd = static_cast<double>(vget<typename VecOf<TM>::type, TM>(mv, v));
        t[v] = static_cast<TA>(fma(-k, mvd, bk[v] - sh));
        if (MASK && v >= valid_cols) t[v] = ninf<TA>();
        cm = vmax(cm, t[v]);
    }
    rm = vmax(rm, cm);
    if (reshift_wave(sh, s, S, rm)) {
#pragma unroll
        for (int v = 0; v < V; v++) {
            const double mvd = static_cast<double>(vget<typename VecOf<TM>::type, TM>(mv, v));
            t[v] = static_cast<TA>(fma(-k, mvd, bk[v] - sh));
            if (MASK && v >= valid_cols) t[v] = ninf<TA>();
        }
    }
#pragma unroll
    for (int v = 0; v < V; v++) s += ex2(t[v]);
}

// Sums of NV values over the 256 threads of the workgroup: butterfly within
// each wave, then the four wave totals in fixed order.  Every thread gets the
// same (bitwise) results; one barrier pair per call.
template <int NV>
__device__ __forceinline__ void block_sums_256(double (&v)[NV])
{
    __shared__ double sm[4][NV];
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = group_sum<64>(v[q]);
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) sm[wave][q] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = ((sm[0][q] + sm[1][q]) + sm[2][q]) + sm[3][q];
}

template <typename TM, typename TA, int V, int CH, bool MASK>
__device__ __forceinline__ void col_chunk(double* sh, TA* s, double* S, TA* rm,
                                          const typename VecOf<TM>::type* mv, const double* ak,
                                          double k, int valid_rows, int valid_cols)
{
#pragma unroll
    for (int v = 0; v < V; v++) {
        TA t[CH];
        TA cm = ninf<TA>();
#pragma unroll
        for (int c = 0; c < CH; c++) {
            const double mvd = static_cast<double>(vget<typename VecOf<TM>::type, TM>(mv[c], v));
            t[c] = static_cast<TA>(fma(-k, mvd, ak[c] - sh[v]));
            if (MASK && (c >= valid_rows || v >= valid_cols)) t[c] = ninf<TA>();
            cm = vmax(cm, t[c]);
        }
        rm[v] = vmax(rm[v], cm);
        if (reshift(sh[v], s[v], S[v], rm[v])) {
#pragma unroll
            for (int c = 0; c < CH; c++) {
                const double mvd =
                    static_cast<double>(vget<typename VecOf<TM>::type, TM>(mv[c], v));
                t[c] = static_cast<TA>(fma(-k, mvd, ak[c] - sh[v]));
                if (MASK && (c >= valid_rows || v >= valid_cols)) t[c] = ninf<TA>();
            }
        }
#pragma unroll
        for (int c = 0; c < CH; c++) s[v] += ex2(t[c]);
    }
}

// ---------------------------------------------------------------------------
// Whole-row and column-tile sweeps
// ---------------------------------------------------------------------------

// One wave, R consecutive rows row0.. of an n x m matrix, columns
// [c_begin, c_end) (c_begin a multiple of 64V): L[r] = log2 sum_j
// 2^(k (beta_j - M_ij)) (-inf past n or for an empty range; the same value on
// every lane).  Lane l reads the V-wide vector at column c_begin + c*64V + l*V
// of each row (1 KiB per row per step, fully coalesced); the beta chunk is
// loaded once per step and reused by all R rows.  The shift guess is
// log2 a_i - k rowpot_i (loga or rowpot may be null).
template <typename TM, typename TA, int R>
__device__ __forceinline__ void row_lse(const TM* __restrict__ M, long ldm, int n, int c_begin,
                                        int c_end, const double* __restrict__ beta, double k,
                                        const double* __restrict__ loga,
                                        const double* __restrict__ rowpot, long row0, int lane,
                                        double (&L)[R], bool nt)
{
    using VT = typename VecOf<TM>::type;
    constexpr int V = VecOf<TM>::N;
    const TM* rowp[R];
#pragma unroll
    for (int r = 0; r < R; r++) rowp[r] = M + (row0 + r < n ? row0 + r : row0) * ldm;

    const int step = 64 * V;
    const int nfull = c_end > c_begin ? (c_end - c_begin) / step : 0;
    // Rows past n read row0 again (in bounds) and are not accumulated.  The
    // next step's vectors of M and beta are loaded before the current ones are
    // consumed, so each wave keeps its loads in flight while it computes; the
    // first step's loads go out before the shift guesses are read.
    VT mv[R];
    double bc[V];
    if (nfull > 0) {
#pragma unroll
        for (int r = 0; r < R; r++) mv[r] = ld_m(rowp[r] + c_begin + lane * V, nt);
#pragma unroll
        for (int v = 0; v < V; v++) bc[v] = beta[c_begin + lane * V + v];
    }

    double sh[R], S[R];
    TA s[R], rm[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const long i = row0 + r;
        double g = 0.0;
        if (i < n && loga != nullptr) {
            g = loga[i] * kLog2e;
            if (rowpot != nullptr) g -= rowpot[i] * k;
        }
        sh[r] = g;
        s[r] = TA(0);
        S[r] = 0.0;
        rm[r] = ninf<TA>();
    }

    for (int c = 0; c < nfull; c++) {
        const int j0 = c_begin + c * step + lane * V;
        VT nx[R];
        double bn[V];
        if (c + 1 < nfull) {
#pragma unroll
            for (int r = 0; r < R; r++) nx[r] = ld_m(rowp[r] + j0 + step, nt);
#pragma unroll
            for (int v = 0; v < V; v++) bn[v] = beta[j0 + step + v];
        }
        double bk[V];
#pragma unroll
        for (int v = 0; v < V; v++) bk[v] = bc[v] * k;
#pragma unroll
        for (int r = 0; r < R; r++)
            if (row0 + r < n)
                row_accumulate<TM, TA, V, false>(sh[r], s[r], S[r], rm[r], bk, mv[r], k, V);
        if (c % kFlushSteps == kFlushSteps - 1) {
#pragma unroll
            for (int r = 0; r < R; r++) flush(s[r], S[r]);
        }
#pragma unroll
        for (int r = 0; r < R; r++) mv[r] = nx[r];
#pragma unroll
        for (int v = 0; v < V; v++) bc[v] = bn[v];
    }
    {
        // partial last step: every lane takes part (the shift is wave-wide);
        // lanes past c_end load an in-bounds vector and mask all of it
        const int jt = c_begin + nfull * step;
        if (jt < c_end) {
            const int j0 = jt + lane * V;
            int valid = c_end - j0;
            valid = valid < 0 ? 0 : (valid > V ? V : valid);
            const int jl = valid > 0 ? j0 : jt;
            double bk[V];
#pragma unroll
            for (int v = 0; v < V; v++) bk[v] = (v < valid) ? beta[j0 + v] * k : 0.0;
#pragma unroll
            for (int r = 0; r < R; r++) {
                if (row0 + r < n) {
                    // jl is a multiple of V below m <= pitch (a multiple of V)
                    const VT mv1 = ld_m(rowp[r] + jl, nt);
                    row_accumulate<TM, TA, V, true>(sh[r], s[r], S[r], rm[r], bk, mv1, k, valid);
                }
            }
        }
    }
    // the shift is wave-uniform: the lanes' totals add up directly
#pragma unroll
    for (int r = 0; r < R; r++) {
        const double tot = wave_sum0(S[r] + static_cast<double>(s[r]));
        L[r] = (row0 + r < n && tot > 0.0) ? sh[r] + log2(tot) : -INFINITY;
    }
}

// One workgroup (4 waves) over one column tile of 64V columns starting at
// tile_col0 and rows [rb, re): per column j, part_row[j] = log2 sum_i
// 2^(k (alpha_i - M_ij)) over those rows.  Lane l owns the V adjacent columns
// tile_col0 + l*V (one 16-byte load per row, a 1 KiB coalesced row segment
// per wave); the waves take CH-row slabs round robin and merge through LDS.
// The shift guess is log2 b_j - k colpot_j (logb or colpot may be null).
template <typename TM, typename TA>
__device__ __forceinline__ void col_partial_tile(const TM* __restrict__ M, long ldm, int m,
                                                 const double* __restrict__ alpha,
                                                 const double* __restrict__ colpot,
                                                 const double* __restrict__ logb, double k,
                                                 int tile_col0, long rb, long re,
                                                 double* __restrict__ part_row, bool nt)
{
    using VT = typename VecOf<TM>::type;
    constexpr int V = VecOf<TM>::N;
    constexpr int CH = 4;
    constexpr int TW = 64 * V;
    __shared__ double red[4][TW];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int j0 = tile_col0 + lane * V;
    const int valid_cols = (m - j0) < V ? (m - j0) : V;   // <= 0: lane has no column

    double sh[V], S[V];
    TA s[V], rm[V];
#pragma unroll
    for (int v = 0; v < V; v++) {
        const int j = j0 + v;
        double g = 0.0;
        if (v < valid_cols && logb != nullptr) {
            g = logb[j] * kLog2e;
            if (colpot != nullptr) g -= colpot[j] * k;
        }
        sh[v] = g;
        s[v] = TA(0);
        S[v] = 0.0;
        rm[v] = ninf<TA>();
    }

    if (valid_cols > 0) {
        const TM* base = M + j0;
        long r = rb + wave * CH;
        for (int q = 0; r + CH <= re; r += 4 * CH, q++) {
            VT mv[CH];
            double ak[CH];
#pragma unroll
            for (int c = 0; c < CH; c++) {
                mv[c] = ld_m(base + (r + c) * ldm, nt);
                ak[c] = alpha[r + c] * k;
            }
            if (valid_cols == V)
                col_chunk<TM, TA, V, CH, false>(sh, s, S, rm, mv, ak, k, CH, V);
            else
                col_chunk<TM, TA, V, CH, true>(sh, s, S, rm, mv, ak, k, CH, valid_cols);
            if (q % kFlushSteps == kFlushSteps - 1) {
#pragma unroll
                for (int v = 0; v < V; v++) flush(s[v], S[v]);
            }
        }
        if (r < re) {
            const int vr = static_cast<int>(re - r);
            VT mv[CH];
            double ak[CH];
#pragma unroll
            for (int c = 0; c < CH; c++) {
                const long rr = (c < vr) ? r + c : r;
                mv[c] = ld_m(base + rr * ldm, nt);
                ak[c] = alpha[rr] * k;
            }
            col_chunk<TM, TA, V, CH, true>(sh, s, S, rm, mv, ak, k, vr, valid_cols);
        }
    }

#pragma unroll
    for (int v = 0; v < V; v++)
        red[wave][lane * V + v] = acc_log2(sh[v], S[v] + static_cast<double>(s[v]));
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int v = 0; v < V; v++) {
            if (v >= valid_cols) break;
            double Lm = -INFINITY;
#pragma unroll
            for (int w = 0; w < 4; w++) Lm = fmax(Lm, red[w][lane * V + v]);
            double out = -INFINITY;
            if (Lm != -INFINITY) {
                double Ssum = 0.0;
#pragma unroll
                for (int w = 0; w < 4; w++) {
                    const double Lw = red[w][lane * V + v];
                    Ssum += (Lw == -INFINITY) ? 0.0 : exp2(Lw - Lm);
                }
                out = Lm + log2(Ssum);
            }
            part_row[j0 + v] = out;
        }
    }
}

// Merge of one column's nblk partial log2 sums (stride m), 4 threads per
// column (sub = threadIdx & 3), fixed order; the result is on all 4 threads.
__device__ __forceinline__ double combine_partials(const double* __restrict__ part, int nblk,
                                                   long stride, int j, bool valid, int sub)
{
    double Lm = -INFINITY;
    if (valid)
        for (int p = sub; p < nblk; p += 4) Lm = fmax(Lm, part[static_cast<long>(p) * stride + j]);
    Lm = fmax(Lm, __shfl_xor(Lm, 1));
    Lm = fmax(Lm, __shfl_xor(Lm, 2));
    double S = 0.0;
    if (valid && Lm != -INFINITY)
        for (int p = sub; p < nblk; p += 4) {
            const double x = part[static_cast<long>(p) * stride + j];
            S += (x == -INFINITY) ? 0.0 : exp2(x - Lm);
        }
    S += __shfl_xor(S, 1);
    S += __shfl_xor(S, 2);
    return (Lm == -INFINITY) ? -INFINITY : Lm + log2(S);
}

}  // namespace dev
}  // namespace regot
