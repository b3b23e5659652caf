// dense_kernels.hip -- streaming passes over the cost matrix M for gfx950.
//
// Every pass evaluates x_ij = (p_i + q_j - M_ij) / reg in FP64 (M is FP32 data
// upcast exactly, or FP64 in FP64 mode), converts the shifted value to the
// bulk type, and accumulates 2^(x - shift) with a per-lane adaptive shift:
//
//   * the shift starts at a caller-supplied guess (log2 of the expected sum),
//   * it moves down once if the first values all sit more than 8 below it,
//   * it moves up whenever a value exceeds it by more than 1,
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
#include "dense.hpp"

#include <cmath>

namespace regot {
namespace {

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

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ double wave_sum0(double v)
{
    // Butterfly sum; lane 0's value is used so the order is fixed.
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return __shfl(v, 0);
}

// log2 of sum over the wave of 2^L (L = -inf for empty lanes).
__device__ __forceinline__ double wave_log2sum(double L)
{
    const double Lmax = wave_max(L);
    if (Lmax == -INFINITY) return -INFINITY;
    const double e = (L == -INFINITY) ? 0.0 : exp2(L - Lmax);
    return Lmax + log2(wave_sum0(e));
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

// ---------------------------------------------------------------------------
// Row pass
// ---------------------------------------------------------------------------

// One row x V columns: t_v = (bk_v - sh) - k M_v, masked entries = -inf.
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
    if (reshift(sh, s, S, rm)) {
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

__device__ __forceinline__ double ld_agent(const double* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sums of NV values over the 256 threads of the workgroup: butterfly within
// each wave, then the four wave totals in fixed order.  Every thread gets the
// same (bitwise) results; one barrier pair per call.
template <int NV>
__device__ __forceinline__ void block_sums_256(double (&v)[NV])
{
    __shared__ double sm[4][NV];
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double x = v[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
        v[q] = __shfl(x, 0);
    }
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) sm[wave][q] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = ((sm[0][q] + sm[1][q]) + sm[2][q]) + sm[3][q];
}

// Each wave owns R consecutive rows and streams them left to right, lane l
// reading the V-wide vector at column c*64V + l*V of each row (1 KiB per row
// per step, fully coalesced).  The beta chunk is loaded once per step and
// reused by all R rows.  With FIN, per-workgroup sums of (g^2, r, alpha a) are
// published and the last workgroup to arrive runs the BCD bookkeeping.
template <typename TM, typename TA, int R, bool FIN>
__global__ __launch_bounds__(256) void k_row_pass(const TM* __restrict__ M, long ldm, int n,
                                                  int m, const double* __restrict__ beta,
                                                  double k, RowEpi epi, BcdState fin,
                                                  const int* __restrict__ ctl)
{
    using VT = typename VecOf<TM>::type;
    constexpr int V = VecOf<TM>::N;
    if (ctl != nullptr && ctl[0] != 0) return;
    const int lane = threadIdx.x & 63;
    // wave-uniform: row tests below become scalar branches, not exec masks
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long row0 = (static_cast<long>(blockIdx.x) * 4 + wave) * R;

    const double* rowpot = epi.rowpot;
    double* aout = epi.alpha_out;
    if (epi.parity) {
        const int kk = ctl[1];
        rowpot = epi.apair[kk & 1];
        aout = epi.apair[(kk + 1) & 1];
    }
    double wg2 = 0.0, wr = 0.0, waa = 0.0;

    if (row0 < n) {
        double sh[R], S[R];
        TA s[R], rm[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const long i = row0 + r;
            double g = 0.0;
            if (i < n && epi.loga != nullptr) {
                g = epi.loga[i] * kLog2e;
                if (rowpot != nullptr) g -= rowpot[i] * k;
            }
            sh[r] = g;
            s[r] = TA(0);
            S[r] = 0.0;
            rm[r] = ninf<TA>();
        }
        const TM* rowp[R];
#pragma unroll
        for (int r = 0; r < R; r++) rowp[r] = M + (row0 + r < n ? row0 + r : row0) * ldm;

        const int step = 64 * V;
        const int nfull = m / step;
        // Rows past n read row0 again (in bounds) and are not accumulated.
        // The next step's vectors are loaded before the current ones are
        // consumed, so each wave keeps R loads in flight while it computes.
        VT mv[R];
        if (nfull > 0) {
#pragma unroll
            for (int r = 0; r < R; r++) mv[r] = ld_stream(rowp[r] + lane * V);
        }
        for (int c = 0; c < nfull; c++) {
            const int j0 = c * step + lane * V;
            VT nx[R];
            if (c + 1 < nfull) {
#pragma unroll
                for (int r = 0; r < R; r++) nx[r] = ld_stream(rowp[r] + j0 + step);
            }
            double bk[V];
#pragma unroll
            for (int v = 0; v < V; v++) bk[v] = beta[j0 + v] * k;
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
        }
        {
            const int j0 = nfull * step + lane * V;
            if (j0 < m) {
                const int valid = (m - j0) < V ? (m - j0) : V;
                double bk[V];
#pragma unroll
                for (int v = 0; v < V; v++) bk[v] = (v < valid) ? beta[j0 + v] * k : 0.0;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    if (row0 + r < n) {
                        // ldm >= m rounded up to V: the vector stays inside the row pitch.
                        const VT mv = ld_stream(rowp[r] + j0);
                        row_accumulate<TM, TA, V, true>(sh[r], s[r], S[r], rm[r], bk, mv, k, valid);
                    }
                }
            }
        }

#pragma unroll
        for (int r = 0; r < R; r++) {
            const long i = row0 + r;
            if (i >= n) break;
            const double L = wave_log2sum(acc_log2(sh[r], S[r] + static_cast<double>(s[r])));
            if (lane == 0) {
                if (epi.lse2) epi.lse2[i] = L;
                if (aout) aout[i] = epi.reg * (epi.loga[i] - kLn2 * L);
                if (epi.rowsum || FIN) {
                    const double ri = exp2(rowpot[i] * k + L);
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
            if (blockIdx.x == 0) fin.ctl[3] = ctl[1];
        }
    }
}

// ---------------------------------------------------------------------------
// Column pass
// ---------------------------------------------------------------------------

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

// Workgroup = 4 waves over one 64V-column tile and one block of rows; the
// waves take CH-row slabs round robin.  Output: one FP64 log2 partial sum per
// (row block, column).
template <typename TM, typename TA>
__global__ __launch_bounds__(256) void k_col_partial(
    const TM* __restrict__ M, long ldm, int n, int m, const double* __restrict__ alpha,
    const double* __restrict__ ap0, const double* __restrict__ ap1,
    const double* __restrict__ colpot, const double* __restrict__ logb, double k, int rpb,
    double* __restrict__ part, const int* __restrict__ ctl)
{
    using VT = typename VecOf<TM>::type;
    constexpr int V = VecOf<TM>::N;
    constexpr int CH = 4;
    constexpr int TW = 64 * V;
    __shared__ double red[4][TW];
    if (ctl != nullptr && ctl[0] != 0) return;
    if (ap0 != nullptr) alpha = (ctl[1] & 1) ? ap1 : ap0;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int j0 = blockIdx.x * TW + lane * V;
    const int valid_cols = (m - j0) < V ? (m - j0) : V;   // <= 0: lane has no column
    const long rb = static_cast<long>(blockIdx.y) * rpb;
    const long re = (rb + rpb < n) ? rb + rpb : n;

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
                mv[c] = ld_stream(base + (r + c) * ldm);
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
                mv[c] = ld_stream(base + rr * ldm);
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
                double S = 0.0;
#pragma unroll
                for (int w = 0; w < 4; w++) {
                    const double L = red[w][lane * V + v];
                    S += (L == -INFINITY) ? 0.0 : exp2(L - Lm);
                }
                out = Lm + log2(S);
            }
            part[static_cast<long>(blockIdx.y) * m + j0 + v] = out;
        }
    }
}

constexpr int kCombineCols = 64;   // columns per combine workgroup (4 threads each)

// Stop test of the on-device BCD loop for iterate k = ctl[3], computed
// identically (same data, same order) by every combine workgroup:
// g_k = [r - a ; column residual of the update that produced beta_k],
// f_k = reg sum r - (alpha_k + beta_{k,m-1})'a - (beta_k - beta_{k,m-1})'b over
// j < m-1.  The test order is that of src/sinkhorn_bcd.cpp:65-79.
__device__ void bcd_decide(const BcdState& st, int kk, double* f_out, double* g_out,
                           bool* done_out)
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
    const double bm = st.beta[st.m - 1];
    const double gnorm = sqrt(v[0] + v[3]);
    const double da = v[2] + bm * st.sum_a;
    const double db = v[4] - bm * st.sum_b1;
    *f_out = st.reg * v[1] - da - db;
    *g_out = gnorm;
    *done_out = (gnorm < st.tol) || (kk >= st.max_iter);
}

// Combine the nblk partials of each column (4 threads per column, fixed
// order), then the column epilogue.  With BCD, first take the stop decision
// for iterate ctl[3]; when it stops, beta is left untouched (beta_k is the
// result), otherwise beta <- beta_{k+1} and the tile sums of the new column
// residual go to the other parity half of cpart.  Workgroup 0 records the
// history entry and advances or stops the loop.
template <bool BCD>
__global__ __launch_bounds__(256) void k_col_combine(const double* __restrict__ part, int nblk,
                                                     int m, double k, ColEpi epi, BcdState st,
                                                     const int* __restrict__ ctl)
{
    int kk = 0;
    double f = 0.0, gnorm = 0.0;
    bool done = false;
    if (BCD) {
        if (st.ctl[0] != 0) return;
        kk = st.ctl[3];
        bcd_decide(st, kk, &f, &gnorm, &done);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            const int h = kk % st.hist_cap;
            st.hist_f[h] = f;
            st.hist_g[h] = gnorm;
            st.hist_ts[h] = static_cast<long long>(wall_clock64());
        }
    } else if (ctl != nullptr && ctl[0] != 0) {
        return;
    }

    const int sub = threadIdx.x & 3;
    const int j = blockIdx.x * kCombineCols + (threadIdx.x >> 2);
    double r2 = 0.0, bb = 0.0;
    if (!done) {
        double Lm = -INFINITY;
        if (j < m)
            for (int p = sub; p < nblk; p += 4) Lm = fmax(Lm, part[static_cast<long>(p) * m + j]);
        Lm = fmax(Lm, __shfl_xor(Lm, 1));
        Lm = fmax(Lm, __shfl_xor(Lm, 2));
        double S = 0.0;
        if (j < m && Lm != -INFINITY)
            for (int p = sub; p < nblk; p += 4) {
                const double x = part[static_cast<long>(p) * m + j];
                S += (x == -INFINITY) ? 0.0 : exp2(x - Lm);
            }
        S += __shfl_xor(S, 1);
        S += __shfl_xor(S, 2);
        S = __shfl(S, threadIdx.x & ~3);
        if (sub == 0 && j < m) {
            const double L = (Lm == -INFINITY) ? -INFINITY : Lm + log2(S);
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
            st.ctl[0] = 1;
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

int row_pass_blocks(int bulk, int n)
{
    return ceil_div(n, 4 * (bulk == 0 ? kRowsPerWave<float> : kRowsPerWave<double>));
}

void launch_row_pass(int bulk, const void* M, long ldm, int n, int m, const double* beta,
                     double k, const RowEpi& epi, const int* ctl, hipStream_t s, bool precise)
{
    if (n <= 0) return;
    BcdState none{};
    const dim3 grid(row_pass_blocks(bulk, n));
    if (bulk == 0 && !precise) {
        constexpr int R = kRowsPerWave<float>;
        hipLaunchKernelGGL((k_row_pass<float, float, R, false>), grid, dim3(256), 0, s,
                           static_cast<const float*>(M), ldm, n, m, beta, k, epi, none, ctl);
    } else if (bulk == 0) {
        constexpr int R = kRowsPerWave<float>;
        hipLaunchKernelGGL((k_row_pass<float, double, R, false>), grid, dim3(256), 0, s,
                           static_cast<const float*>(M), ldm, n, m, beta, k, epi, none, ctl);
    } else {
        constexpr int R = kRowsPerWave<double>;
        hipLaunchKernelGGL((k_row_pass<double, double, R, false>), grid, dim3(256), 0, s,
                           static_cast<const double*>(M), ldm, n, m, beta, k, epi, none, ctl);
    }
    KERNEL_CHECK();
}

void launch_row_pass_bcd(int bulk, const void* M, long ldm, int n, int m, const double* beta,
                         double k, const RowEpi& epi, const BcdState& st, hipStream_t s)
{
    const dim3 grid(row_pass_blocks(bulk, n));
    if (bulk == 0) {
        constexpr int R = kRowsPerWave<float>;
        hipLaunchKernelGGL((k_row_pass<float, float, R, true>), grid, dim3(256), 0, s,
                           static_cast<const float*>(M), ldm, n, m, beta, k, epi, st, st.ctl);
    } else {
        constexpr int R = kRowsPerWave<double>;
        hipLaunchKernelGGL((k_row_pass<double, double, R, true>), grid, dim3(256), 0, s,
                           static_cast<const double*>(M), ldm, n, m, beta, k, epi, st, st.ctl);
    }
    KERNEL_CHECK();
}

int col_pass_blocks(int n, int m)
{
    // Aim for >= 512 workgroups (2 per CU on 256 CUs) with >= 16 rows each:
    // long column streams per workgroup measured fastest (16384^2 FP32:
    // 163 us at 512 workgroups, 175 us at 2048).
    const int tiles = ceil_div(m, 256);
    int nblk = ceil_div(512, tiles);
    const int maxblk = ceil_div(n, 16);
    if (nblk > maxblk) nblk = maxblk;
    if (nblk < 1) nblk = 1;
    return nblk;
}

int col_combine_blocks(int m) { return ceil_div(m, kCombineCols); }

void launch_col_pass(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                     const double* alpha_pair0, const double* alpha_pair1,
                     const double* colpot_guess, const double* logb, double k, double* part,
                     int nblk, const ColEpi& epi, const BcdState* st, const int* ctl,
                     hipStream_t s, bool precise)
{
    if (m <= 0 || n <= 0) return;
    const int rpb = ceil_div(n, nblk);
    const dim3 grid(ceil_div(m, bulk == 0 ? 256 : 128), nblk);
    if (bulk == 0 && !precise)
        hipLaunchKernelGGL((k_col_partial<float, float>), grid, dim3(256), 0, s,
                           static_cast<const float*>(M), ldm, n, m, alpha, alpha_pair0,
                           alpha_pair1, colpot_guess, logb, k, rpb, part, ctl);
    else if (bulk == 0)
        hipLaunchKernelGGL((k_col_partial<float, double>), grid, dim3(256), 0, s,
                           static_cast<const float*>(M), ldm, n, m, alpha, alpha_pair0,
                           alpha_pair1, colpot_guess, logb, k, rpb, part, ctl);
    else
        hipLaunchKernelGGL((k_col_partial<double, double>), grid, dim3(256), 0, s,
                           static_cast<const double*>(M), ldm, n, m, alpha, alpha_pair0,
                           alpha_pair1, colpot_guess, logb, k, rpb, part, ctl);
    KERNEL_CHECK();
    if (st != nullptr)
        hipLaunchKernelGGL(k_col_combine<true>, dim3(col_combine_blocks(m)), dim3(256), 0, s,
                           part, nblk, m, k, epi, *st, ctl);
    else
        hipLaunchKernelGGL(k_col_combine<false>, dim3(col_combine_blocks(m)), dim3(256), 0, s,
                           part, nblk, m, k, epi, BcdState{}, ctl);
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
