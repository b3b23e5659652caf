// hess_mf.hip -- products with the sparsified Hessian of the FP32 bulk as a
// dense FP32 matrix of its kept entries (the "matrix-free" storage: no CSR /
// CSC, no indices).
//
// The threshold rule keeps entry (i, j), j < m-1, iff T_ij >= min(tau_j,
// rho_i) (or T_ij underflows, which adds nothing to a product), and sigma_ij
// = T_ij / reg.  At the densities SSNS reaches (tens of percent) the two
// stored copies cost 2 x 12 bytes per kept entry and product, more than a
// dense FP32 read (4 bytes per entry).  The build writes the kept matrix K
// (K_ij = T_ij where kept, else 0; row pitch ldt) from the same pass that
// decides the kept set, with the sparsifier's own element arithmetic (log2 T
// in FP64, exp2f), so its values are the stored copies' bit for bit; within
// one Newton step every product then reads K instead of re-forming T from M
// (no exponent arguments, no exponentials, no kept test per element).
//
// The Schur-complement product Delta u = (h2 + s) u - sigma' ((sigma u) /
// (h1 + s)) needs both halves: one pass does both, since a workgroup that
// holds whole rows knows y_i = (sigma u)_i / (h1_i + s) as soon as row i is
// summed, and adds sigma_ij y_i into column accumulators from the same
// registers.
//
// Layout: one workgroup of 512 threads per compute unit holds a contiguous
// block of rows; thread t owns the 4-column vectors at 4t + 2048c (c < NC,
// m - 1 <= 2048 NC <= 16384) with FP64 column accumulators in registers; the
// row input vector is staged in LDS (FP64).  Products and sums are FP64.
// Each workgroup writes one FP64 column partial per column; mf_fold sums
// them in a fixed order.
#include "dense_device.hpp"
#include "hess_mf.hpp"

#include <cmath>

namespace regot {
namespace {

using namespace dev;

constexpr int kMaxCols = 16384;          // m - 1 <= 16384

// Sum over the wave in a fixed order (the same bits on every lane).
__device__ __forceinline__ double wave_total(double v) { return group_sum<64>(v); }

// Row r of an FP32 matrix (row pointer Mi) in the pass layout: out[c] holds
// columns 4t + SPAN c .. + 3; the last chunk past m - 1 is zero.
template <int NC, bool FULL, int SPAN>
__device__ __forceinline__ void load_row(const float* __restrict__ Mi, int t, int vlast,
                                         float4* out, bool nt)
{
#pragma unroll
    for (int c = 0; c < NC; c++) {
        if (FULL || c < NC - 1 || vlast > 0)
            out[c] = ld_m(Mi + 4 * t + c * SPAN, nt);
        else
            out[c] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// Store chunk c of a row in the pass layout (the last chunk's columns past
// m - 1, vlast of them valid, are left alone).
template <int NC, bool FULL>
__device__ __forceinline__ void store4(float* __restrict__ Ti, int t, int c, int span, int vlast,
                                       const float (&v)[4])
{
    float* p = Ti + 4 * t + c * span;
    if (FULL || c < NC - 1 || vlast >= 4) {
        *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (q < vlast) p[q] = v[q];
    }
}

// The full Hessian's kept matrix: K_ij = T_ij for every j < m - 1 (the
// sparsifier's arithmetic), the build of the top-k rule at density 1.
template <int NC, bool FULL>
__global__ __launch_bounds__(512) void k_mf_full(MfArgs a)
{
    constexpr int kSpan = 2048;
    const int t = threadIdx.x;
    const int m1 = a.m1;
    const long r0 = static_cast<long>(blockIdx.x) * a.rpw;
    const long r1 = (r0 + a.rpw < a.n) ? r0 + a.rpw : a.n;
    int vlast = m1 - (4 * t + (NC - 1) * kSpan);
    vlast = FULL ? 4 : (vlast < 0 ? 0 : (vlast > 4 ? 4 : vlast));
    auto ok = [vlast](int c, int q) { return FULL || c < NC - 1 || q < vlast; };
    double bj[NC][4];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int j0 = 4 * t + c * kSpan;
#pragma unroll
        for (int q = 0; q < 4; q++) bj[c][q] = ok(c, q) ? a.beta[j0 + q] : 0.0;
    }
    // the row in flight and its alpha_i load together
    float4 pf[NC];
    double pa = 0.0;
    auto prefetch = [&](long r) {
        load_row<NC, FULL, kSpan>(a.M + r * a.ldm, t, vlast, pf, a.nt != 0);
        pa = a.alpha[r];
    };
    if (r0 < r1) prefetch(r0);
#pragma unroll 1
    for (long i = r0; i < r1; i++) {
        float e[NC][4];
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) e[c][q] = vget<float4, float>(pf[c], q);
        const double ai = pa;
        if (i + 1 < r1) prefetch(i + 1);
        float* Ti = a.Tout + i * a.ldt;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            float k4[4];
#pragma unroll
            for (int q = 0; q < 4; q++)
                k4[q] = exp2f(static_cast<float>(
                    fma(-a.k, static_cast<double>(e[c][q]), (ai + bj[c][q]) * a.k)));
            store4<NC, FULL>(Ti, t, c, kSpan, vlast, k4);
        }
    }
}

// Workgroup-wide exclusive scan of one double per thread (512 threads), in
// thread order; *total gets the sum.  Two barriers.
__device__ __forceinline__ double scan512(double v, double* wt, double* total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double u = __shfl_up(inc, o);
        if (lane >= o) inc += u;
    }
    if (lane == 63) wt[wave] = inc;
    __syncthreads();
    double off = 0.0, tot = 0.0;
#pragma unroll
    for (int w = 0; w < 8; w++) {
        if (w < wave) off += wt[w];
        tot += wt[w];
    }
    *total = tot;
    __syncthreads();
    return off + inc - v;
}

// Descending bucket search over hist[0..nb) (nb <= 2048, 512 threads, four
// buckets each): the largest b with above + sum(hist[b..nb-1]) >= thr;
// *before = above + sum(hist[b+1..]).  -1 when none qualifies.
__device__ int select_desc512(const double* hist, int nb, double above, double thr,
                              double* before, double* wt, int* claim, double* claim_before)
{
    const int t = threadIdx.x;
    if (t == 0) *claim = -1;
    const int u = 511 - t;   // thread t owns buckets [4u, 4u + 4), top buckets first
    double own = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (4 * u + q < nb) own += hist[4 * u + q];
    double total;
    double cum = above + scan512(own, wt, &total);
    int found = -1;
    double fb = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int b = 4 * u + 3 - q;
        if (b >= nb) continue;
        const double nxt = cum + hist[b];
        if (nxt >= thr && found < 0) {
            found = b;
            fb = cum;
        }
        cum = nxt;
    }
    if (found >= 0) {
        claim_before[u] = fb;
        atomicMax(claim, found);
    }
    __syncthreads();
    const int b = *claim;
    int res = -1;
    if (b >= 0) {
        res = b;
        *before = claim_before[b / 4];
    }
    __syncthreads();
    return res;
}

// Row side of the threshold rule in the pass layout (FP32 bulk, m - 1 <=
// 16384), one read of M: per row the dropped mass D_i = sum_j [T_ij < tau_j]
// T_ij; rho_i = +inf when D_i <= 2 delta, else the radix select of
// k_row_threshold (the largest dropped value v with sum_{d >= v} d >= D_i -
// 2 delta: 11-, 11- and 10-bit digits, FP64 masses) over the row held in
// registers; the stored count (T_ij >= min(tau_j, rho_i), T_ij != 0), the
// FP64-zero count (exp2 argument < -1075) and, when a.Tout is not null, the
// row of the dense kept matrix (T_ij where kept, else 0).  The element
// arithmetic is Elem<float>::val's.
template <int NC, bool FULL>
__global__ __launch_bounds__(512) void k_mf_rule(MfArgs a, double delta2,
                                                 double* __restrict__ rho, int* __restrict__ cnt,
                                                 int* __restrict__ zcnt)
{
    constexpr int kSpan = 2048;
    constexpr int kWaves = 8;
    // tau'_j = max(tau_j, 2^-149) (FP32 values; +inf past m - 1): T_ij < tau'_j
    // is the dropped test and its negation the kept test (T_ij >= tau_j and
    // T_ij != 0), since the only value below 2^-149 is 0
    __shared__ float4 stau[NC * kSpan / 4];
    __shared__ double hist[2048];
    __shared__ double rd[2][kWaves];
    __shared__ int rn[2][kWaves], rn2[kWaves];
    __shared__ double wt[kWaves];
    __shared__ double claim_before[512];
    __shared__ int claim;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int m1 = a.m1;
    const long r0 = static_cast<long>(blockIdx.x) * a.rpw;
    const long r1 = (r0 + a.rpw < a.n) ? r0 + a.rpw : a.n;
    int vlast = m1 - (4 * t + (NC - 1) * kSpan);
    vlast = FULL ? 4 : (vlast < 0 ? 0 : (vlast > 4 ? 4 : vlast));
    auto ok = [vlast](int c, int q) { return FULL || c < NC - 1 || q < vlast; };
    double bj[NC][4];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int j0 = 4 * t + c * kSpan;
        float tq[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            bj[c][q] = ok(c, q) ? a.beta[j0 + q] : 0.0;
            tq[q] = ok(c, q) ? fmaxf(static_cast<float>(a.tau[j0 + q]), 0x1p-149f) : INFINITY;
        }
        stau[t + c * (kSpan / 4)] = make_float4(tq[0], tq[1], tq[2], tq[3]);
    }
    __syncthreads();
    // tau of this thread's columns (c, q)
    auto tau_of = [&](int c, int q) { return vget<float4, float>(stau[t + c * (kSpan / 4)], q); };
    float4 pf[NC];
    double pa = 0.0;
    auto prefetch = [&](long r) {
        load_row<NC, FULL, kSpan>(a.M + r * a.ldm, t, vlast, pf, a.nt != 0);
        pa = a.alpha[r];
    };
    if (r0 < r1) prefetch(r0);
#pragma unroll 1
    for (long i = r0; i < r1; i++) {
        float e[NC][4];
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) e[c][q] = vget<float4, float>(pf[c], q);
        const double ai = pa;
        if (i + 1 < r1) prefetch(i + 1);
        // kept bits per element (the stored count is their popcount)
        unsigned int bits = 0u;
        int nz = 0;
        double dsum = 0.0;
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                double x = fma(-a.k, static_cast<double>(e[c][q]), (ai + bj[c][q]) * a.k);
                float tv = exp2f(static_cast<float>(x));
                if (!ok(c, q)) {   // past m - 1 (last chunk only): no mass, not kept
                    x = 0.0;
                    tv = 0.f;
                }
                e[c][q] = tv;
                const bool dropped = tv < tau_of(c, q);
                dsum += static_cast<double>(dropped ? tv : 0.f);
                bits |= dropped ? 0u : (1u << (4 * c + q));
                nz += x < -1075.0 ? 1 : 0;
            }
        const int slot = static_cast<int>((i - r0) & 1);
        {
            const double dw = group_sum<64>(dsum);
            const int cw = wave_isum(__popc(bits) | (nz << 16));   // both <= 2^14 per wave
            if (lane == 0) {
                rd[slot][wave] = dw;
                rn[slot][wave] = cw;
            }
        }
        __syncthreads();
        double D = 0.0;
        int NS = 0, NZ = 0;
#pragma unroll
        for (int w = 0; w < kWaves; w++) {
            D += rd[slot][w];
            NS += rn[slot][w] & 0xffff;
            NZ += rn[slot][w] >> 16;
        }
        double r_i = INFINITY;
        if (D > delta2) {   // uniform: D is the same in every lane
            const double target = D - delta2;
            unsigned int prefix = 0u;
            double above = 0.0;
#pragma unroll 1
            for (int p = 0; p < 3; p++) {
                const int shift = p == 0 ? 21 : (p == 1 ? 10 : 0);
                const int width = p == 2 ? 10 : 11;
                for (int q = t; q < 2048; q += 512) hist[q] = 0.0;
                __syncthreads();
#pragma unroll
                for (int c = 0; c < NC; c++)
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        if (!ok(c, q)) continue;
                        const float d = e[c][q] < tau_of(c, q) ? e[c][q] : 0.f;
                        const unsigned int key = __float_as_uint(d);
                        if (p == 0 || (key >> (shift + width)) == prefix)
                            atomicAdd(&hist[(key >> shift) & ((1u << width) - 1u)],
                                      static_cast<double>(d));
                    }
                __syncthreads();
                double before = 0.0;
                const int b = select_desc512(hist, 1 << width, above, target, &before, wt,
                                             &claim, claim_before);
                if (b < 0) break;   // cannot happen: the dropped mass exceeds target
                above = before;
                prefix = (prefix << width) | static_cast<unsigned int>(b);
            }
            r_i = static_cast<double>(__uint_as_float(prefix));
            const float rf = __uint_as_float(prefix);
            bits = 0u;
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (!ok(c, q)) continue;
                    const bool kept = e[c][q] >= fminf(tau_of(c, q), rf) && e[c][q] != 0.f;
                    bits |= kept ? (1u << (4 * c + q)) : 0u;
                }
            int nw = __popc(bits);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) nw += __shfl_xor(nw, o);
            if (lane == 0) rn2[wave] = nw;
            __syncthreads();
            NS = 0;
#pragma unroll
            for (int w = 0; w < kWaves; w++) NS += rn2[w];
            __syncthreads();
        }
        if (t == 0) {
            rho[i] = r_i;
            cnt[i] = NS;
            zcnt[i] = NZ;
        }
        if (a.Tout != nullptr) {
            float* Ti = a.Tout + i * a.ldt;
#pragma unroll
            for (int c = 0; c < NC; c++) {
                float k4[4];
#pragma unroll
                for (int q = 0; q < 4; q++) k4[q] = ((bits >> (4 * c + q)) & 1u) ? e[c][q] : 0.f;
                store4<NC, FULL>(Ti, t, c, kSpan, vlast, k4);
            }
        }
    }
}

// MODE (see hess_mf.hpp): which halves run and what weights the columns.
// K rows are read two ahead (they and the per-row operands load together:
// a load of its own inside the loop would wait for the whole prefetch).
template <int MODE, int NC, bool FULL, int TH>
__global__ __launch_bounds__(TH) void k_mf_pass(MfArgs a)
{
    constexpr bool ROW = MODE == MF_CG || MODE == MF_HS || MODE == MF_ROWS;
    constexpr bool COL = MODE != MF_ROWS;
    constexpr bool WH1 = MODE == MF_CG || MODE == MF_DIAG;
    constexpr bool WXR = MODE == MF_HS || MODE == MF_COLS;
    constexpr int kThreads = TH;
    constexpr int kWaves = TH / 64;
    constexpr int kSpan = 4 * TH;
    __shared__ double sx[ROW ? NC * kSpan : 1];
    __shared__ double red[2][kWaves];
    if (a.ctl != nullptr && a.ctl[0] != 0) return;

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    if (MODE == MF_CG && a.rr_part != nullptr) {
        // every workgroup sums the partials in the same order, so all decide
        // alike; only workgroup 0 writes, after its own test
        if (wave == 0) {
            double v = 0.0;
            for (int b = lane; b < a.rr_n; b += 64) v += a.rr_part[b];
            v = wave_sum0(v);
            if (lane == 0) red[0][0] = v;
        }
        __syncthreads();
        const double rr = red[0][0];
        const bool stop = rr < a.scal[5];
        if (blockIdx.x == 0 && t == 0) {
            a.scal[4] = rr;
            if (stop) {
                a.ctl[1] = a.it;   // the converged update is not counted
                a.ctl[0] = 1;
            }
        }
        if (stop) return;
        __syncthreads();   // red[0][0] is reused by the row reduction
    }
    const int m1 = a.m1;
    const int r0 = blockIdx.x * a.rpw;
    const int r1 = (r0 + a.rpw < a.n) ? r0 + a.rpw : a.n;

    int vlast = m1 - (4 * t + (NC - 1) * kSpan);
    vlast = FULL ? 4 : (vlast < 0 ? 0 : (vlast > 4 ? 4 : vlast));
    auto ok = [vlast](int c, int q) { return FULL || c < NC - 1 || q < vlast; };

    if (ROW)
        for (int j = t; j < m1; j += kThreads) sx[j] = a.xc[j];
    double acc[NC][4];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[c][q] = 0.0;
    if (ROW) __syncthreads();

    float4 pf0[NC], pf1[NC];
    double w0 = 0.0, w1 = 0.0;   // the row weight input (h1_i or xr_i) of the rows in flight
    auto load = [&](int r, float4* pf, double* w) {
        load_row<NC, FULL, kSpan>(a.T + static_cast<long>(r) * a.ldt, t, vlast, pf, a.nt != 0);
        if (WH1) *w = a.h1[r];
        if (WXR) *w = a.xr[r];
    };
    if (r0 < r1) load(r0, pf0, &w0);
    if (r0 + 1 < r1) load(r0 + 1, pf1, &w1);
#pragma unroll 1
    for (int i = r0; i < r1; i++) {
        float e[NC][4];
#pragma unroll
        for (int c = 0; c < NC; c++) {
#pragma unroll
            for (int q = 0; q < 4; q++) e[c][q] = ok(c, q) ? vget<float4, float>(pf0[c], q) : 0.f;
            pf0[c] = pf1[c];
        }
        const double wi = w0;
        w0 = w1;
        if (i + 2 < r1) load(i + 2, pf1, &w1);
        double cw = 0.0;
        if (ROW) {
            double racc = 0.0;
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int q = 0; q < 4; q++)
                    if (ok(c, q))
                        racc = fma(static_cast<double>(e[c][q]), sx[4 * t + c * kSpan + q], racc);
            const int slot = (i - r0) & 1;
            const double S = wave_total(racc);
            if (lane == 0) red[slot][wave] = S;
            __syncthreads();
            // every lane sums the wave totals in the same order
            double rv = 0.0;
#pragma unroll
            for (int w = 0; w < kWaves; w++) rv += red[slot][w];
            rv *= a.inv_reg;   // (sigma xc)_i
            if (MODE == MF_CG) {
                cw = rv / (wi + a.shift) * a.inv_reg;
            } else {
                if (t == 0) a.rowout[i] = rv;
                cw = (MODE == MF_HS) ? wi * a.inv_reg : 0.0;
            }
        } else if (MODE == MF_COLS) {
            cw = wi * a.inv_reg;
        } else {   // MF_DIAG: sum_i sigma_ij^2 / (h1_i + s)
            cw = a.inv_reg * a.inv_reg / (wi + a.shift);
        }
        if (COL) {
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const double tv = static_cast<double>(e[c][q]);
                    acc[c][q] = fma(MODE == MF_DIAG ? tv * tv : tv, cw, acc[c][q]);
                }
        }
    }
    if (COL) {
        double* pw = a.part + static_cast<long>(blockIdx.x) * m1;
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (ok(c, q)) pw[4 * t + c * kSpan + q] = acc[c][q];
    }
}

template <int MODE, int NC, int TH>
void pass_launch(const MfArgs& a, int nwg, bool full, hipStream_t s)
{
    if (full)
        hipLaunchKernelGGL((k_mf_pass<MODE, NC, true, TH>), dim3(nwg), dim3(TH), 0, s, a);
    else
        hipLaunchKernelGGL((k_mf_pass<MODE, NC, false, TH>), dim3(nwg), dim3(TH), 0, s, a);
}

template <int MODE>
void pass_nc(const MfArgs& a, int nwg, hipStream_t s)
{
    const int NC = ceil_div(a.m1, 2048);
    const bool full = a.m1 == NC * 2048;
    switch (NC) {
    case 1: pass_launch<MODE, 1, 512>(a, nwg, full, s); break;
    case 2: pass_launch<MODE, 2, 512>(a, nwg, full, s); break;
    case 3: pass_launch<MODE, 3, 512>(a, nwg, full, s); break;
    case 4: pass_launch<MODE, 4, 512>(a, nwg, full, s); break;
    case 5: pass_launch<MODE, 5, 512>(a, nwg, full, s); break;
    case 6: pass_launch<MODE, 6, 512>(a, nwg, full, s); break;
    case 7: pass_launch<MODE, 7, 512>(a, nwg, full, s); break;
    default: pass_launch<MODE, 8, 512>(a, nwg, full, s); break;
    }
}

// out_j = sum_w part[w][j] (fixed order), then by epi: 0 out = sum; 1 out =
// z_j - sum (sigma' applied and subtracted); 2 out = (h2_j + shift) z_j -
// sum (Delta form); 3 out = h2_j z_j + sum (lower block of Hs);
// 4 out = 1 / max(h2_j + shift - sum, floor) (Jacobi diagonal of Delta).
__global__ __launch_bounds__(256) void k_mf_fold(const double* __restrict__ part, int nwg, int m1,
                                                 int epi, const double* __restrict__ z,
                                                 const double* __restrict__ h2, double shift,
                                                 double* __restrict__ out)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= m1) return;
    double s = 0.0;
    int w = 0;
    for (; w + 8 <= nwg; w += 8) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; u++) x[u] = part[static_cast<long>(w + u) * m1 + j];
#pragma unroll
        for (int u = 0; u < 8; u++) s += x[u];
    }
    for (; w < nwg; w++) s += part[static_cast<long>(w) * m1 + j];
    switch (epi) {
    case 0: out[j] = s; break;
    case 1: out[j] = z[j] - s; break;
    case 2: out[j] = h2[j] * z[j] + shift * z[j] - s; break;
    case 3: out[j] = h2[j] * z[j] + s; break;
    default: {
        double d = h2[j] + shift - s;
        const double floor = 1e-12 * (h2[j] + shift);
        if (!(d > floor)) d = (h2[j] + shift > 0.0) ? h2[j] + shift : 1.0;
        out[j] = 1.0 / d;
    }
    }
}

}  // namespace

bool mf_shape_ok(int n, int m) { return n >= 1 && m >= 2 && m - 1 <= kMaxCols; }

long mf_pitch(int m) { return (static_cast<long>(m - 1) + 63) / 64 * 64; }

int mf_blocks(int n, int ncu)
{
    int g = ncu > 0 ? ncu : 256;
    if (g > n) g = n;
    const int rpw = ceil_div(n, g);
    return ceil_div(n, rpw);
}

template <int NC>
void full_launch(const MfArgs& a, int nwg, bool full, hipStream_t s)
{
    if (full)
        hipLaunchKernelGGL((k_mf_full<NC, true>), dim3(nwg), dim3(512), 0, s, a);
    else
        hipLaunchKernelGGL((k_mf_full<NC, false>), dim3(nwg), dim3(512), 0, s, a);
}

void launch_mf_full(MfArgs a, int nwg, hipStream_t s)
{
    if (!mf_shape_ok(a.n, a.m1 + 1)) fail(E_INTERNAL, "kept matrix: %d x %d", a.n, a.m1 + 1);
    if (a.Tout == nullptr || a.ldt < a.m1) fail(E_INTERNAL, "kept matrix: no output");
    a.rpw = ceil_div(a.n, nwg);
    a.nt = 0;
    const int NC = ceil_div(a.m1, 2048);
    const bool full = a.m1 == NC * 2048;
    switch (NC) {
    case 1: full_launch<1>(a, nwg, full, s); break;
    case 2: full_launch<2>(a, nwg, full, s); break;
    case 3: full_launch<3>(a, nwg, full, s); break;
    case 4: full_launch<4>(a, nwg, full, s); break;
    case 5: full_launch<5>(a, nwg, full, s); break;
    case 6: full_launch<6>(a, nwg, full, s); break;
    case 7: full_launch<7>(a, nwg, full, s); break;
    default: full_launch<8>(a, nwg, full, s); break;
    }
    KERNEL_CHECK();
}

template <int NC>
void rule_launch(const MfArgs& a, int nwg, bool full, double delta2, double* rho, int* cnt,
                 int* zcnt, hipStream_t s)
{
    if (full)
        hipLaunchKernelGGL((k_mf_rule<NC, true>), dim3(nwg), dim3(512), 0, s, a, delta2, rho, cnt,
                           zcnt);
    else
        hipLaunchKernelGGL((k_mf_rule<NC, false>), dim3(nwg), dim3(512), 0, s, a, delta2, rho,
                           cnt, zcnt);
}

void launch_mf_rule(MfArgs a, int nwg, double delta2, double* rho, int* cnt, int* zcnt,
                    hipStream_t s)
{
    if (!mf_shape_ok(a.n, a.m1 + 1)) fail(E_INTERNAL, "threshold rows: %d x %d", a.n, a.m1 + 1);
    if (a.Tout != nullptr && a.ldt < a.m1) fail(E_INTERNAL, "kept matrix pitch %ld", a.ldt);
    a.rpw = ceil_div(a.n, nwg);
    a.nt = 0;
    const int NC = ceil_div(a.m1, 2048);
    const bool full = a.m1 == NC * 2048;
    switch (NC) {
    case 1: rule_launch<1>(a, nwg, full, delta2, rho, cnt, zcnt, s); break;
    case 2: rule_launch<2>(a, nwg, full, delta2, rho, cnt, zcnt, s); break;
    case 3: rule_launch<3>(a, nwg, full, delta2, rho, cnt, zcnt, s); break;
    case 4: rule_launch<4>(a, nwg, full, delta2, rho, cnt, zcnt, s); break;
    case 5: rule_launch<5>(a, nwg, full, delta2, rho, cnt, zcnt, s); break;
    case 6: rule_launch<6>(a, nwg, full, delta2, rho, cnt, zcnt, s); break;
    case 7: rule_launch<7>(a, nwg, full, delta2, rho, cnt, zcnt, s); break;
    default: rule_launch<8>(a, nwg, full, delta2, rho, cnt, zcnt, s); break;
    }
    KERNEL_CHECK();
}

void launch_mf_pass(int mode, MfArgs a, int nwg, hipStream_t s)
{
    if (!mf_shape_ok(a.n, a.m1 + 1)) fail(E_INTERNAL, "matrix-free pass: %d x %d", a.n, a.m1 + 1);
    if (a.T == nullptr || a.ldt < a.m1) fail(E_INTERNAL, "matrix-free pass: no kept matrix");
    a.rpw = ceil_div(a.n, nwg);
    a.nt = static_cast<long>(a.n) * a.ldt * 4 > (256L << 20) ? 1 : 0;
    switch (mode) {
    case MF_CG: pass_nc<MF_CG>(a, nwg, s); break;
    case MF_HS: pass_nc<MF_HS>(a, nwg, s); break;
    case MF_ROWS: pass_nc<MF_ROWS>(a, nwg, s); break;
    case MF_COLS: pass_nc<MF_COLS>(a, nwg, s); break;
    default: pass_nc<MF_DIAG>(a, nwg, s); break;
    }
    KERNEL_CHECK();
}

void launch_mf_fold(const double* part, int nwg, int m1, int epi, const double* z,
                    const double* h2, double shift, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_mf_fold, dim3(ceil_div(m1, 256)), dim3(256), 0, s, part, nwg, m1, epi,
                       z, h2, shift, out);
    KERNEL_CHECK();
}

}  // namespace regot
