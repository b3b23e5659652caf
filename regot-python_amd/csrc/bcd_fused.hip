// bcd_fused.hip -- one BCD iteration in one read of the cost matrix (FP32 bulk).
//
// A BCD iteration is alpha <- optimal_alpha(beta), then beta <-
// optimal_beta(alpha) (src/sinkhorn_bcd.cpp:65-103).  Both half-steps need a
// full sweep of M, but the second only needs alpha_{k+1,i} for the rows it
// sums, and row i's new potential is known as soon as row i has been summed.
// So one workgroup of 512 threads holds whole rows: thread t owns the
// 4-column vectors at 4t + 2048c (c < NC, m <= 2048 NC <= 16384), a row
// arrives as one coalesced sweep, and per row
//
//   e_ij  = 2^((beta_j - M_ij) k - sh_i)          (FP64 argument, FP32 exp2)
//   tot_i = sum_j e_ij                            (FP32 over the thread's
//                                                  <= 32 terms, FP64 above)
//   alpha_{k+1,i} = reg (log a_i - ln2 (sh_i + log2 tot_i))
//   u_ij  = e_ij * (a_i / tot_i) = exp((alpha_{k+1,i} + beta_j - M_ij) / reg)
//
// u is the plan at (alpha_{k+1}, beta_k), whose column sums give beta_{k+1}:
// each thread adds u into FP32 column accumulators, moved into FP64 LDS
// accumulators every 16 rows, and the workgroup writes one FP64 column sum
// per column.  The combine kernel then sums the workgroups' column sums in a
// fixed order and takes the stop decision as in the two-pass iteration.  One
// exp2 per element, one read of M per iteration.  beta_k lives in registers
// (FP64, 8 NC per thread), the next one or two rows' vectors are in flight
// while a row is reduced; one workgroup per CU.
//
// Range: sh_i = log2 a_i - k alpha_{k,i} makes tot_i = r_i / a_i, which is ~1
// near the solution; if tot_i leaves [2^-20, 2^20] (or over/underflows) the
// row is redone with its largest term as the shift.  The column terms u_ij
// are then <= a_i and a flushed e_ij (< 2^-126) drops < 2^-106 a_i of a column
// sum.  The host uses this pass only when the previous half-step was an exact
// beta update (column sums b_j): then every column sum is >= b_j min_i a_i /
// sum b, and the host requires min a min b >= 2^-60 sum a sum b, so the
// dropped mass is below 2^-46 of any column sum.
#include "dense.hpp"
#include "dense_device.hpp"

#include <cmath>

namespace regot {
namespace {

using namespace dev;

constexpr int kFThreads = 512;           // 2 waves per SIMD: 256 VGPRs per wave
constexpr int kFWaves = kFThreads / 64;
constexpr int kFSpan = 4 * kFThreads;     // columns covered by one vector step
constexpr int kFMaxNC = 8;                // m <= 8 kFSpan = 16384
constexpr int kFFlush = 16;               // rows between FP32 -> FP64 column flushes
constexpr double kTotLo = 0x1p-20;
constexpr double kTotHi = 0x1p20;

// Row i's vectors of the thread's columns.  Only the last vector (c = NC-1)
// can reach past m: vlast = its number of valid columns (0..4); lanes with
// vlast = 0 load nothing.
template <int NC, bool FULL>
__device__ __forceinline__ void load_row(const float* __restrict__ Mi, int t, int vlast,
                                         float4* out, bool nt)
{
#pragma unroll
    for (int c = 0; c < NC; c++) {
        if (FULL || c < NC - 1 || vlast > 0)
            out[c] = ld_m(Mi + 4 * t + c * kFSpan, nt);
        else
            out[c] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// t_ij = (beta_j - M_ij) k - sh in FP64, rounded to FP32 (-inf past m).
template <bool FULL>
__device__ __forceinline__ float targ(float mij, double bkj, double sh, double k, bool ok)
{
    const float x = static_cast<float>(fma(-k, static_cast<double>(mij), bkj - sh));
    return (FULL || ok) ? x : ninf<float>();
}

// Sum of the 8 wave totals v[] in a fixed order, read by every thread (the
// same bits everywhere).
__device__ __forceinline__ double sum8(const double* v)
{
    return ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
}

// Sum of the kFWaves wave totals v[] in a fixed butterfly order, computed by
// the low lanes of every wave: the result is bitwise the same on every thread.
__device__ __forceinline__ double sum_waves(const double* v, int lane)
{
    double x = v[lane & (kFWaves - 1)];
#pragma unroll
    for (int o = kFWaves / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kFWaves);
    return __shfl(x, 0);
}

__device__ __forceinline__ double max_waves(const double* v, int lane)
{
    double x = v[lane & (kFWaves - 1)];
#pragma unroll
    for (int o = kFWaves / 2; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, kFWaves));
    return __shfl(x, 0);
}

// Rows per workgroup are bounded by the per-row (shift, total) record kept
// for the epilogue.
constexpr int kFMaxRows = 1024;

template <int NC, bool FULL, int PF>
__global__ __launch_bounds__(kFThreads) void k_bcd_fused(
    const float* __restrict__ M, long ldm, int n, int m, const double* __restrict__ beta,
    double k, double reg, const double* __restrict__ loga, const double* __restrict__ a,
    const double* __restrict__ alpha, double* __restrict__ alpha_out, int rpw,
    double* __restrict__ part, double* __restrict__ rpart, double* __restrict__ bm,
    const int* __restrict__ ctl, int nt)
{
    if (ctl[0] != 0) return;
    __shared__ double acc[4 * NC][kFThreads];
    __shared__ double rsh[kFMaxRows], rtot[kFMaxRows];
    __shared__ double red[2][kFWaves];
    __shared__ double redx[2][kFWaves];

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const long r0 = static_cast<long>(blockIdx.x) * rpw;
    const long r1 = (r0 + rpw < n) ? r0 + rpw : n;

    int vlast = m - (4 * t + (NC - 1) * kFSpan);
    vlast = FULL ? 4 : (vlast < 0 ? 0 : (vlast > 4 ? 4 : vlast));
    // column (c, q) of this thread is valid iff ok(c, q)
    auto ok = [vlast](int c, int q) { return FULL || c < NC - 1 || q < vlast; };
    double bk[NC][4];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int j0 = 4 * t + c * kFSpan;
#pragma unroll
        for (int q = 0; q < 4; q++) bk[c][q] = ok(c, q) ? beta[j0 + q] * k : 0.0;
    }
    float acc32[NC][4];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            acc32[c][q] = 0.f;
            acc[4 * c + q][t] = 0.0;   // each thread's own slots: no barrier needed
        }

    // PF rows in flight: pf0 = row i, pf1 = row i + 1 (PF = 2)
    float4 pf0[NC], pf1[NC];
    if (r0 < r1) load_row<NC, FULL>(M + r0 * ldm, t, vlast, pf0, nt != 0);
    if (PF == 2 && r0 + 1 < r1) load_row<NC, FULL>(M + (r0 + 1) * ldm, t, vlast, pf1, nt != 0);

#pragma unroll 1
    for (long i = r0; i < r1; i++) {
        float e[NC][4];
#pragma unroll
        for (int c = 0; c < NC; c++) {
#pragma unroll
            for (int q = 0; q < 4; q++) e[c][q] = vget<float4, float>(pf0[c], q);
            if (PF == 2) pf0[c] = pf1[c];
        }
        if (PF == 2) {
            if (i + 2 < r1) load_row<NC, FULL>(M + (i + 2) * ldm, t, vlast, pf1, nt != 0);
        } else {
            if (i + 1 < r1) load_row<NC, FULL>(M + (i + 1) * ldm, t, vlast, pf0, nt != 0);
        }

        double sh = loga[i] * kLog2e - alpha[i] * k;
        float s = 0.f;
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                e[c][q] = ex2(targ<FULL>(e[c][q], bk[c][q], sh, k, ok(c, q)));
                s += e[c][q];
            }
        const int slot = static_cast<int>((i - r0) & 1);
        double S[1] = {static_cast<double>(s)};
        wave_sums<1>(S);
        if (lane == 0) red[slot][wave] = S[0];
        __syncthreads();
        double tot = sum8(red[slot]);

        if (!(tot >= kTotLo && tot <= kTotHi)) {
            // Rare (first iterations from a poor start): shift by the row's
            // exact maximum.  tot is bitwise the same on every thread, so the
            // whole workgroup takes this branch.  The row is read again.
            const float* Mi = M + i * ldm;
            double tm = -INFINITY;
#pragma unroll
            for (int c = 0; c < NC; c++) {
                float4 rv = make_float4(0.f, 0.f, 0.f, 0.f);
                if (ok(c, 0)) rv = ld_m(Mi + 4 * t + c * kFSpan, false);
#pragma unroll
                for (int q = 0; q < 4; q++)
                    if (ok(c, q))
                        tm = fmax(tm, fma(-k, static_cast<double>(vget<float4, float>(rv, q)),
                                          bk[c][q]));
            }
            const double dmw = wave_max(tm);
            if (lane == 0) redx[0][wave] = dmw;
            __syncthreads();
            const double dm = max_waves(redx[0], lane);
            tot = 0.0;
            if (dm > -INFINITY && dm < INFINITY) {
                sh = dm;   // log2 of the row's largest term
                s = 0.f;
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    float4 rv = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (ok(c, 0)) rv = ld_m(Mi + 4 * t + c * kFSpan, false);
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        e[c][q] = ex2(targ<FULL>(vget<float4, float>(rv, q), bk[c][q], sh, k,
                                                 ok(c, q)));
                        s += e[c][q];
                    }
                }
                const double S2 = wave_sum0(static_cast<double>(s));
                if (lane == 0) redx[1][wave] = S2;
                __syncthreads();
                tot = sum_waves(redx[1], lane);
            }
        }

        if (t == 0) {
            rsh[i - r0] = sh;
            rtot[i - r0] = tot;
        }
        const float wi = tot > 0.0 ? static_cast<float>(a[i] / tot) : 0.f;
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int q = 0; q < 4; q++) acc32[c][q] = fmaf(e[c][q], wi, acc32[c][q]);
        if (((i - r0) % kFFlush) == kFFlush - 1) {
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    acc[4 * c + q][t] += static_cast<double>(acc32[c][q]);
                    acc32[c][q] = 0.f;
                }
        }
    }

    double* pw = part + static_cast<long>(blockIdx.x) * m;
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (ok(c, q)) pw[4 * t + c * kFSpan + q] = acc[4 * c + q][t] + acc32[c][q];

    // Row epilogue, rows strided over the threads: alpha_{k+1} and the
    // stop-test sums at iterate k (r_i = sum_j exp((alpha_i + beta_j - M_ij) /
    // reg)).
    __syncthreads();
    double v[3] = {0.0, 0.0, 0.0};
    for (int rr = t; rr < r1 - r0; rr += kFThreads) {
        const long i = r0 + rr;
        const double L = rtot[rr] > 0.0 ? rsh[rr] + log2(rtot[rr]) : -INFINITY;
        const double ali = alpha[i], ai = a[i];
        alpha_out[i] = reg * (loga[i] - kLn2 * L);
        const double ri = exp2(ali * k + L);
        const double g = ri - ai;
        v[0] += g * g;
        v[1] += ri;
        v[2] += ali * ai;
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const double x = wave_sum0(v[q]);
        if (lane == 0) redx[q & 1][wave] = x;   // redx is free again here
        __syncthreads();
        const double tq = sum_waves(redx[q & 1], lane);
        if (t == 0) rpart[3 * blockIdx.x + q] = tq;
        __syncthreads();
    }
    // beta_k[m-1] for the objective (the combine overwrites beta)
    if (t == 0 && blockIdx.x == 0) bm[0] = beta[m - 1];
}

// Two rows in flight (PF = 2) wherever that fits the 256 VGPRs of a wave
// without spills: every full-width case, ragged m up to NC = 6.
template <int NC>
void fused_launch(bool full, dim3 grid, hipStream_t s, const float* M, long ldm, int n, int m,
                  const double* beta, double k, double reg, const double* loga, const double* a,
                  const double* alpha, double* alpha_out, int rpw, double* part, double* rpart,
                  double* bm, const int* ctl, int nt)
{
    constexpr int PFR = NC <= 6 ? 2 : 1;
    if (full)
        hipLaunchKernelGGL((k_bcd_fused<NC, true, 2>), grid, dim3(kFThreads), 0, s, M, ldm, n,
                           m, beta, k, reg, loga, a, alpha, alpha_out, rpw, part, rpart, bm, ctl,
                           nt);
    else
        hipLaunchKernelGGL((k_bcd_fused<NC, false, PFR>), grid, dim3(kFThreads), 0, s, M, ldm,
                           n, m, beta, k, reg, loga, a, alpha, alpha_out, rpw, part, rpart, bm,
                           ctl, nt);
}

}  // namespace

int bcd_fused_blocks(int n, int ncu)
{
    int g = ncu > 0 ? ncu : 256;
    if (g > n) g = n;
    const int rpw = ceil_div(n, g);
    return ceil_div(n, rpw);
}

bool bcd_fused_shape_ok(int n, int m, int ncu)
{
    if (n < 1 || m <= 1024 || m > kFMaxNC * kFSpan) return false;
    return ceil_div(n, bcd_fused_blocks(n, ncu)) <= kFMaxRows;
}

void launch_bcd_fused(const void* M, long ldm, int n, int m, const double* beta, double k,
                      double reg, const double* loga, const double* a, const double* alpha,
                      double* alpha_out, int nblk, double* part, double* rpart, double* bm,
                      const int* ctl, hipStream_t s)
{
    const int rpw = ceil_div(n, nblk);
    if (n < 1 || m <= 1024 || m > kFMaxNC * kFSpan || rpw > kFMaxRows)
        fail(E_INTERNAL, "single-read BCD pass: shape %d x %d over %d workgroups", n, m, nblk);
    if (static_cast<long>(rpw) * (nblk - 1) >= n)
        fail(E_INTERNAL, "single-read BCD pass: %d workgroups for %d rows", nblk, n);
    const int NC = ceil_div(m, kFSpan);
    const bool full = (m == NC * kFSpan);
    const int nt = static_cast<long>(n) * ldm * 4 > (256L << 20) ? 1 : 0;
    const float* Mf = static_cast<const float*>(M);
    const dim3 grid(nblk);
    switch (NC) {
    case 1:
        fused_launch<1>(full, grid, s, Mf, ldm, n, m, beta, k, reg, loga, a, alpha, alpha_out,
                        rpw, part, rpart, bm, ctl, nt);
        break;
    case 2:
        fused_launch<2>(full, grid, s, Mf, ldm, n, m, beta, k, reg, loga, a, alpha, alpha_out,
                        rpw, part, rpart, bm, ctl, nt);
        break;
    case 3:
        fused_launch<3>(full, grid, s, Mf, ldm, n, m, beta, k, reg, loga, a, alpha, alpha_out,
                        rpw, part, rpart, bm, ctl, nt);
        break;
    case 4:
        fused_launch<4>(full, grid, s, Mf, ldm, n, m, beta, k, reg, loga, a, alpha, alpha_out,
                        rpw, part, rpart, bm, ctl, nt);
        break;
    case 5:
        fused_launch<5>(full, grid, s, Mf, ldm, n, m, beta, k, reg, loga, a, alpha, alpha_out,
                        rpw, part, rpart, bm, ctl, nt);
        break;
    case 6:
        fused_launch<6>(full, grid, s, Mf, ldm, n, m, beta, k, reg, loga, a, alpha, alpha_out,
                        rpw, part, rpart, bm, ctl, nt);
        break;
    case 7:
        fused_launch<7>(full, grid, s, Mf, ldm, n, m, beta, k, reg, loga, a, alpha, alpha_out,
                        rpw, part, rpart, bm, ctl, nt);
        break;
    default:
        fused_launch<8>(full, grid, s, Mf, ldm, n, m, beta, k, reg, loga, a, alpha, alpha_out,
                        rpw, part, rpart, bm, ctl, nt);
        break;
    }
    KERNEL_CHECK();
}

}  // namespace regot
