// batch.hip -- many independent EROT problems solved by BCD (Sinkhorn) in
// shared launches.
//
// All cost matrices live in one packed device buffer (problem p row-major at
// moff[p] with a 256-byte row pitch).  One BCD round of every active problem
// is a few launches whose workgroups come from tile tables:
//
//   single-read    (p, t): rows t fr .. t fr + fr - 1 of problem p whole ->
//   tiles          alpha_{k+1}, the row partials of g_k and the column sums at
//                  alpha_{k+1} (FP32 problems with 1024 < m <= 4096: one read
//                  of M per iteration; one launch per column-chunk count);
//   row tiles      (p, t): rows 4R t .. 4R t + 4R - 1 of problem p -> alpha_{k+1}
//                  and the row partials of g_k;
//   column tiles   (p, s, b): column strip s (64V columns) over row block b of
//                  problem p -> partial log2 column sums at alpha_{k+1};
//   combine tiles  (p, c): columns 64c .. 64c + 63 of problem p -> every tile
//                  of p takes p's stop decision for iterate k itself (same data,
//                  same order), then beta_{k+1} and its column residual.
//
// The other problems' rounds are row tiles then column tiles (two reads of M).
//
// Each problem follows the single-problem semantics of sinkhorn_bcd
// (src/sinkhorn_bcd.cpp:14-124): test ||g_k|| < tol first, then update, and
// niter = k at the stop.  A stopped problem's tiles return at once; every few
// rounds the host drops them from the tables (the work queue shrinks as
// problems converge).  When few tiles remain (the long tail of slow
// problems), the tables switch to fine tiles -- one-row waves, 64-row column
// blocks -- so that a few problems still spread over the whole chip, and M
// is read through the caches once the active matrices fit the 256 MB
// Infinity Cache.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <vector>

#include "dense_device.hpp"
#include "regot_hip.h"

namespace regot {

int set_error(int code, const char* what);

namespace {

using namespace dev;

constexpr int kBlockRows = 512;      // rows per column-pass row block (coarse tiles)
constexpr int kBlockRowsFine = 64;   // rows per column-pass row block (fine tiles)
constexpr int kCombCols = 64;        // columns per combine tile (4 threads each)
constexpr long kFineBelow = 2048;    // coarse row tiles below which fine tiles are used

template <typename TM>
constexpr int kRows = sizeof(TM) == 4 ? 4 : 2;   // rows per wave in row tiles

struct Desc {
    long moff;          // element offset of M_p in the packed buffer
    int n, m, ldm;
    int aoff, boff;     // offsets into the packed n-sized / m-sized vectors
    int rt0, nrt;       // row tiles of p: partial slots rt0 .. rt0 + nrt - 1
    int rpt;            // rows per row tile (4R coarse, R fine)
    int nblk, rpb;      // column-pass row blocks, rows per block
    long poff;          // offset of p's column partials (nblk x m doubles)
    int ct0, nct;       // combine tiles of p: residual slots (2 parities)
    double reg, k;      // k = log2(e) / reg
    double sum_a, sum_b1;
    // single-read rounds (fused = 1): nft tiles of fr whole rows, column
    // chunks fnc = ceil(m / 1024), column sums at fpoff (nft x m doubles);
    // then nrt = nft and the tiles' row partials take the row tiles' slots
    int fused, fnc, fr, nft;
    long fpoff;
};

struct Vecs {
    const double* a;
    const double* b;
    const double* loga;
    const double* logb;
    double* ap0;        // alpha, double buffer by round parity
    double* ap1;
    double* beta;
    double* rpart;      // 3 per row tile
    double* part;       // column partials (log2 sums, two-pass rounds)
    double* fpart;      // column sums of the single-read tiles
    double* cpart;      // 2 per combine tile per parity
    double* bm;         // beta_k[m-1] per problem (stored by row tile 0)
    int* done;          // per problem
    int* niter;
    double* fval;       // f and ||g|| at the decision
    double* gval;
};

// Each wave owns R whole rows, 4R rows per tile (R = kRows coarse, 1 fine).
template <typename TM, int R>
__global__ __launch_bounds__(256) void k_batch_row(const TM* __restrict__ M,
                                                   const Desc* __restrict__ D,
                                                   const int2* __restrict__ tiles, Vecs v,
                                                   const int* __restrict__ kp, int nt)
{
    const int2 tl = tiles[blockIdx.x];
    const int p = tl.x;
    if (v.done[p]) return;
    const int k = *kp;
    const Desc d = D[p];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const double* rowpot = ((k & 1) ? v.ap1 : v.ap0) + d.aoff;
    double* aout = ((k & 1) ? v.ap0 : v.ap1) + d.aoff;
    const double* loga = v.loga + d.aoff;
    const double* a = v.a + d.aoff;
    double wg2 = 0.0, wr = 0.0, waa = 0.0;
    const long row0 = (static_cast<long>(tl.y) * 4 + wave) * R;
    if (row0 < d.n) {
        double L[R];
        row_lse<TM, TM, R>(M + d.moff, d.ldm, d.n, 0, d.m, v.beta + d.boff, d.k, loga, rowpot,
                           row0, lane, L, nt != 0);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const long i = row0 + r;
            if (i >= d.n) break;
            if (lane == 0) {
                aout[i] = d.reg * (loga[i] - kLn2 * L[r]);
                const double ri = exp2(rowpot[i] * d.k + L[r]);
                const double g = ri - a[i];
                wg2 += g * g;
                wr += ri;
                waa += rowpot[i] * a[i];
            }
        }
    }
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
            v.rpart[3L * (d.rt0 + tl.y) + q] = ((red[0][q] + red[1][q]) + red[2][q]) + red[3][q];
        if (tl.y == 0) v.bm[p] = v.beta[d.boff + d.m - 1];
    }
}

// init: alpha_0 = 0 and no stop flag is read.
template <typename TM>
__global__ __launch_bounds__(256) void k_batch_colpart(const TM* __restrict__ M,
                                                       const Desc* __restrict__ D,
                                                       const int4* __restrict__ tiles, Vecs v,
                                                       const int* __restrict__ kp, int init,
                                                       int nt)
{
    constexpr int TW = 64 * VecOf<TM>::N;
    const int4 tl = tiles[blockIdx.x];
    const int p = tl.x;
    if (!init && v.done[p]) return;
    const int k = *kp;
    const Desc d = D[p];
    const double* alpha = (((k + 1) & 1) ? v.ap1 : v.ap0) + d.aoff;   // alpha_{k+1}
    const long rb = static_cast<long>(tl.z) * d.rpb;
    const long re = (rb + d.rpb < d.n) ? rb + d.rpb : d.n;
    col_partial_tile<TM, TM>(M + d.moff, d.ldm, d.m, alpha, init ? nullptr : v.beta + d.boff,
                                v.logb + d.boff, d.k, tl.y * TW, rb, re,
                                v.part + d.poff + static_cast<long>(tl.z) * d.m, nt != 0);
}

// ---------------------------------------------------------------------------
// Single-read round (FP32, 1024 < m <= 4096): one BCD iteration in one read of
// M.  alpha_{k+1,i} needs only row i, and the beta half-step needs only
// alpha_{k+1,i} for the rows it sums, so tile (p, t) owns rows
// t fr .. t fr + fr - 1 of problem p whole: thread th holds the 4-column
// vectors at 4 th + 1024 c (c < fnc) and beta_{k,j} k for them in FP64
// registers, and per row
//
//   e_ij  = 2^((beta_j - M_ij) k - sh_i)   (FP64 argument, FP32 exp2)
//   tot_i = sum_j e_ij                     (FP32 per thread, FP64 across)
//   alpha_{k+1,i} = reg (log a_i - ln2 (sh_i + log2 tot_i))
//   u_ij  = e_ij a_i / tot_i = exp((alpha_{k+1,i} + beta_j - M_ij) / reg)
//
// u is the plan at (alpha_{k+1}, beta_k): its FP32 column accumulators move
// into FP64 LDS accumulators every 16 rows and the tile writes one FP64
// column sum per column; the combine adds the tiles' sums in tile order.
// sh_i = log2 a_i - k alpha_{k,i} makes tot_i = r_i / a_i (about 1 near the
// solution); a row whose tot_i leaves [2^-20, 2^20] is redone with its
// largest term as the shift.  The host puts a problem on these rounds only
// when min a min b >= 2^-60 sum a sum b, so the terms flushed below 2^-126
// (u_ij <= a_i) drop less than 2^-46 of any column sum (which is at least
// b_j min a / sum b after an exact beta update, as every round's beta is).
// ---------------------------------------------------------------------------
constexpr int kBFThreads = 256;
constexpr int kBFWaves = kBFThreads / 64;
constexpr int kBFSpan = 4 * kBFThreads;   // columns per vector step
constexpr int kBFMaxNC = 4;               // m <= 4 * 1024
constexpr int kBFMinM = 1025;             // narrower problems keep two-pass rounds
constexpr int kBFMaxRows = 128;           // rows per tile (the row record below)
constexpr int kBFMinRows = 16;
constexpr int kBFTiles = 4096;            // tile count aimed at once tiles shrink
constexpr int kBFFlush = 16;              // rows between FP32 -> FP64 flushes
constexpr double kBFTotLo = 0x1p-20;
constexpr double kBFTotHi = 0x1p20;

// Sum / maximum of the 4 wave values v[] in a fixed butterfly order, bitwise
// the same on every thread.
__device__ __forceinline__ double bf_sum_waves(const double* v, int lane)
{
    double x = v[lane & (kBFWaves - 1)];
#pragma unroll
    for (int o = kBFWaves / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, kBFWaves);
    return __shfl(x, 0);
}

__device__ __forceinline__ double bf_max_waves(const double* v, int lane)
{
    double x = v[lane & (kBFWaves - 1)];
#pragma unroll
    for (int o = kBFWaves / 2; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, kBFWaves));
    return __shfl(x, 0);
}

// Rows per step for NC column chunks: a step's rows arrive together (about
// 32 KB of M per workgroup) and are reduced with one barrier.
template <int NC>
constexpr int kBFRows = NC == 1 ? 8 : (NC == 2 ? 4 : (NC == 3 ? 3 : 2));

// Thread th's NC 4-column vectors of one row.  Every load is issued (the
// compiler then counts the loads in flight exactly): a vector that would
// start at or past m reads the row's first vector instead, and the caller
// masks it.  A vector that starts below m ends inside the row pitch (a
// multiple of 64 floats).
template <int NC, bool NT>
__device__ __forceinline__ void bf_load(const float* __restrict__ Mi, const int* off,
                                        float4* out)
{
#pragma unroll
    for (int c = 0; c < NC; c++) out[c] = ld_m(Mi + off[c], NT);
}

// NT: nontemporal loads (the active matrices exceed the Infinity Cache).
template <int NC, bool NT>
__global__ __launch_bounds__(kBFThreads) void k_batch_fused(const float* __restrict__ M,
                                                            const Desc* __restrict__ D,
                                                            const int2* __restrict__ tiles,
                                                            Vecs v, const int* __restrict__ kp)
{
    constexpr int R = kBFRows<NC>;
    constexpr int FLUSH = kBFFlush / R;   // steps between FP32 -> FP64 flushes
    const int2 tl = tiles[blockIdx.x];
    const int p = tl.x;
    if (v.done[p]) return;
    const int k = *kp;
    const Desc d = D[p];
    __shared__ double acc[4 * NC][kBFThreads];
    __shared__ double rsh[kBFMaxRows], rtot[kBFMaxRows], ra[kBFMaxRows];
    __shared__ double red[2][R][kBFWaves];
    __shared__ double redx[2][kBFWaves];

    const int th = threadIdx.x;
    const int lane = th & 63;
    const int wave = __builtin_amdgcn_readfirstlane(th >> 6);
    const int m = d.m;
    const long ldm = d.ldm;
    const long r0 = static_cast<long>(tl.y) * d.fr;
    const long r1 = (r0 + d.fr < d.n) ? r0 + d.fr : d.n;
    const float* Mp = M + d.moff;
    const double kk = d.k;
    const double* alpha = ((k & 1) ? v.ap1 : v.ap0) + d.aoff;   // alpha_k
    double* aout = ((k & 1) ? v.ap0 : v.ap1) + d.aoff;          // alpha_{k+1}
    const double* loga = v.loga + d.aoff;
    const double* a = v.a + d.aoff;
    const double* beta = v.beta + d.boff;

    // beta_j k, or -inf for columns at or past m: their terms are 0 (M is
    // finite there: row padding, or the row's first vector)
    int off[NC];
    double bk[NC][4];
    float acc32[NC][4];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int j0 = 4 * th + c * kBFSpan;
        off[c] = j0 < m ? j0 : 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            bk[c][q] = j0 + q < m ? beta[j0 + q] * kk : -INFINITY;
            acc32[c][q] = 0.f;
            acc[4 * c + q][th] = 0.0;   // each thread's own slots: no barrier needed
        }
    }
    // the rows' shifts and marginals in LDS: inside the row loop only the
    // loads of M go to memory
    for (int rr = th; rr < r1 - r0; rr += kBFThreads) {
        const long i = r0 + rr;
        rsh[rr] = loga[i] * kLog2e - alpha[i] * kk;
        ra[rr] = a[i];
    }
    __syncthreads();

    // the next step's rows are in flight while a step is reduced (past the
    // tile's end a load repeats the tile's last row)
    float4 pf[R][NC];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const long ir = (r0 + r < r1) ? r0 + r : r1 - 1;
        bf_load<NC, NT>(Mp + ir * ldm, off, pf[r]);
    }

    int step = 0;
#pragma unroll 1
    for (long i0 = r0; i0 < r1; i0 += R, step++) {
        float e[R][NC][4];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int q = 0; q < 4; q++) e[r][c][q] = vget<float4, float>(pf[r][c], q);
#pragma unroll
        for (int r = 0; r < R; r++) {
            const long ir = (i0 + R + r < r1) ? i0 + R + r : r1 - 1;
            bf_load<NC, NT>(Mp + ir * ldm, off, pf[r]);
        }

        // e_ij in place of M_ij; rows past r1 (the same for the whole
        // workgroup) contribute nothing
        double sh[R];
        float s[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const bool rv = i0 + r < r1;
            sh[r] = rv ? rsh[i0 + r - r0] : 0.0;
            s[r] = 0.f;
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const float x = rv ? static_cast<float>(fma(-kk,
                                                                static_cast<double>(e[r][c][q]),
                                                                bk[c][q] - sh[r]))
                                       : ninf<float>();
                    e[r][c][q] = ex2(x);
                    s[r] += e[r][c][q];
                }
        }
        const int slot = step & 1;
        double S[R];
#pragma unroll
        for (int r = 0; r < R; r++) S[r] = static_cast<double>(s[r]);
        wave_sums<R>(S);
        if (lane == 0) {
#pragma unroll
            for (int r = 0; r < R; r++) red[slot][r][wave] = S[r];
        }
        __syncthreads();

#pragma unroll
        for (int r = 0; r < R; r++) {
            const long i = i0 + r;
            if (i >= r1) break;   // the same for the whole workgroup
            // the four wave totals in a fixed order (the same on every thread)
            double tot = ((red[slot][r][0] + red[slot][r][1]) + red[slot][r][2]) + red[slot][r][3];
            if (!(tot >= kBFTotLo && tot <= kBFTotHi)) {
                // Rare (first rounds): shift by the row's exact largest term.
                // tot is the same on every thread, so the whole workgroup is
                // here; the row is read again.
                float4 rw[NC];
#pragma unroll
                for (int c = 0; c < NC; c++) rw[c] = ld_m(Mp + i * ldm + off[c], false);
                double tm = -INFINITY;
#pragma unroll
                for (int c = 0; c < NC; c++)
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        tm = fmax(tm, fma(-kk, static_cast<double>(vget<float4, float>(rw[c], q)),
                                          bk[c][q]));
                const double dmw = wave_max(tm);
                if (lane == 0) redx[0][wave] = dmw;
                __syncthreads();
                const double dm = bf_max_waves(redx[0], lane);
                tot = 0.0;
                if (dm > -INFINITY && dm < INFINITY) {
                    sh[r] = dm;   // log2 of the row's largest term
                    float s2 = 0.f;
#pragma unroll
                    for (int c = 0; c < NC; c++)
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            const float x = static_cast<float>(
                                fma(-kk, static_cast<double>(vget<float4, float>(rw[c], q)),
                                    bk[c][q] - sh[r]));
                            e[r][c][q] = ex2(x);
                            s2 += e[r][c][q];
                        }
                    const double S2 = wave_sum0(static_cast<double>(s2));
                    if (lane == 0) redx[1][wave] = S2;
                    __syncthreads();
                    tot = bf_sum_waves(redx[1], lane);
                }
            }
            // rsh[i - r0] was read by every thread before the step's barrier
            if (th == 0) {
                rsh[i - r0] = sh[r];
                rtot[i - r0] = tot;
            }
            const float wi = tot > 0.0 ? static_cast<float>(ra[i - r0] / tot) : 0.f;
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int q = 0; q < 4; q++) acc32[c][q] = fmaf(e[r][c][q], wi, acc32[c][q]);
        }
        if ((step % FLUSH) == FLUSH - 1) {
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    acc[4 * c + q][th] += static_cast<double>(acc32[c][q]);
                    acc32[c][q] = 0.f;
                }
        }
    }

    double* pw = v.fpart + d.fpoff + static_cast<long>(tl.y) * m;
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (4 * th + c * kBFSpan + q < m)
                pw[4 * th + c * kBFSpan + q] = acc[4 * c + q][th] + acc32[c][q];

    // Row epilogue, rows strided over the threads: alpha_{k+1} and the row
    // partials of g_k (r_i = sum_j exp((alpha_{k,i} + beta_j - M_ij) / reg)).
    __syncthreads();
    double s3[3] = {0.0, 0.0, 0.0};
    for (int rr = th; rr < r1 - r0; rr += kBFThreads) {
        const long i = r0 + rr;
        const double L = rtot[rr] > 0.0 ? rsh[rr] + log2(rtot[rr]) : -INFINITY;
        const double ali = alpha[i], ai = a[i];
        aout[i] = d.reg * (loga[i] - kLn2 * L);
        const double ri = exp2(ali * kk + L);
        const double g = ri - ai;
        s3[0] += g * g;
        s3[1] += ri;
        s3[2] += ali * ai;
    }
    block_sums_256<3>(s3);
    if (th == 0) {
#pragma unroll
        for (int q = 0; q < 3; q++) v.rpart[3L * (d.rt0 + tl.y) + q] = s3[q];
        if (tl.y == 0) v.bm[p] = beta[m - 1];
    }
}

// Stop decision for iterate k of problem p (identical in every tile of p):
// g_k = [r - a ; residual of the update that produced beta_k].
__device__ void batch_decide(const Desc& d, const Vecs& v, int p, int k, double tol,
                             int max_iter, double* f, double* gn, bool* done)
{
    const double* cp = v.cpart + 2L * (2L * d.ct0 + static_cast<long>(k & 1) * d.nct);
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < d.nrt; q += 256) {
        s[0] += v.rpart[3L * (d.rt0 + q)];
        s[1] += v.rpart[3L * (d.rt0 + q) + 1];
        s[2] += v.rpart[3L * (d.rt0 + q) + 2];
    }
    for (int q = threadIdx.x; q < d.nct; q += 256) {
        s[3] += cp[2 * q];
        s[4] += cp[2 * q + 1];
    }
    block_sums_256<5>(s);
    const double bm = v.bm[p];
    *gn = sqrt(s[0] + s[3]);
    *f = d.reg * s[1] - (s[2] + bm * d.sum_a) - (s[4] - bm * d.sum_b1);
    *done = (*gn < tol) || (k >= max_iter);
}

__global__ __launch_bounds__(256) void k_batch_combine(const Desc* __restrict__ D,
                                                       const int2* __restrict__ tiles, Vecs v,
                                                       const int* __restrict__ kp, int init,
                                                       double tol, int max_iter)
{
    const int2 tl = tiles[blockIdx.x];
    const int p = tl.x;
    if (!init && v.done[p]) return;
    const int k = *kp;
    const Desc d = D[p];
    if (!init) {
        double f, gn;
        bool done;
        batch_decide(d, v, p, k, tol, max_iter, &f, &gn, &done);
        if (tl.y == 0 && threadIdx.x == 0) {
            v.fval[p] = f;
            v.gval[p] = gn;
            if (done) {
                v.niter[p] = k;
                v.done[p] = 1;
            }
        }
        if (done) return;
    }
    const int sub = threadIdx.x & 3;
    const int j = tl.y * kCombCols + (threadIdx.x >> 2);
    const bool valid = j < d.m;
    double L;
    if (d.fused && !init) {
        // the single-read tiles' plain column sums of exp((alpha_{k+1,i} +
        // beta_{k,j} - M_ij) / reg), added in tile order p = sub, sub + 4, ...
        // (8 loads in flight), so log2 sum_i 2^((alpha_{k+1,i} - M_ij) k) =
        // log2(sum) - beta_{k,j} k; beta_{k,j} is read before sub 0 of the
        // same column overwrites it below
        double cs = 0.0;
        if (valid) {
            const double* fp = v.fpart + d.fpoff + j;
            int q = sub;
            for (; q + 28 < d.nft; q += 32) {
                double x[8];
#pragma unroll
                for (int u = 0; u < 8; u++) x[u] = fp[static_cast<long>(q + 4 * u) * d.m];
#pragma unroll
                for (int u = 0; u < 8; u++) cs += x[u];
            }
            for (; q < d.nft; q += 4) cs += fp[static_cast<long>(q) * d.m];
        }
        cs += __shfl_xor(cs, 1);
        cs += __shfl_xor(cs, 2);
        L = (valid && cs > 0.0) ? log2(cs) - v.beta[d.boff + j] * d.k : -INFINITY;
    } else {
        L = combine_partials(v.part + d.poff, d.nblk, d.m, j, valid, sub);
    }
    double r2 = 0.0, bb = 0.0;
    if (sub == 0 && valid) {
        const double lb = v.logb[d.boff + j];
        const double bj = d.reg * (lb - kLn2 * L);
        v.beta[d.boff + j] = bj;
        // residual of the update in FP64: b_j (exp(beta_j/reg + ln2 L - log b_j) - 1)
        const double g = v.b[d.boff + j] * expm1(bj / d.reg + kLn2 * L - lb);
        if (j < d.m - 1) {
            r2 = g * g;
            bb = bj * v.b[d.boff + j];
        }
    }
    double s[2] = {r2, bb};
    block_sums_256<2>(s);
    if (threadIdx.x == 0) {
        // residual slots of beta_{k+1}: parity (k + 1) & 1 (init: beta_0, parity 0)
        const int par = init ? 0 : ((k + 1) & 1);
        double* cp = v.cpart + 2L * (2L * d.ct0 + static_cast<long>(par) * d.nct);
        cp[2 * tl.y] = s[0];
        cp[2 * tl.y + 1] = s[1];
    }
}

// gamma_p = [alpha + beta_{m-1}; beta_{0..m-2} - beta_{m-1}] with alpha =
// alpha_{niter}; one workgroup per problem.
__global__ __launch_bounds__(256) void k_batch_dual(const Desc* __restrict__ D, Vecs v,
                                                    double* __restrict__ dual)
{
    const int p = blockIdx.x;
    const Desc d = D[p];
    const int it = v.niter[p];
    const double* alpha = ((it & 1) ? v.ap1 : v.ap0) + d.aoff;
    const double* beta = v.beta + d.boff;
    const double bm = beta[d.m - 1];
    double* out = dual + (d.aoff + d.boff - p);
    for (int i = threadIdx.x; i < d.n; i += 256) out[i] = alpha[i] + bm;
    for (int j = threadIdx.x; j < d.m - 1; j += 256) out[d.n + j] = beta[j] - bm;
}

template <typename Tin, typename Tout>
__global__ void k_pack(const Tin* __restrict__ src, long lds, Tout* __restrict__ dst, long ldd,
                       int n, int m)
{
    const long total = static_cast<long>(n) * ldd;
    for (long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; t < total;
         t += static_cast<long>(gridDim.x) * 256) {
        const long i = t / ldd, j = t % ldd;
        dst[t] = (j < m) ? static_cast<Tout>(src[i * lds + j]) : Tout(0);
    }
}

// Advances the device round counter read by the round's kernels; the last
// launch of every round, so one captured sequence of rounds replays as is.
__global__ void k_round_next(int* kp)
{
    if (threadIdx.x == 0) *kp += 1;
}

__global__ void k_batch_init(double* ap0, const double* a, double* loga, long na,
                             const double* b, double* logb, long nb, int* done, int* niter,
                             int np, int* kp)
{
    const long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
    if (t == 0) *kp = -1;   // the init round (beta_0) runs as round -1
    if (t < na) {
        ap0[t] = 0.0;
        loga[t] = log(a[t]);
    }
    if (t < nb) logb[t] = log(b[t]);
    if (t < np) {
        done[t] = 0;
        niter[t] = 0;
    }
}

}  // namespace
}  // namespace regot

using namespace regot;

struct regot_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int nprob = 0;
    int bulk = 0;
    std::vector<Desc> desc;
    long total_n = 0, total_m = 0, total_dual = 0;
    DBuf<unsigned char> M;
    DBuf<Desc> d_desc;
    DBuf<double> a, b, loga, logb, ap0, ap1, beta, rpart, part, fpart, cpart, bm, fval, gval;
    DBuf<int> done, niter, kdev;
    DBuf<int2> t_row, t_comb, t_fus;
    DBuf<int4> t_col;
    HBuf<int> h_done;
    HBuf<int2> h_row[2], h_comb[2], h_fus[2];
    HBuf<int4> h_col[2];
    HBuf<Desc> h_desc[2];
    int hd_slot = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    long long total_elems = 0;
    bool fine = false;   // tile granularity of the current tables
    // coarse tiles: rows per wave of the row tiles and rows per column-pass
    // block; fixed rules unless an autotune run picked them
    int coarse_R = 4;
    int coarse_rows = kBlockRows;
    bool autotune = false, tuned = false;
    bool started_fine = false;   // the last solve began in fine tiles
    double tune_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool nt = true;      // nontemporal loads of M (active matrices beyond the cache)
    // single-read rounds: problems that qualify (fused_ok) take them unless
    // the option turns them off; fpart holds fpart_cap doubles
    std::vector<char> fused_ok;
    bool use_fused = true;
    long fpart_cap = 0;
    // K rounds captured once per tile-table configuration (and tol,
    // max_iter) and replayed
    bool use_graph = false;
    hipGraphExec_t gexec = nullptr;
    static constexpr int kGraphKey = 13;
    int g_key[kGraphKey] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
    double g_tol = -1.0;
    long long graph_builds = 0, graph_launches = 0;

    ~regot_batch()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    Vecs vecs()
    {
        Vecs v;
        v.a = a;
        v.b = b;
        v.loga = loga;
        v.logb = logb;
        v.ap0 = ap0;
        v.ap1 = ap1;
        v.beta = beta;
        v.rpart = rpart;
        v.part = part;
        v.fpart = fpart;
        v.cpart = cpart;
        v.bm = bm;
        v.done = done;
        v.niter = niter;
        v.fval = fval;
        v.gval = gval;
        return v;
    }

    // Tile granularity of the two-pass tiles in every problem's descriptor.
    void set_mode(bool f)
    {
        for (Desc& d : desc) {
            d.rpt = f ? 4 : 4 * coarse_R;
            d.nrt = ceil_div(d.n, d.rpt);
            d.nblk = ceil_div(d.n, f ? kBlockRowsFine : coarse_rows);
            d.rpb = ceil_div(d.n, d.nblk);
        }
        fine = f;
    }

    bool takes_fused(int p) const { return use_fused && fused_ok[p]; }

    // Tile tables of the listed problems into staging slot s, and the
    // descriptors.  Problems on single-read rounds get single-read tiles of
    // fr rows, fr chosen so that about kBFTiles tiles cover their rows
    // (kBFMinRows .. kBFMaxRows); the others get row and column tiles, which
    // switch to fine tiles when the coarse tables would be small.  init_cols:
    // every problem gets column tiles (the init round computes beta_0 by
    // column tiles alone).
    void build_tables(const std::vector<int>& act, int s, int* nrow, int* ncol, int* ncomb,
                      int* nfus, bool init_cols, bool refresh = false)
    {
        // the coarse / fine switch counts tiles of the default rows per wave,
        // so it does not move with the tuned tile shape
        const int R0 = bulk == 0 ? 4 : 2;
        long ntiles = 0, frows = 0;
        double bytes = 0.0;
        for (int p : act) {
            if (takes_fused(p)) frows += desc[p].n;
            if (!takes_fused(p) || init_cols) ntiles += ceil_div(desc[p].n, 4 * R0);
            bytes += static_cast<double>(desc[p].n) * desc[p].ldm * (bulk == 0 ? 4 : 8);
        }
        nt = bytes > static_cast<double>(256L << 20);
        const bool want_fine = ntiles < kFineBelow;
        if (want_fine != fine || refresh) set_mode(want_fine);
        const int fr = static_cast<int>(std::min<long>(
            kBFMaxRows, std::max<long>(kBFMinRows, ceil_div(frows, static_cast<long>(kBFTiles)))));
        long cr = 0, cc = 0, cb = 0, cf = 0, foff = 0;
        long cfc[kBFMaxNC] = {0, 0, 0, 0};
        const int vw = bulk == 0 ? 4 : 2;
        for (int p : act) {
            Desc& d = desc[p];
            d.fused = takes_fused(p) ? 1 : 0;
            if (d.fused) {
                d.nft = ceil_div(d.n, fr);
                d.fr = ceil_div(d.n, d.nft);
                d.fnc = ceil_div(d.m, kBFSpan);
                d.fpoff = foff;
                foff += static_cast<long>(d.nft) * d.m;
                // the single-read tiles' row partials use the row-tile slots
                d.nrt = d.nft;
                d.rpt = d.fr;
                cf += d.nft;
                cfc[d.fnc - 1] += d.nft;
            } else {
                cr += d.nrt;
            }
            if (!d.fused || init_cols) cc += static_cast<long>(ceil_div(d.m, 64 * vw)) * d.nblk;
            cb += d.nct;
        }
        if (foff > fpart_cap)
            fail(E_INTERNAL, "batched rounds: %ld single-read column sums above the %ld held",
                 foff, fpart_cap);
        h_row[s].ensure(std::max<long>(cr, 1));
        h_col[s].ensure(std::max<long>(cc, 1));
        h_comb[s].ensure(std::max<long>(cb, 1));
        h_fus[s].ensure(std::max<long>(cf, 1));
        // single-read tiles grouped by column chunks (one launch per group)
        long ir = 0, ic = 0, ib = 0, jf[kBFMaxNC];
        jf[0] = 0;
        for (int c = 1; c < kBFMaxNC; c++) jf[c] = jf[c - 1] + cfc[c - 1];
        for (int p : act) {
            const Desc& d = desc[p];
            if (d.fused) {
                for (int t = 0; t < d.nft; t++) h_fus[s].p[jf[d.fnc - 1]++] = make_int2(p, t);
            } else {
                for (int t = 0; t < d.nrt; t++) h_row[s].p[ir++] = make_int2(p, t);
            }
            if (!d.fused || init_cols) {
                const int strips = ceil_div(d.m, 64 * vw);
                for (int bk = 0; bk < d.nblk; bk++)
                    for (int st = 0; st < strips; st++)
                        h_col[s].p[ic++] = make_int4(p, st, bk, 0);
            }
            for (int c = 0; c < d.nct; c++) h_comb[s].p[ib++] = make_int2(p, c);
        }
        *nrow = static_cast<int>(cr);
        *ncol = static_cast<int>(cc);
        *ncomb = static_cast<int>(cb);
        for (int c = 0; c < kBFMaxNC; c++) nfus[c] = static_cast<int>(cfc[c]);
        h_desc[hd_slot].ensure(desc.size());
        std::memcpy(h_desc[hd_slot].p, desc.data(), sizeof(Desc) * desc.size());
        HIP_CHECK(hipMemcpyAsync(d_desc.p, h_desc[hd_slot].p, sizeof(Desc) * desc.size(),
                                 hipMemcpyHostToDevice, stream));
        hd_slot ^= 1;
        HIP_CHECK(hipMemcpyAsync(t_row.p, h_row[s].p, sizeof(int2) * std::max<long>(cr, 1),
                                 hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(t_col.p, h_col[s].p, sizeof(int4) * std::max<long>(cc, 1),
                                 hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(t_comb.p, h_comb[s].p, sizeof(int2) * std::max<long>(cb, 1),
                                 hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(t_fus.p, h_fus[s].p, sizeof(int2) * std::max<long>(cf, 1),
                                 hipMemcpyHostToDevice, stream));
    }

    template <bool NT>
    void launch_fused(int nc, int cnt, const float* Mf, const int2* tb, const Vecs& v,
                      const int* kp)
    {
        const dim3 g(cnt), b(kBFThreads);
        switch (nc) {
        case 1:
            hipLaunchKernelGGL((k_batch_fused<1, NT>), g, b, 0, stream, Mf, d_desc.p, tb, v, kp);
            break;
        case 2:
            hipLaunchKernelGGL((k_batch_fused<2, NT>), g, b, 0, stream, Mf, d_desc.p, tb, v, kp);
            break;
        case 3:
            hipLaunchKernelGGL((k_batch_fused<3, NT>), g, b, 0, stream, Mf, d_desc.p, tb, v, kp);
            break;
        default:
            hipLaunchKernelGGL((k_batch_fused<4, NT>), g, b, 0, stream, Mf, d_desc.p, tb, v, kp);
            break;
        }
        KERNEL_CHECK();
    }

    // One BCD round (round index from the device counter, advanced at its end).
    template <typename TM>
    void launch_round(int init, int nrow, int ncol, int ncomb, const int* nfus, double tol,
                      int max_iter)
    {
        const TM* Mp = reinterpret_cast<const TM*>(M.p);
        const Vecs v = vecs();
        const int ntf = nt ? 1 : 0;
        const int* kp = kdev.p;
        if (!init) {
            const float* Mf = reinterpret_cast<const float*>(M.p);
            long base = 0;
            for (int c = 0; c < kBFMaxNC; c++) {
                const int cnt = nfus[c];
                if (cnt > 0) {
                    if (sizeof(TM) != 4) fail(E_INTERNAL, "single-read batched round on FP64 bulk");
                    const int2* tb = t_fus.p + base;
                    if (nt)
                        launch_fused<true>(c + 1, cnt, Mf, tb, v, kp);
                    else
                        launch_fused<false>(c + 1, cnt, Mf, tb, v, kp);
                }
                base += cnt;
            }
        }
        if (!init && nrow > 0) {
            const int R = fine ? 1 : coarse_R;
            if (R == 1)
                hipLaunchKernelGGL((k_batch_row<TM, 1>), dim3(nrow), dim3(256), 0, stream, Mp,
                                   d_desc.p, t_row.p, v, kp, ntf);
            else if (R == 2)
                hipLaunchKernelGGL((k_batch_row<TM, 2>), dim3(nrow), dim3(256), 0, stream, Mp,
                                   d_desc.p, t_row.p, v, kp, ntf);
            else
                hipLaunchKernelGGL((k_batch_row<TM, kRows<TM>>), dim3(nrow), dim3(256), 0, stream,
                                   Mp, d_desc.p, t_row.p, v, kp, ntf);
            KERNEL_CHECK();
        }
        if (ncol > 0) {
            hipLaunchKernelGGL(k_batch_colpart<TM>, dim3(ncol), dim3(256), 0, stream, Mp,
                               d_desc.p, t_col.p, v, kp, init, 1);
            KERNEL_CHECK();
        }
        if (ncomb > 0) {
            hipLaunchKernelGGL(k_batch_combine, dim3(ncomb), dim3(256), 0, stream, d_desc.p,
                               t_comb.p, v, kp, init, tol, max_iter);
            KERNEL_CHECK();
        }
        hipLaunchKernelGGL(k_round_next, dim3(1), dim3(64), 0, stream, kdev.p);
        KERNEL_CHECK();
    }

    void round(int init, int nrow, int ncol, int ncomb, const int* nfus, double tol,
               int max_iter)
    {
        if (bulk == 0)
            launch_round<float>(init, nrow, ncol, ncomb, nfus, tol, max_iter);
        else
            launch_round<double>(init, nrow, ncol, ncomb, nfus, tol, max_iter);
    }

    // Reference initial point of every problem: alpha_0 = 0, flags cleared,
    // round counter at -1 (the init round computes beta_0).
    void reset()
    {
        const long nmax = std::max<long>(std::max(total_n, total_m), nprob);
        hipLaunchKernelGGL(k_batch_init, dim3(static_cast<int>(ceil_div(nmax, 256L))), dim3(256),
                           0, stream, ap0.p, a.p, loga.p, total_n, b.p, logb.p, total_m, done.p,
                           niter.p, nprob, kdev.p);
        KERNEL_CHECK();
    }

    // Autotuner of the coarse tiles (rows per wave of the row tiles, rows per
    // column-pass block): every candidate runs a whole batched solve of this
    // handle's problems to the requested tolerance (host wall time, the
    // solve's own graph / launch mode); the fastest is kept for the handle.
    // Each solve starts from the reference's initial point, so results depend
    // only on the tiles finally chosen.  A batch too small for coarse tiles
    // (it starts in fine tiles) keeps the defaults.
    void tune(double tol, int max_iter)
    {
        // the candidates shape only the two-pass tiles
        bool any_two_pass = false;
        for (int p = 0; p < nprob; p++) any_two_pass = any_two_pass || !takes_fused(p);
        if (!any_two_pass) {
            tuned = true;
            return;
        }
        const int R32[2] = {2, 4}, R64[2] = {1, 2};
        const int rows[3] = {256, 512, 1024};
        double best = 1e300;
        int bR = coarse_R, bB = coarse_rows, c = 0;
        for (int ri = 0; ri < 2; ri++) {
            for (int bi = 0; bi < 3; bi++, c++) {
                coarse_R = (bulk == 0 ? R32 : R64)[ri];
                coarse_rows = rows[bi];
                const auto t0 = std::chrono::steady_clock::now();
                run(tol, max_iter);
                const double ms = std::chrono::duration<double, std::milli>(
                                      std::chrono::steady_clock::now() - t0).count();
                tune_ms[c] = ms;
                if (ms < best) {
                    best = ms;
                    bR = coarse_R;
                    bB = coarse_rows;
                }
                if (started_fine) break;
            }
            if (started_fine) break;
        }
        coarse_R = bR;
        coarse_rows = bB;
        tuned = true;
    }

    // One batched solve from the reference's initial point to the stop of
    // every problem (or max_iter rounds); returns the rounds that did work.
    long long run(double tol, int max_iter)
    {
        hipStream_t s = stream;
        const int np = nprob;
        reset();

        std::vector<int> act(np);
        for (int p = 0; p < np; p++) act[p] = p;
        int slot = 0, nrow = 0, ncol = 0, ncomb = 0;
        int nfus[kBFMaxNC] = {0, 0, 0, 0};
        build_tables(act, slot, &nrow, &ncol, &ncomb, nfus, true, true);
        started_fine = fine;
        // beta_0 = optimal_beta(alpha_0 = 0) and its residual (parity 0), by
        // column tiles for every problem
        round(1, nrow, ncol, ncomb, nfus, tol, max_iter);
        // the rounds' tables: single-read problems leave the column tiles
        slot ^= 1;
        build_tables(act, slot, &nrow, &ncol, &ncomb, nfus, false);

        // K rounds per batch, two batches in flight.  Every problem is stopped
        // by round max_iter, so rounds past it (the tail of the last batch)
        // return at once.
        const int K = 8;
        int k = 0;
        long long nround = 0;
        auto launch_batch = [&](int e) {
            rounds(K, nrow, ncol, ncomb, nfus, tol, max_iter);
            const int work = std::min(K, max_iter + 1 - k);
            nround += work;
            k += work;
            HIP_CHECK(hipMemcpyAsync(h_done.p + static_cast<long>(e) * np, done.p,
                                     sizeof(int) * np, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipEventRecord(ev[e], s));
        };
        std::deque<int> q;
        try {
            launch_batch(0);
            q.push_back(0);
            if (k <= max_iter) {
                launch_batch(1);
                q.push_back(1);
            }
            while (!q.empty()) {
                const int e = q.front();
                q.pop_front();
                HIP_CHECK(hipEventSynchronize(ev[e]));
                const int* dn = h_done.p + static_cast<long>(e) * np;
                std::vector<int> still;
                still.reserve(act.size());
                for (int p : act)
                    if (!dn[p]) still.push_back(p);
                if (still.empty() || k > max_iter) {
                    // every problem stopped, or every round up to max_iter is queued
                    if (still.empty()) break;
                    continue;
                }
                if (still.size() != act.size()) {
                    act.swap(still);
                    slot ^= 1;
                    build_tables(act, slot, &nrow, &ncol, &ncomb, nfus, false);
                }
                launch_batch(e);
                q.push_back(e);
            }
            HIP_CHECK(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipStreamSynchronize(s);
            throw;
        }
        return nround;
    }

    // K rounds as one graph launch.  The graph is captured from the stream
    // once per configuration (grid sizes, tile mode, load mode; tables and
    // descriptors are device buffers refreshed outside it) and replayed;
    // the round index lives on the device, so replays need no parameter
    // updates.  The rounds contain kernels only (no memset nodes).
    void rounds(int K, int nrow, int ncol, int ncomb, const int* nfus, double tol,
                int max_iter)
    {
        if (!use_graph) {
            for (int q = 0; q < K; q++) round(0, nrow, ncol, ncomb, nfus, tol, max_iter);
            return;
        }
        const int key[kGraphKey] = {K,          nrow,     ncol,     ncomb,       fine ? 1 : 0,
                                    nt ? 1 : 0, max_iter, coarse_R, coarse_rows, nfus[0],
                                    nfus[1],    nfus[2],  nfus[3]};
        if (!gexec || !std::equal(key, key + kGraphKey, g_key) || tol != g_tol) {
            if (gexec) {
                HIP_CHECK(hipGraphExecDestroy(gexec));
                gexec = nullptr;
            }
            hipGraph_t g = nullptr;
            HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            try {
                for (int q = 0; q < K; q++) round(0, nrow, ncol, ncomb, nfus, tol, max_iter);
            } catch (...) {
                (void)hipStreamEndCapture(stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIP_CHECK(hipStreamEndCapture(stream, &g));
            const hipError_t e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIP_CHECK(e);
            std::copy(key, key + kGraphKey, g_key);
            g_tol = tol;
            graph_builds++;
        }
        HIP_CHECK(hipGraphLaunch(gexec, stream));
        graph_launches++;
    }
};

namespace {

#define CAPI_BEGIN try {
#define CAPI_END                                                                          \
    }                                                                                     \
    catch (const regot::Error& e) { return set_error(e.code, e.what()); }                 \
    catch (const std::bad_alloc&) { return set_error(E_NOMEM, "host allocation failed"); } \
    catch (const std::exception& e) { return set_error(E_INTERNAL, e.what()); }            \
    return OK;

}  // namespace

extern "C" {

int regot_batch_create(regot_batch** out, int nprob, const int* n, const int* m,
                       const void* const* M, int M_dtype, int bulk, int M_device,
                       const double* const* a, const double* const* b, const double* reg,
                       int device)
{
    CAPI_BEGIN
    if (!out || nprob <= 0 || !n || !m || !M || !a || !b || !reg)
        fail(E_INVALID, "regot_batch_create: null argument or nprob <= 0");
    if ((M_dtype != REGOT_F32 && M_dtype != REGOT_F64) || (bulk != REGOT_F32 && bulk != REGOT_F64))
        fail(E_INVALID, "regot_batch_create: dtype must be REGOT_F32 or REGOT_F64");
    *out = nullptr;
    int ndev = 0;
    HIP_CHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) fail(E_INVALID, "device %d not available", device);
    HIP_CHECK(hipSetDevice(device));
    auto B = std::make_unique<regot_batch>();
    B->device = device;
    B->nprob = nprob;
    B->bulk = bulk;
    B->coarse_R = bulk == REGOT_F32 ? 4 : 2;
    HIP_CHECK(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    const int esz = bulk == REGOT_F32 ? 4 : 8;
    const int R = bulk == REGOT_F32 ? 4 : 2;
    B->desc.resize(nprob);
    long moff = 0, aoff = 0, boff = 0, rt = 0, poff = 0, ct = 0;
    long fbig = 0, fsum_m = 0;
    int fmax_m = 0;
    for (int p = 0; p < nprob; p++) {
        if (n[p] < 1 || m[p] < 1 || !M[p] || !a[p] || !b[p] || !(reg[p] > 0.0))
            fail(E_INVALID, "problem %d: bad sizes, null pointer or reg <= 0", p);
        if (static_cast<long>(n[p]) * m[p] > (1L << 31))
            fail(E_INVALID, "problem %d: n*m above 2^31", p);
        Desc& d = B->desc[p];
        d.n = n[p];
        d.m = m[p];
        d.ldm = static_cast<int>(ceil_div(m[p], 256 / esz) * (256 / esz));
        d.moff = moff;
        moff += static_cast<long>(d.n) * d.ldm;
        d.aoff = static_cast<int>(aoff);
        d.boff = static_cast<int>(boff);
        aoff += d.n;
        boff += d.m;
        // partial slots sized for fine tiles (the larger count)
        d.rpt = 4 * R;
        d.nrt = ceil_div(d.n, d.rpt);
        d.rt0 = static_cast<int>(rt);
        rt += ceil_div(d.n, 4);
        d.nblk = ceil_div(d.n, kBlockRows);
        d.rpb = ceil_div(d.n, d.nblk);
        d.poff = poff;
        poff += static_cast<long>(ceil_div(d.n, kBlockRowsFine)) * d.m;
        d.nct = ceil_div(d.m, kCombCols);
        d.ct0 = static_cast<int>(ct);
        ct += d.nct;
        d.reg = reg[p];
        d.k = 1.4426950408889634074 / reg[p];
        double sa = 0.0, sb = 0.0, sbf = 0.0, mina = a[p][0], minb = b[p][0];
        for (int i = 0; i < d.n; i++) {
            sa += a[p][i];
            mina = std::min(mina, a[p][i]);
        }
        for (int j = 0; j < d.m; j++) {
            if (j < d.m - 1) sb += b[p][j];
            sbf += b[p][j];
            minb = std::min(minb, b[p][j]);
        }
        d.sum_a = sa;
        d.sum_b1 = sb;
        d.fused = d.fnc = d.fr = d.nft = 0;
        d.fpoff = 0;
        // single-read rounds: FP32 bulk, 1024 < m <= 4096, and every column
        // sum far above the terms the FP32 exponential flushes
        const bool fok = bulk == REGOT_F32 && d.m >= kBFMinM && d.m <= kBFMaxNC * kBFSpan &&
                         mina > 0.0 && minb > 0.0 && mina * minb >= 0x1p-60 * sa * sbf;
        B->fused_ok.push_back(fok ? 1 : 0);
        if (fok) {
            fbig += static_cast<long>(ceil_div(d.n, kBFMaxRows)) * d.m;
            fsum_m += d.m;
            fmax_m = std::max(fmax_m, d.m);
        }
        B->total_elems += static_cast<long long>(d.n) * d.m;
    }
    // Single-read column sums: with fr = kBFMaxRows the tables hold at most
    // fbig; below it fr >= max(kBFMinRows, rows / kBFTiles) bounds the tiles'
    // column sums by kBFTiles max m + sum m (build_tables checks the bound).
    B->fpart_cap = std::max(fbig, static_cast<long>(kBFTiles) * fmax_m + fsum_m);
    if (aoff > 0x7fffffffL || boff > 0x7fffffffL || rt > 0x7fffffffL)
        fail(E_INVALID, "batch too large (sum of n or m above 2^31)");
    B->total_n = aoff;
    B->total_m = boff;
    B->total_dual = aoff + boff - nprob;
    B->M.alloc(static_cast<size_t>(moff) * esz);
    B->d_desc.alloc(nprob);
    HIP_CHECK(hipMemcpyAsync(B->d_desc.p, B->desc.data(), sizeof(Desc) * nprob,
                             hipMemcpyHostToDevice, B->stream));
    // pack M (convert / pad on the device; host sources go through one staging copy)
    const int in_esz = M_dtype == REGOT_F32 ? 4 : 8;
    DBuf<unsigned char> stage;
    for (int p = 0; p < nprob; p++) {
        const Desc& d = B->desc[p];
        const void* src = M[p];
        if (!M_device) {
            stage.ensure(static_cast<size_t>(d.n) * d.m * in_esz);
            HIP_CHECK(hipMemcpyAsync(stage.p, M[p], static_cast<size_t>(d.n) * d.m * in_esz,
                                     hipMemcpyHostToDevice, B->stream));
            src = stage.p;
        }
        unsigned char* dst = B->M.p + static_cast<size_t>(d.moff) * esz;
        const long total = static_cast<long>(d.n) * d.ldm;
        const int g = static_cast<int>(std::min<long>(ceil_div(total, 256L), 65536L));
        if (M_dtype == REGOT_F32 && bulk == REGOT_F32)
            hipLaunchKernelGGL((k_pack<float, float>), dim3(g), dim3(256), 0, B->stream,
                               static_cast<const float*>(src), static_cast<long>(d.m),
                               reinterpret_cast<float*>(dst), static_cast<long>(d.ldm), d.n, d.m);
        else if (M_dtype == REGOT_F64 && bulk == REGOT_F32)
            hipLaunchKernelGGL((k_pack<double, float>), dim3(g), dim3(256), 0, B->stream,
                               static_cast<const double*>(src), static_cast<long>(d.m),
                               reinterpret_cast<float*>(dst), static_cast<long>(d.ldm), d.n, d.m);
        else if (M_dtype == REGOT_F32)
            hipLaunchKernelGGL((k_pack<float, double>), dim3(g), dim3(256), 0, B->stream,
                               static_cast<const float*>(src), static_cast<long>(d.m),
                               reinterpret_cast<double*>(dst), static_cast<long>(d.ldm), d.n, d.m);
        else
            hipLaunchKernelGGL((k_pack<double, double>), dim3(g), dim3(256), 0, B->stream,
                               static_cast<const double*>(src), static_cast<long>(d.m),
                               reinterpret_cast<double*>(dst), static_cast<long>(d.ldm), d.n, d.m);
        KERNEL_CHECK();
        if (!M_device) HIP_CHECK(hipStreamSynchronize(B->stream));   // stage is reused
    }
    // marginals
    std::vector<double> ha(aoff), hb(boff);
    for (int p = 0; p < nprob; p++) {
        std::memcpy(ha.data() + B->desc[p].aoff, a[p], sizeof(double) * n[p]);
        std::memcpy(hb.data() + B->desc[p].boff, b[p], sizeof(double) * m[p]);
    }
    B->a.alloc(aoff);
    B->b.alloc(boff);
    B->loga.alloc(aoff);
    B->logb.alloc(boff);
    B->ap0.alloc(aoff);
    B->ap1.alloc(aoff);
    B->beta.alloc(boff);
    B->rpart.alloc(3 * rt);
    B->part.alloc(poff);
    B->fpart.alloc(std::max(B->fpart_cap, 1L));
    B->cpart.alloc(4 * ct);
    B->bm.alloc(nprob);
    B->fval.alloc(nprob);
    B->gval.alloc(nprob);
    B->done.alloc(nprob);
    B->niter.alloc(nprob);
    B->kdev.alloc(1);
    HIP_CHECK(hipMemcpyAsync(B->a.p, ha.data(), sizeof(double) * aoff, hipMemcpyHostToDevice,
                             B->stream));
    HIP_CHECK(hipMemcpyAsync(B->b.p, hb.data(), sizeof(double) * boff, hipMemcpyHostToDevice,
                             B->stream));
    // tables sized for fine tiles (the larger counts)
    long tr = 0, tc = 0, tb = 0;
    for (const Desc& d : B->desc) {
        tr += ceil_div(d.n, 4);
        tc += static_cast<long>(ceil_div(d.m, 64 * (bulk == REGOT_F32 ? 4 : 2))) *
              ceil_div(d.n, kBlockRowsFine);
        tb += d.nct;
    }
    B->t_row.alloc(tr);
    B->t_col.alloc(tc);
    B->t_comb.alloc(tb);
    // single-read tiles: at most ceil(n / kBFMinRows) per problem
    long tf = 0;
    for (const Desc& d : B->desc) tf += ceil_div(d.n, kBFMinRows);
    B->t_fus.alloc(tf);
    B->h_done.ensure(2L * nprob);
    for (auto& e : B->ev) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_CHECK(hipStreamSynchronize(B->stream));
    *out = B.release();
    CAPI_END
}

int regot_batch_set_option(regot_batch* B, int option, long long value)
{
    CAPI_BEGIN
    if (!B) fail(E_INVALID, "null batch");
    switch (option) {
        case REGOT_BATCH_GRAPH:
            B->use_graph = value != 0;
            break;
        case REGOT_BATCH_FUSED:
            B->use_fused = value != 0;
            break;
        case REGOT_BATCH_AUTOTUNE:
            B->autotune = value != 0;
            if (!B->autotune) {
                B->tuned = false;
                B->coarse_R = B->bulk == REGOT_F32 ? 4 : 2;
                B->coarse_rows = kBlockRows;
            }
            break;
        default:
            fail(E_INVALID, "unknown batch option %d", option);
    }
    CAPI_END
}

int regot_batch_stats(const regot_batch* B, double* out)
{
    CAPI_BEGIN
    if (!B || !out) fail(E_INVALID, "null argument");
    out[0] = B->coarse_R;
    out[1] = B->coarse_rows;
    out[2] = B->tuned ? 1.0 : 0.0;
    out[3] = static_cast<double>(B->graph_builds);
    out[4] = static_cast<double>(B->graph_launches);
    for (int c = 0; c < 6; c++) out[5 + c] = B->tune_ms[c];
    int nf = 0;
    for (int p = 0; p < B->nprob; p++) nf += B->takes_fused(p) ? 1 : 0;
    out[11] = nf;
    CAPI_END
}

int regot_batch_round_kinds(const regot_batch* B, int* kind)
{
    CAPI_BEGIN
    if (!B || !kind) fail(E_INVALID, "null argument");
    for (int p = 0; p < B->nprob; p++) kind[p] = B->takes_fused(p) ? 1 : 0;
    CAPI_END
}

int regot_batch_destroy(regot_batch* B)
{
    CAPI_BEGIN
    delete B;
    CAPI_END
}

int regot_batch_info(const regot_batch* B, int* nprob, long long* total_elems,
                     long long* dual_len)
{
    CAPI_BEGIN
    if (!B) fail(E_INVALID, "null batch");
    if (nprob) *nprob = B->nprob;
    if (total_elems) *total_elems = B->total_elems;
    if (dual_len) *dual_len = B->total_dual;
    CAPI_END
}

int regot_batch_solve_bcd(regot_batch* B, double tol, int max_iter, int* niter, double* obj,
                          double* mar_err, double* dual, long long* rounds)
{
    CAPI_BEGIN
    if (!B) fail(E_INVALID, "null batch");
    if (max_iter < 0) fail(E_INVALID, "max_iter < 0");
    HIP_CHECK(hipSetDevice(B->device));
    hipStream_t s = B->stream;
    const int np = B->nprob;
    if (B->autotune && !B->tuned) B->tune(tol, max_iter);
    const long long nround = B->run(tol, max_iter);

    if (niter) HIP_CHECK(hipMemcpyAsync(niter, B->niter.p, sizeof(int) * np, hipMemcpyDeviceToHost, s));
    if (obj) HIP_CHECK(hipMemcpyAsync(obj, B->fval.p, sizeof(double) * np, hipMemcpyDeviceToHost, s));
    if (mar_err)
        HIP_CHECK(hipMemcpyAsync(mar_err, B->gval.p, sizeof(double) * np, hipMemcpyDeviceToHost, s));
    if (dual) {
        DBuf<double> dd(static_cast<size_t>(B->total_dual));
        hipLaunchKernelGGL(k_batch_dual, dim3(np), dim3(256), 0, s, B->d_desc.p, B->vecs(), dd.p);
        KERNEL_CHECK();
        HIP_CHECK(hipMemcpyAsync(dual, dd.p, sizeof(double) * B->total_dual,
                                 hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
    HIP_CHECK(hipStreamSynchronize(s));
    if (rounds) *rounds = nround;
    CAPI_END
}

}  // extern "C"
