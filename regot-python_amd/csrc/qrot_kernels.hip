// qrot_kernels.hip -- streaming passes of the quadratically regularised OT
// (QROT) dual over the cost matrix M (FP32 or FP64 storage, FP64 arithmetic).
//
// With D_ij = (alpha_i + beta_j - M_ij)_+ (src/qrot_problem.cpp:37-96):
//   k_qrot_tile       row sums, column sums and sum of squares of
//                     W_ij = ((D_ij / div) x_i) y_j (x, y optional weights;
//                     div = 1 for the dual, reg for the plan D / reg);
//   k_qrot_round_tile the rounded plan P_ij = W_ij + ea_i eb_j / l1 of
//                     approx_proj (src/approx_proj.h:6-26) with <P, M> and
//                     ||P||^2, P optionally written;
//   k_qrot_pl_tile    one Newton step of the piecewise-linear roots behind
//                     optimal_beta / optimal_alpha (src/qrot_problem.cpp:365-418):
//                     per column (or row) the sum and count of y = M_ij - alpha_i
//                     (or M_ij - beta_j) below the current iterate;
//   k_qrot_pl_update  the step x <- (c + S) / K per column or row.
//   k_qrot_mask       the generalised Hessian's 0/1 pattern sigma_ij =
//                     [alpha_i + beta_j - M_ij > 0] (src/qrot_hess.cpp:42-96)
//                     as a bitmask, one bit per entry (n m / 8 bytes);
//   k_qrot_sig        sigma y (row sums) and sigma' x (column sums) read from
//                     the bitmask (src/qrot_hess.cpp:99-125).
// Tiles are 256-column strips (64 lanes x 4 adjacent columns) over blocks of
// `rows` rows, one row per wave at a time; every reduction has a fixed order.
#include <cmath>

#include "qrot.hpp"

namespace regot {
namespace {

__device__ __forceinline__ double wsum64(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// One DPP step on a double (both halves moved with the same control).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

// Sum over each 16-lane row of the wave, left in every lane of the row:
// quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror
// (fixed order, no LDS).
__device__ __forceinline__ double rsum16(double v)
{
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    v += dpp_f64<0x141>(v);
    v += dpp_f64<0x140>(v);
    return v;
}

// The lane's 4 adjacent entries of a row: one 16-byte load (FP32) or two
// (FP64) when all 4 are in range (rows are 16-byte aligned and c0 is a
// multiple of 4), else element by element.
__device__ __forceinline__ void load4(const float* row, int c0, int m, double* v)
{
    if (c0 + 3 < m) {
        const float4 x = *reinterpret_cast<const float4*>(row + c0);
        v[0] = x.x;
        v[1] = x.y;
        v[2] = x.z;
        v[3] = x.w;
        return;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = (c0 + q < m) ? static_cast<double>(row[c0 + q]) : 0.0;
}

__device__ __forceinline__ void load4(const double* row, int c0, int m, double* v)
{
    if (c0 + 3 < m) {
        const double2 x = *reinterpret_cast<const double2*>(row + c0);
        const double2 y = *reinterpret_cast<const double2*>(row + c0 + 2);
        v[0] = x.x;
        v[1] = x.y;
        v[2] = y.x;
        v[3] = y.y;
        return;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = (c0 + q < m) ? row[c0 + q] : 0.0;
}

// DIV: divide D by div (the plan D / reg of the rounding passes, in the
// reference's order of operations); without it div is 1 and skipped.
template <typename TM, bool WT, bool DIV>
__global__ __launch_bounds__(256) void k_qrot_tile(const TM* __restrict__ M, long ldm, int n,
                                                   int m, const double* __restrict__ alpha,
                                                   const double* __restrict__ beta, double div,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ y, int rows,
                                                   double* __restrict__ rowpart,
                                                   double* __restrict__ colpart,
                                                   double* __restrict__ sqpart)
{
    __shared__ double red[4][257];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int strip = blockIdx.x, blk = blockIdx.y;
    const int c0 = strip * 256 + lane * 4;
    const int r0 = blk * rows;
    const int r1 = min(n, r0 + rows);
    double bq[4], yq[4], cs[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        bq[q] = (c0 + q < m) ? beta[c0 + q] : 0.0;
        yq[q] = (WT && y != nullptr && c0 + q < m) ? y[c0 + q] : 1.0;
    }
    double sq = 0.0;
    for (int r = r0 + wave; r < r1; r += 4) {
        const double ai = alpha[r];
        const double xi = (WT && x != nullptr) ? x[r] : 1.0;
        double mv[4];
        load4(M + static_cast<long>(r) * ldm, c0, m, mv);
        double rs = 0.0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (c0 + q < m) {
                const double c = (ai + bq[q]) - mv[q];
                double d = c > 0.0 ? c : 0.0;
                if (DIV) d = d / div;
                if (WT) {
                    if (x != nullptr) d = d * xi;
                    if (y != nullptr) d = d * yq[q];
                }
                rs += d;
                cs[q] += d;
                sq += d * d;
            }
        }
        rs = wsum64(rs);
        if (lane == 0 && rowpart) rowpart[static_cast<long>(strip) * n + r] = rs;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) red[wave][lane * 4 + q] = cs[q];
    sq = wsum64(sq);
    if (lane == 0) red[wave][256] = sq;
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int j = c0 + q;
            const double v = ((red[0][lane * 4 + q] + red[1][lane * 4 + q]) +
                              red[2][lane * 4 + q]) + red[3][lane * 4 + q];
            if (j < m && colpart) colpart[static_cast<long>(blk) * m + j] = v;
        }
        if (lane == 0 && sqpart)
            sqpart[static_cast<long>(blk) * gridDim.x + strip] =
                ((red[0][256] + red[1][256]) + red[2][256]) + red[3][256];
    }
}

template <typename TM>
__global__ __launch_bounds__(256) void k_qrot_round_tile(
    const TM* __restrict__ M, long ldm, int n, int m, const double* __restrict__ alpha,
    const double* __restrict__ beta, double reg, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ ea, const double* __restrict__ eb,
    const double* __restrict__ l1p, int rows, double* __restrict__ P, long ldp,
    double* __restrict__ scal)
{
    __shared__ double ss[2][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int strip = blockIdx.x, blk = blockIdx.y;
    const int c0 = strip * 256 + lane * 4;
    const int r0 = blk * rows;
    const int r1 = min(n, r0 + rows);
    const double l1 = *l1p;
    double bq[4], yq[4], eq[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const bool in = c0 + q < m;
        bq[q] = in ? beta[c0 + q] : 0.0;
        yq[q] = in ? y[c0 + q] : 0.0;
        eq[q] = in ? eb[c0 + q] : 0.0;
    }
    double mp = 0.0, sq = 0.0;
    for (int r = r0 + wave; r < r1; r += 4) {
        const double ai = alpha[r], xi = x[r], ei = ea[r];
        double mv[4];
        load4(M + static_cast<long>(r) * ldm, c0, m, mv);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int j = c0 + q;
            if (j < m) {
                const double c = (ai + bq[q]) - mv[q];
                const double v = (((c > 0.0 ? c : 0.0) / reg) * xi) * yq[q] + ei * eq[q] / l1;
                if (P) P[static_cast<long>(r) * ldp + j] = v;
                mp += v * mv[q];
                sq += v * v;
            }
        }
    }
    mp = wsum64(mp);
    sq = wsum64(sq);
    if (lane == 0) {
        ss[0][wave] = mp;
        ss[1][wave] = sq;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const long wg = static_cast<long>(blk) * gridDim.x + strip;
        scal[2 * wg] = ((ss[0][0] + ss[0][1]) + ss[0][2]) + ss[0][3];
        scal[2 * wg + 1] = ((ss[1][0] + ss[1][1]) + ss[1][2]) + ss[1][3];
    }
}

// COLS: per column j, y_i = M_ij - pot_i over the block's rows; else per row
// i, y_j = M_ij - pot_j over the strip's columns.  Sum and count of y < xc.
template <typename TM, bool COLS>
__global__ __launch_bounds__(256) void k_qrot_pl_tile(const TM* __restrict__ M, long ldm, int n,
                                                      int m, const double* __restrict__ pot,
                                                      const double* __restrict__ xc,
                                                      const int* __restrict__ done, int rows,
                                                      double* __restrict__ spart,
                                                      int* __restrict__ kpart)
{
    __shared__ double rs_[4][256];
    __shared__ int rk_[4][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int strip = blockIdx.x, blk = blockIdx.y;
    const int c0 = strip * 256 + lane * 4;
    const int r0 = blk * rows;
    const int r1 = min(n, r0 + rows);
    double pc[4], thr[4], S[4] = {0.0, 0.0, 0.0, 0.0};
    int K[4] = {0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const bool in = c0 + q < m;
        pc[q] = (!COLS && in) ? pot[c0 + q] : 0.0;
        thr[q] = (COLS && in) ? xc[c0 + q] : 0.0;
    }
    for (int r = r0 + wave; r < r1; r += 4) {
        double mv[4];
        load4(M + static_cast<long>(r) * ldm, c0, m, mv);
        if (COLS) {
            const double pr = pot[r];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const double yv = mv[q] - pr;
                if (c0 + q < m && yv < thr[q]) {
                    S[q] += yv;
                    K[q] += 1;
                }
            }
        } else {
            if (done[r]) continue;   // wave-uniform: r is the same on every lane
            const double xr = xc[r];
            double s = 0.0;
            int kk = 0;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const double yv = mv[q] - pc[q];
                if (c0 + q < m && yv < xr) {
                    s += yv;
                    kk += 1;
                }
            }
            s = wsum64(s);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) kk += __shfl_xor(kk, o);
            if (lane == 0) {
                spart[static_cast<long>(strip) * n + r] = s;
                kpart[static_cast<long>(strip) * n + r] = kk;
            }
        }
    }
    if (COLS) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            rs_[wave][lane * 4 + q] = S[q];
            rk_[wave][lane * 4 + q] = K[q];
        }
        __syncthreads();
        if (wave == 0) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int j = c0 + q;
                if (j >= m) continue;
                const int t = lane * 4 + q;
                spart[static_cast<long>(blk) * m + j] =
                    ((rs_[0][t] + rs_[1][t]) + rs_[2][t]) + rs_[3][t];
                kpart[static_cast<long>(blk) * m + j] =
                    ((rk_[0][t] + rk_[1][t]) + rk_[2][t]) + rk_[3][t];
            }
        }
    }
}

// Per entry t < count (a column or a row): S = sum of the nparts partials
// (fixed order), K likewise; x_t <- (reg w_t + S) / K.  The iteration stops
// for t once K no longer shrinks (the active set is final; from the right
// Newton's iterates only shrink it), counted in *open.
__global__ void k_qrot_pl_update(const double* __restrict__ spart, const int* __restrict__ kpart,
                                 int nparts, long stride, int count,
                                 const double* __restrict__ w, double reg,
                                 double* __restrict__ xc, int* __restrict__ kprev,
                                 int* __restrict__ done, int* __restrict__ open)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= count || done[t]) return;
    double S = 0.0;
    int K = 0;
    for (int s = 0; s < nparts; s++) {
        S += spart[static_cast<long>(s) * stride + t];
        K += kpart[static_cast<long>(s) * stride + t];
    }
    if (kprev[t] >= 0 && K >= kprev[t]) {
        done[t] = 1;
        return;
    }
    if (K > 0) xc[t] = (reg * w[t] + S) / K;
    kprev[t] = K;
    atomicAdd(open, 1);
}

__global__ void k_qrot_pl_init(double* xc, int* kprev, int* done, int count, int* open)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t == 0) open[0] = 0;
    if (t >= count) return;
    xc[t] = INFINITY;
    kprev[t] = -1;
    done[t] = 0;
}

// APDAGD's / PDAAM's primal plan: P = cold P + tau (alpha_i + beta_j - M_ij)_+ / reg
// (INIT: P = the plan itself, P not read), written back, with row sums per
// strip, column sums per block and <M, P>, ||P||^2 per workgroup
// (mpart, spart[blk * strips + strip]).
template <typename TM, bool INIT>
__global__ __launch_bounds__(256) void k_qrot_apd_plan(
    const TM* __restrict__ M, long ldm, int n, int m, const double* __restrict__ alpha,
    const double* __restrict__ beta, double reg, double cold, double tau,
    double* __restrict__ P, long ldp,
    int rows, double* __restrict__ rowpart, double* __restrict__ colpart,
    double* __restrict__ mpart, double* __restrict__ spart)
{
    __shared__ double red[4][258];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int strip = blockIdx.x, blk = blockIdx.y;
    const int c0 = strip * 256 + lane * 4;
    const int r0 = blk * rows;
    const int r1 = min(n, r0 + rows);
    double bq[4], cs[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 4; q++) bq[q] = (c0 + q < m) ? beta[c0 + q] : 0.0;
    double mp = 0.0, sq = 0.0;
    for (int r = r0 + wave; r < r1; r += 4) {
        const double ai = alpha[r];
        double mv[4];
        load4(M + static_cast<long>(r) * ldm, c0, m, mv);
        double* Pr = P + static_cast<long>(r) * ldp;
        double rs = 0.0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int j = c0 + q;
            if (j < m) {
                const double c = (ai + bq[q]) - mv[q];
                const double t = (c > 0.0 ? c : 0.0) / reg;
                const double v = INIT ? t : cold * Pr[j] + tau * t;
                Pr[j] = v;
                rs += v;
                cs[q] += v;
                mp += mv[q] * v;
                sq += v * v;
            }
        }
        rs = wsum64(rs);
        if (lane == 0) rowpart[static_cast<long>(strip) * n + r] = rs;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) red[wave][lane * 4 + q] = cs[q];
    mp = wsum64(mp);
    sq = wsum64(sq);
    if (lane == 0) {
        red[wave][256] = mp;
        red[wave][257] = sq;
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int j = c0 + q;
            const int t = lane * 4 + q;
            if (j < m)
                colpart[static_cast<long>(blk) * m + j] =
                    ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
        }
        if (lane == 0) {
            const long wg = static_cast<long>(blk) * gridDim.x + strip;
            mpart[wg] = ((red[0][256] + red[1][256]) + red[2][256]) + red[3][256];
            spart[wg] = ((red[0][257] + red[1][257]) + red[2][257]) + red[3][257];
        }
    }
}

// part[w] = sum over the workgroup's slice of (rs - a)^2 (t < n) and
// (cs - b)^2 (t < m), fixed order.
__global__ __launch_bounds__(256) void k_qrot_resid2(const double* __restrict__ rs,
                                                     const double* __restrict__ a, int n,
                                                     const double* __restrict__ cs,
                                                     const double* __restrict__ b, int m,
                                                     double* __restrict__ part)
{
    __shared__ double sm[4];
    const int t = blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (t < n) v += (rs[t] - a[t]) * (rs[t] - a[t]);
    if (t < m) v += (cs[t] - b[t]) * (cs[t] - b[t]);
    v = wsum64(v);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((sm[0] + sm[1]) + sm[2]) + sm[3];
}

__global__ void k_qrot_zero_int(int* p) { p[0] = 0; }

// out[t] = sum over s of part[s * stride + t] (fixed order), t < count.
__global__ void k_qrot_fold(const double* __restrict__ part, int nparts, long stride, int count,
                            double* __restrict__ out)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= count) return;
    double v = 0.0;
    for (int s = 0; s < nparts; s++) v += part[static_cast<long>(s) * stride + t];
    out[t] = v;
}

// One workgroup: out[0] = sum of part[0..count) (fixed order).
__global__ __launch_bounds__(256) void k_qrot_sum1(const double* __restrict__ part, int count,
                                                   double* __restrict__ out)
{
    __shared__ double sm[4];
    double v = 0.0;
    for (int i = threadIdx.x; i < count; i += 256) v += part[i];
    v = wsum64(v);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = ((sm[0] + sm[1]) + sm[2]) + sm[3];
}

template <typename TM>
__global__ __launch_bounds__(256) void k_qrot_plan(const TM* __restrict__ M, long ldm, int m,
                                                   const double* __restrict__ alpha,
                                                   const double* __restrict__ beta, double reg,
                                                   double* __restrict__ plan, long r0)
{
    const long i = r0 + blockIdx.y;
    const double ai = alpha[i];
    const TM* Mi = M + i * ldm;
    double* Pi = plan + static_cast<long>(blockIdx.y) * m;
    for (int j = blockIdx.x * 256 + threadIdx.x; j < m; j += gridDim.x * 256) {
        const double c = (ai + beta[j]) - static_cast<double>(Mi[j]);
        Pi[j] = (c > 0.0 ? c : 0.0) / reg;
    }
}

// Workgroup w of the finish: partial sums over its 256-entry slices of
// alpha'a, beta'b, 1'alpha, 1'beta (fixed order) into part[4 w ..].
__global__ __launch_bounds__(256) void k_qrot_fin_part(
    const double* __restrict__ alpha, const double* __restrict__ beta,
    const double* __restrict__ a, const double* __restrict__ b, int n, int m,
    double* __restrict__ part)
{
    __shared__ double sm[4][4];
    const int t = blockIdx.x * 256 + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (t < n) {
        v[0] = alpha[t] * a[t];
        v[2] = alpha[t];
    }
    if (t < m) {
        v[1] = beta[t] * b[t];
        v[3] = beta[t];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        v[q] = wsum64(v[q]);
        if (lane == 0) sm[wave][q] = v[q];
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int q = threadIdx.x;
        part[4 * blockIdx.x + q] = ((sm[0][q] + sm[1][q]) + sm[2][q]) + sm[3][q];
    }
}

// Gradient (if requested) and its squared norm per workgroup; every
// workgroup folds the nw partials of k_qrot_fin_part in the same order, so d
// is the same everywhere.  gpart[w] = sum of g^2 over the slice.
__global__ __launch_bounds__(256) void k_qrot_fin_grad(
    const double* __restrict__ part, int nw, const double* __restrict__ a,
    const double* __restrict__ b, int n, int m, double reg, double tau,
    const double* __restrict__ rsum, const double* __restrict__ csum,
    double* __restrict__ grad, double* __restrict__ gpart)
{
    __shared__ double sm[4];
    double sa = 0.0, sb = 0.0;
    for (int w = 0; w < nw; w++) {
        sa += part[4 * w + 2];
        sb += part[4 * w + 3];
    }
    const double d = sa - sb;
    const double td = tau > 0.0 ? tau * d : 0.0;
    const int t = blockIdx.x * 256 + threadIdx.x;
    double g2 = 0.0;
    if (t < n) {
        const double g = (rsum[t] - reg * a[t]) + td;
        grad[t] = g;
        g2 += g * g;
    }
    if (t < m) {
        const double g = (csum[t] - reg * b[t]) - td;
        grad[n + t] = g;
        g2 += g * g;
    }
    g2 = wsum64(g2);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = g2;
    __syncthreads();
    if (threadIdx.x == 0) gpart[blockIdx.x] = ((sm[0] + sm[1]) + sm[2]) + sm[3];
}

// out[0] = f = 0.5 sq - reg (alpha'a + beta'b) + tau/2 d^2, out[1] = ||g||^2.
__global__ __launch_bounds__(64) void k_qrot_fin_total(const double* __restrict__ part, int nw,
                                                       const double* __restrict__ gpart,
                                                       int ngw, double reg, double tau,
                                                       const double* __restrict__ sq,
                                                       double* __restrict__ out)
{
    if (threadIdx.x != 0) return;
    double tot[4] = {0.0, 0.0, 0.0, 0.0};
    for (int w = 0; w < nw; w++)
        for (int q = 0; q < 4; q++) tot[q] += part[4 * w + q];
    const double d = tot[2] - tot[3];
    double f = 0.5 * sq[0] - reg * (tot[0] + tot[1]);
    if (tau > 0.0) f += 0.5 * tau * d * d;
    double g2 = 0.0;
    for (int w = 0; w < ngw; w++) g2 += gpart[w];
    out[0] = f;
    out[1] = g2;
}

// The bitmask shares the tiles' lane layout: for row r and 256-column strip
// st, word q of mask[(r * nstrip + st) * 4 + q] holds in bit l the entry of
// column st * 256 + 4 l + q (the lane-l, q-th column of the strip), so a wave
// writes a row's strip with four ballots and reads it back as four
// wave-uniform words.  Columns past m read as 0.
template <typename TM>
__global__ __launch_bounds__(256) void k_qrot_mask(const TM* __restrict__ M, long ldm, int n,
                                                   int m, const double* __restrict__ alpha,
                                                   const double* __restrict__ beta, int rows,
                                                   unsigned long long* __restrict__ mask)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int strip = blockIdx.x, nstrip = gridDim.x;
    const int c0 = strip * 256 + lane * 4;
    const int r0 = blockIdx.y * rows;
    const int r1 = min(n, r0 + rows);
    double bq[4];
#pragma unroll
    for (int q = 0; q < 4; q++) bq[q] = (c0 + q < m) ? beta[c0 + q] : 0.0;
    for (int r = r0 + wave; r < r1; r += 4) {
        const double ai = alpha[r];
        double mv[4];
        load4(M + static_cast<long>(r) * ldm, c0, m, mv);
        unsigned long long w[4];
#pragma unroll
        for (int q = 0; q < 4; q++) w[q] = __ballot(c0 + q < m && (ai + bq[q]) - mv[q] > 0.0);
        if (lane < 4) {
            const unsigned long long v = lane == 0 ? w[0] : lane == 1 ? w[1] : lane == 2 ? w[2] : w[3];
            mask[(static_cast<long>(r) * nstrip + strip) * 4 + lane] = v;
        }
    }
}

// ROWS: rowpart[(strip * 4 + k) * n + r] = sum_j sigma_rj y_j over the k-th
// 64 columns of the strip (16 lanes; y null: y = 1); COLS: colpart[blk * m + j] = sum_r sigma_rj x_r over the row block
// (x null: x = 1).  Same grid and fixed reduction order as k_qrot_tile.
template <bool ROWS, bool COLS>
__global__ __launch_bounds__(256) void k_qrot_sig(const unsigned long long* __restrict__ mask,
                                                  int n, int m, const double* __restrict__ x,
                                                  const double* __restrict__ y, int rows,
                                                  double* __restrict__ rowpart,
                                                  double* __restrict__ colpart)
{
    __shared__ double red[4][256];
    // the wave index as a wave-uniform value: the mask words of a row are
    // then one uniform address (scalar loads instead of 64-lane broadcasts)
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int strip = blockIdx.x, blk = blockIdx.y, nstrip = gridDim.x;
    const int c0 = strip * 256 + lane * 4;
    const int r0 = blk * rows;
    const int r1 = min(n, r0 + rows);
    double yq[4], cs[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 4; q++) yq[q] = (ROWS && c0 + q < m) ? (y ? y[c0 + q] : 1.0) : 0.0;
    // four rows per step (rows r, r + 4, r + 8, r + 12 of this wave): their
    // words (two 16-byte loads each) and weights are loaded before any is used
    constexpr int U = 4;
    for (int rb = r0 + wave; rb < r1; rb += 4 * U) {
        ulonglong2 w[U][2];
        double xr[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int r = rb + 4 * u;
            if (r < r1) {
                const ulonglong2* wr = reinterpret_cast<const ulonglong2*>(
                    mask + (static_cast<long>(r) * nstrip + strip) * 4);
                w[u][0] = wr[0];
                w[u][1] = wr[1];
                xr[u] = (COLS && x) ? x[r] : 1.0;
            } else {
                w[u][0] = make_ulonglong2(0ull, 0ull);
                w[u][1] = make_ulonglong2(0ull, 0ull);
                xr[u] = 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int r = rb + 4 * u;
            if (r >= r1) break;
            const bool on[4] = {((w[u][0].x >> lane) & 1ull) != 0, ((w[u][0].y >> lane) & 1ull) != 0,
                                ((w[u][1].x >> lane) & 1ull) != 0, ((w[u][1].y >> lane) & 1ull) != 0};
            if (ROWS) {
                double rs = 0.0;
#pragma unroll
                for (int q = 0; q < 4; q++) rs += on[q] ? yq[q] : 0.0;
                rs = rsum16(rs);
                if ((lane & 15) == 0)
                    rowpart[(static_cast<long>(strip) * 4 + (lane >> 4)) * n + r] = rs;
            }
            if (COLS) {
#pragma unroll
                for (int q = 0; q < 4; q++) cs[q] += on[q] ? xr[u] : 0.0;
            }
        }
    }
    if (COLS) {
#pragma unroll
        for (int q = 0; q < 4; q++) red[wave][lane * 4 + q] = cs[q];
        __syncthreads();
        if (wave == 0) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int j = c0 + q;
                const double v = ((red[0][lane * 4 + q] + red[1][lane * 4 + q]) +
                                  red[2][lane * 4 + q]) + red[3][lane * 4 + q];
                if (j < m) colpart[static_cast<long>(blk) * m + j] = v;
            }
        }
    }
}

// out_i = x_i / (h_i + shift) - c / (h_i + shift)
__global__ __launch_bounds__(256) void k_qrot_ainv(const double* __restrict__ x,
                                                   const double* __restrict__ h, double shift,
                                                   double c, double* __restrict__ out, int count)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) {
        const double hs = h[i] + shift;
        out[i] = c != 0.0 ? x[i] / hs - c / hs : x[i] / hs;
    }
}

// out_i = h_i x_i + shift x_i - t_i + c
__global__ __launch_bounds__(256) void k_qrot_delta_fin(const double* __restrict__ x,
                                                        const double* __restrict__ h,
                                                        double shift,
                                                        const double* __restrict__ t, double c,
                                                        double* __restrict__ out, int count)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) out[i] = ((h[i] * x[i] + shift * x[i]) - t[i]) + c;
}

// out_i = (t_i + h_i x_i) + c
__global__ __launch_bounds__(256) void k_qrot_hfin(const double* __restrict__ x,
                                                   const double* __restrict__ h,
                                                   const double* __restrict__ t, double c,
                                                   double* __restrict__ out, int count)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) out[i] = (t[i] + h[i] * x[i]) + c;
}

// out_i = 1 / (h_i + shift)
__global__ __launch_bounds__(256) void k_qrot_recip(const double* __restrict__ h, double shift,
                                                    double* __restrict__ out, int count)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) out[i] = 1.0 / (h[i] + shift);
}

}  // namespace

int qrot_rows(int n, int m)
{
    // >= 1024 workgroups where the matrix allows, blocks of 16..256 rows
    const int strips = ceil_div(m, 256);
    int rows = 16;
    while (rows < 256 && static_cast<long>(strips) * ceil_div(n, rows * 2) >= 1024) rows *= 2;
    return rows;
}

template <typename TM>
void qrot_tile_launch(dim3 grid, int rows, hipStream_t s, const TM* M, long ldm, int n, int m,
                      const double* alpha, const double* beta, double div, const double* x,
                      const double* y, double* rowpart, double* colpart, double* sqpart)
{
    const bool wt = x != nullptr || y != nullptr;
    if (wt)
        hipLaunchKernelGGL((k_qrot_tile<TM, true, true>), grid, dim3(256), 0, s, M, ldm, n, m,
                           alpha, beta, div, x, y, rows, rowpart, colpart, sqpart);
    else if (div != 1.0)
        hipLaunchKernelGGL((k_qrot_tile<TM, false, true>), grid, dim3(256), 0, s, M, ldm, n, m,
                           alpha, beta, div, x, y, rows, rowpart, colpart, sqpart);
    else
        hipLaunchKernelGGL((k_qrot_tile<TM, false, false>), grid, dim3(256), 0, s, M, ldm, n, m,
                           alpha, beta, div, x, y, rows, rowpart, colpart, sqpart);
    KERNEL_CHECK();
}

void launch_qrot_tile(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                      const double* beta, double div, const double* x, const double* y,
                      double* rowpart, double* colpart, double* sqpart, hipStream_t s)
{
    const int rows = qrot_rows(n, m);
    const dim3 grid(ceil_div(m, 256), ceil_div(n, rows));
    if (bulk == 0)
        qrot_tile_launch(grid, rows, s, static_cast<const float*>(M), ldm, n, m, alpha, beta, div,
                         x, y, rowpart, colpart, sqpart);
    else
        qrot_tile_launch(grid, rows, s, static_cast<const double*>(M), ldm, n, m, alpha, beta,
                         div, x, y, rowpart, colpart, sqpart);
}

void launch_qrot_round(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                       const double* beta, double reg, const double* x, const double* y,
                       const double* ea, const double* eb, const double* l1, double* P,
                       long ldp, double* scal, hipStream_t s)
{
    const int rows = qrot_rows(n, m);
    const dim3 grid(ceil_div(m, 256), ceil_div(n, rows));
    if (bulk == 0)
        hipLaunchKernelGGL(k_qrot_round_tile<float>, grid, dim3(256), 0, s,
                           static_cast<const float*>(M), ldm, n, m, alpha, beta, reg, x, y, ea,
                           eb, l1, rows, P, ldp, scal);
    else
        hipLaunchKernelGGL(k_qrot_round_tile<double>, grid, dim3(256), 0, s,
                           static_cast<const double*>(M), ldm, n, m, alpha, beta, reg, x, y, ea,
                           eb, l1, rows, P, ldp, scal);
    KERNEL_CHECK();
}

void launch_qrot_pl_init(double* xc, int* kprev, int* done, int count, int* open, hipStream_t s)
{
    hipLaunchKernelGGL(k_qrot_pl_init, dim3(ceil_div(count, 256)), dim3(256), 0, s, xc, kprev,
                       done, count, open);
    KERNEL_CHECK();
}

void launch_qrot_pl_step(int bulk, const void* M, long ldm, int n, int m, bool cols,
                         const double* pot, const double* w, double reg, double* xc,
                         int* kprev, int* done, double* spart, int* kpart, int* open,
                         hipStream_t s)
{
    const int rows = qrot_rows(n, m);
    const dim3 grid(ceil_div(m, 256), ceil_div(n, rows));
    hipLaunchKernelGGL(k_qrot_zero_int, dim3(1), dim3(1), 0, s, open);
    if (bulk == 0) {
        const float* Mf = static_cast<const float*>(M);
        if (cols)
            hipLaunchKernelGGL((k_qrot_pl_tile<float, true>), grid, dim3(256), 0, s, Mf, ldm, n,
                               m, pot, xc, done, rows, spart, kpart);
        else
            hipLaunchKernelGGL((k_qrot_pl_tile<float, false>), grid, dim3(256), 0, s, Mf, ldm,
                               n, m, pot, xc, done, rows, spart, kpart);
    } else {
        const double* Md = static_cast<const double*>(M);
        if (cols)
            hipLaunchKernelGGL((k_qrot_pl_tile<double, true>), grid, dim3(256), 0, s, Md, ldm,
                               n, m, pot, xc, done, rows, spart, kpart);
        else
            hipLaunchKernelGGL((k_qrot_pl_tile<double, false>), grid, dim3(256), 0, s, Md, ldm,
                               n, m, pot, xc, done, rows, spart, kpart);
    }
    KERNEL_CHECK();
    const int count = cols ? m : n;
    const int nparts = cols ? ceil_div(n, rows) : ceil_div(m, 256);
    hipLaunchKernelGGL(k_qrot_pl_update, dim3(ceil_div(count, 256)), dim3(256), 0, s, spart,
                       kpart, nparts, static_cast<long>(count), count, w, reg, xc, kprev, done,
                       open);
    KERNEL_CHECK();
}

void launch_qrot_plan(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                      const double* beta, double reg, double* plan, long r0, int nr,
                      hipStream_t s)
{
    (void)n;
    if (nr <= 0 || m <= 0) return;
    const int gx = ceil_div(m, 256) < 8 ? ceil_div(m, 256) : 8;
    if (bulk == 0)
        hipLaunchKernelGGL(k_qrot_plan<float>, dim3(gx, nr), dim3(256), 0, s,
                           static_cast<const float*>(M), ldm, m, alpha, beta, reg, plan, r0);
    else
        hipLaunchKernelGGL(k_qrot_plan<double>, dim3(gx, nr), dim3(256), 0, s,
                           static_cast<const double*>(M), ldm, m, alpha, beta, reg, plan, r0);
    KERNEL_CHECK();
}

// APDAGD plan pass and the marginal residual: out[0] = <M, P>, out[1] =
// ||P||^2, out[2] = ||P1 - a||^2 + ||P'1 - b||^2.  Scratch: rowpart (strips
// x n), colpart (blocks x m), rsum (n), csum (m), wpart (2 strips blocks +
// ceil(max(n, m) / 256)).
void launch_qrot_apd_plan(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                          const double* beta, double reg, double cold, double tau, bool init,
                          double* P,
                          long ldp, const double* a, const double* b, double* rowpart,
                          double* colpart, double* rsum, double* csum, double* wpart,
                          double* out, hipStream_t s)
{
    const int rows = qrot_rows(n, m);
    const int nstrip = ceil_div(m, 256), nblk = ceil_div(n, rows);
    const dim3 grid(nstrip, nblk);
    double* mpart = wpart;
    double* spart = wpart + static_cast<long>(nstrip) * nblk;
    double* rpart = spart + static_cast<long>(nstrip) * nblk;
    if (bulk == 0) {
        const float* Mf = static_cast<const float*>(M);
        if (init)
            hipLaunchKernelGGL((k_qrot_apd_plan<float, true>), grid, dim3(256), 0, s, Mf, ldm, n,
                               m, alpha, beta, reg, cold, tau, P, ldp, rows, rowpart, colpart, mpart,
                               spart);
        else
            hipLaunchKernelGGL((k_qrot_apd_plan<float, false>), grid, dim3(256), 0, s, Mf, ldm,
                               n, m, alpha, beta, reg, cold, tau, P, ldp, rows, rowpart, colpart, mpart,
                               spart);
    } else {
        const double* Md = static_cast<const double*>(M);
        if (init)
            hipLaunchKernelGGL((k_qrot_apd_plan<double, true>), grid, dim3(256), 0, s, Md, ldm,
                               n, m, alpha, beta, reg, cold, tau, P, ldp, rows, rowpart, colpart, mpart,
                               spart);
        else
            hipLaunchKernelGGL((k_qrot_apd_plan<double, false>), grid, dim3(256), 0, s, Md, ldm,
                               n, m, alpha, beta, reg, cold, tau, P, ldp, rows, rowpart, colpart,
                               mpart, spart);
    }
    KERNEL_CHECK();
    launch_qrot_fold(rowpart, nstrip, n, n, rsum, s);
    launch_qrot_fold(colpart, nblk, m, m, csum, s);
    const int nw = ceil_div(n > m ? n : m, 256);
    hipLaunchKernelGGL(k_qrot_resid2, dim3(nw), dim3(256), 0, s, rsum, a, n, csum, b, m, rpart);
    KERNEL_CHECK();
    launch_qrot_sum1(mpart, nstrip * nblk, out, s);
    launch_qrot_sum1(spart, nstrip * nblk, out + 1, s);
    launch_qrot_sum1(rpart, nw, out + 2, s);
}

void launch_qrot_fold(const double* part, int nparts, long stride, int count, double* out,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_qrot_fold, dim3(ceil_div(count, 256)), dim3(256), 0, s, part, nparts,
                       stride, count, out);
    KERNEL_CHECK();
}

void launch_qrot_sum1(const double* part, int count, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_qrot_sum1, dim3(1), dim3(256), 0, s, part, count, out);
    KERNEL_CHECK();
}

void launch_qrot_finish(const double* alpha, const double* beta, const double* a,
                        const double* b, int n, int m, double reg, double tau, const double* sq,
                        const double* rsum, const double* csum, double* grad, double* part,
                        double* gpart, double* out, hipStream_t s)
{
    const int nw = ceil_div(n > m ? n : m, 256);
    hipLaunchKernelGGL(k_qrot_fin_part, dim3(nw), dim3(256), 0, s, alpha, beta, a, b, n, m, part);
    KERNEL_CHECK();
    if (grad) {
        hipLaunchKernelGGL(k_qrot_fin_grad, dim3(nw), dim3(256), 0, s, part, nw, a, b, n, m, reg,
                           tau, rsum, csum, grad, gpart);
        KERNEL_CHECK();
    }
    hipLaunchKernelGGL(k_qrot_fin_total, dim3(1), dim3(64), 0, s, part, nw, gpart,
                       grad ? nw : 0, reg, tau, sq, out);
    KERNEL_CHECK();
}

}  // namespace regot

namespace regot {

void launch_qrot_mask(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                      const double* beta, unsigned long long* mask, hipStream_t s)
{
    const int rows = qrot_rows(n, m);
    const dim3 grid(ceil_div(m, 256), ceil_div(n, rows));
    if (bulk == 0)
        hipLaunchKernelGGL(k_qrot_mask<float>, grid, dim3(256), 0, s,
                           static_cast<const float*>(M), ldm, n, m, alpha, beta, rows, mask);
    else
        hipLaunchKernelGGL(k_qrot_mask<double>, grid, dim3(256), 0, s,
                           static_cast<const double*>(M), ldm, n, m, alpha, beta, rows, mask);
    KERNEL_CHECK();
}

void launch_qrot_sig(const unsigned long long* mask, int n, int m, bool rows_out, bool cols_out,
                     const double* x, const double* y, double* rowpart, double* colpart,
                     hipStream_t s)
{
    const int rows = qrot_rows(n, m);
    const dim3 grid(ceil_div(m, 256), ceil_div(n, rows));
    if (rows_out && cols_out)
        hipLaunchKernelGGL((k_qrot_sig<true, true>), grid, dim3(256), 0, s, mask, n, m, x, y,
                           rows, rowpart, colpart);
    else if (rows_out)
        hipLaunchKernelGGL((k_qrot_sig<true, false>), grid, dim3(256), 0, s, mask, n, m, x, y,
                           rows, rowpart, colpart);
    else
        hipLaunchKernelGGL((k_qrot_sig<false, true>), grid, dim3(256), 0, s, mask, n, m, x, y,
                           rows, rowpart, colpart);
    KERNEL_CHECK();
}

void launch_qrot_ainv(const double* x, const double* h, double shift, double c, double* out,
                      int count, hipStream_t s)
{
    hipLaunchKernelGGL(k_qrot_ainv, dim3(ceil_div(count, 256)), dim3(256), 0, s, x, h, shift, c,
                       out, count);
    KERNEL_CHECK();
}

void launch_qrot_delta_fin(const double* x, const double* h, double shift, const double* t,
                           double c, double* out, int count, hipStream_t s)
{
    hipLaunchKernelGGL(k_qrot_delta_fin, dim3(ceil_div(count, 256)), dim3(256), 0, s, x, h,
                       shift, t, c, out, count);
    KERNEL_CHECK();
}

void launch_qrot_recip(const double* h, double shift, double* out, int count, hipStream_t s)
{
    hipLaunchKernelGGL(k_qrot_recip, dim3(ceil_div(count, 256)), dim3(256), 0, s, h, shift, out,
                       count);
    KERNEL_CHECK();
}

void launch_qrot_hfin(const double* x, const double* h, const double* t, double c, double* out,
                      int count, hipStream_t s)
{
    hipLaunchKernelGGL(k_qrot_hfin, dim3(ceil_div(count, 256)), dim3(256), 0, s, x, h, t, c, out,
                       count);
    KERNEL_CHECK();
}

}  // namespace regot
