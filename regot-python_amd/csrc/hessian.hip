// hessian.hip -- building the sparsified Hessian from the current dual point,
// its products, and the block-elimination CG solve, plus their C-ABI.
#include <cmath>
#include <cstring>
#include <deque>
#include <vector>

#include "dense_device.hpp"
#include "hess_mf.hpp"
#include "hessian.hpp"
#include "regot_hip.h"

namespace regot {
namespace {

// h1 = (g_head + a) / reg, h2 = (g_tail + b[0:m-1]) / reg.  As in
// Problem::dual_sparsified_hess(_with_density), src/sinkhorn_problem_obj.cpp:432-457,
// the diagonals come from the gradient, i.e. the exact row and column sums.
__global__ void k_hess_diag(const double* grad, const double* a, const double* b, int n, int m,
                            double reg, double* h1, double* h2)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) h1[t] = (grad[t] + a[t]) / reg;
    if (t < m - 1) h2[t] = (grad[n + t] + b[t]) / reg;
}

__global__ void k_diag_axpy(double* out, const double* d, const double* x, double shift, int count,
                            int div)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= count) return;
    out[t] = div ? x[t] / (d[t] + shift) : out[t] + d[t] * x[t];
}

// v = (w - y) / (h1 + shift)
__global__ void k_sub_div(const double* w, const double* y, const double* h1, double shift,
                          int n, double* out)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t < n) out[t] = (w[t] - y[t]) / (h1[t] + shift);
}

// ---- CG on Delta, device-controlled --------------------------------------
// Preconditioned conjugate gradients in the Chronopoulos-Gear form (one
// reduction phase per iteration; the iterates of the textbook recurrence in
// exact arithmetic), Jacobi preconditioner u = D^-1 r (dinv = 1 / diag(Delta);
// null: plain CG, u = r).  Iteration i is three kernels:
//
//   update  alpha_i, beta_i from the partials of gamma_i = r.u, delta_i = w.u
//           (every workgroup reduces them in the same order); p = u + beta p,
//           s = w + beta s, x += alpha p, r -= alpha s, u = D^-1 r; partials
//           of r.r;
//   rows    y = (sigma u) / (h1 + shift);
//   cols    stop if r.r < threshold (every workgroup decides alike), else
//           w = (h2 + shift) u - sigma' y and the partials of r.u, w.u,
//
// with beta_i = gamma_i / gamma_{i-1}, alpha_i = gamma_i / (delta_i - beta_i
// gamma_i / alpha_{i-1}); alpha_0, gamma_0 come from the set-up and beta_0 = 0.
// The iteration index is a launch argument (the host enqueues at most
// max_iter iterations); alpha and gamma are kept in two slots by parity so
// no kernel reads a slot that it writes.
// ctl: [0] done, [1] iterations when done.  scal: [0..1] alpha by parity,
// [2..3] gamma by parity, [4] r.r of the last update, [5] stop threshold,
// [6] rhs norm^2.
struct CgArgs {
    const int* rowptr;
    const int* colind;
    const double* rval;
    const int* colptr;
    const int* rowind;
    const double* cval;
    const double* h1;
    const double* h2;
    double shift;
    int n, m1;
    double* x;
    double* r;
    double* p;
    double* s;
    double* w;
    double* u;
    double* y;      // n
    double* part;   // [2 nbc] r.u, w.u per column workgroup; [nbu] r.r per update workgroup
    double* scal;
    int* ctl;
    int nbc;        // column workgroups
    int nbu;        // update workgroups (256 entries each)
    const double* dinv;   // Jacobi preconditioner, or null
};

// dinv_j = 1 / (h2_j + shift - sum_i sigma_ij^2 / (h1_i + shift)), the diagonal
// of Delta (>= shift since sigma_ij <= h1_i and sum_i sigma_ij <= h2_j); one
// wave per column.
__global__ __launch_bounds__(256) void k_delta_diag(const int* __restrict__ colptr,
                                                    const int* __restrict__ rowind,
                                                    const double* __restrict__ cval,
                                                    const double* __restrict__ h1,
                                                    const double* __restrict__ h2, double shift,
                                                    int m1, double* __restrict__ dinv)
{
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= m1) return;
    double s = 0.0;
    for (int k = colptr[j] + lane; k < colptr[j + 1]; k += 64) {
        const double v = cval[k];
        s += v * v / (h1[rowind[k]] + shift);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) {
        double d = h2[j] + shift - s;
        const double floor = 1e-12 * (h2[j] + shift);
        if (!(d > floor)) d = (h2[j] + shift > 0.0) ? h2[j] + shift : 1.0;
        dinv[j] = 1.0 / d;
    }
}

// u = dinv .* r (or r), partials of r.u and r.r per 1024 entries.
__global__ __launch_bounds__(256) void k_pcg_start(const double* __restrict__ r,
                                                   const double* __restrict__ dinv, int m1,
                                                   double* __restrict__ u,
                                                   double* __restrict__ part)
{
    __shared__ double sm[2][256];
    double rz = 0.0, rr = 0.0;
    for (int j = blockIdx.x * 1024 + threadIdx.x; j < (blockIdx.x + 1) * 1024 && j < m1;
         j += 256) {
        const double rj = r[j];
        const double zj = dinv ? rj * dinv[j] : rj;
        u[j] = zj;
        rz += rj * zj;
        rr += rj * rj;
    }
    sm[0][threadIdx.x] = rz;
    sm[1][threadIdx.x] = rr;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + o];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sm[0][0];
        part[2 * blockIdx.x + 1] = sm[1][0];
    }
}

// Fixed-order sums of NV values over the 256 threads: each wave's by DPP and
// shuffle steps, then the four wave totals in order from LDS (one barrier);
// the same bits on every thread.  sm must not be reused before a barrier.
template <int NV>
__device__ __forceinline__ void wg_sums(double (&v)[NV], double (*sm)[4])
{
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = dev::group_sum<64>(v[q]);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) sm[q][wave] = v[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = ((sm[q][0] + sm[q][1]) + sm[q][2]) + sm[q][3];
}

// The CG kernels serve one or two right-hand sides: grid row blockIdx.y
// takes a0 or a1 (two independent CGs on the same Delta in the same
// launches; every right-hand side runs exactly the single-CG arithmetic).
__global__ __launch_bounds__(256) void k_cg_update(CgArgs a0, CgArgs a1, int it)
{
    __shared__ double sm1[2][4], sm2[1][4];
    const CgArgs a = blockIdx.y == 0 ? a0 : a1;
    // this thread's entries are loaded first, so that they arrive during the
    // stop-flag read and the partial sums
    const int j = blockIdx.x * 256 + threadIdx.x;
    double uj = 0.0, pj0 = 0.0, wj = 0.0, sj0 = 0.0, rj0 = 0.0, xj = 0.0, dj = 0.0;
    if (j < a.m1) {
        uj = a.u[j];
        pj0 = a.p[j];
        wj = a.w[j];
        sj0 = a.s[j];
        rj0 = a.r[j];
        xj = a.x[j];
        if (a.dinv) dj = a.dinv[j];
    }
    if (a.ctl[0]) return;
    double al, be = 0.0;
    const int cur = it & 1, prev = cur ^ 1;
    if (it == 0) {
        al = a.scal[0];
    } else {
        // the column workgroups' partials, b = t, t + 256, ... in order, eight
        // loads in flight
        double v[2] = {0.0, 0.0};
        int b = threadIdx.x;
        for (; b + 7 * 256 < a.nbc; b += 8 * 256) {
            double x0[8], x1[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                x0[u] = a.part[2 * (b + 256 * u)];
                x1[u] = a.part[2 * (b + 256 * u) + 1];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                v[0] += x0[u];
                v[1] += x1[u];
            }
        }
        for (; b < a.nbc; b += 256) {
            v[0] += a.part[2 * b];
            v[1] += a.part[2 * b + 1];
        }
        wg_sums<2>(v, sm1);
        const double g = v[0], d = v[1];
        const double al_prev = a.scal[prev], g_prev = a.scal[2 + prev];
        be = g / g_prev;
        al = g / (d - be * g / al_prev);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.scal[cur] = al;
            a.scal[2 + cur] = g;
        }
    }
    double rr = 0.0;
    if (j < a.m1) {
        const double pj = uj + be * pj0;
        const double sj = wj + be * sj0;
        const double rj = rj0 - al * sj;
        a.p[j] = pj;
        a.s[j] = sj;
        a.x[j] = xj + al * pj;
        a.r[j] = rj;
        a.u[j] = a.dinv ? dj * rj : rj;
        rr = rj * rj;
    }
    double v1[1] = {rr};
    wg_sums<1>(v1, sm2);
    if (threadIdx.x == 0) a.part[2 * a.nbc + blockIdx.x] = v1[0];
}

// y = (sigma u) / (h1 + shift); LPR lanes per row, 64 / LPR rows per wave.
template <int LPR>
__global__ __launch_bounds__(256) void k_cg_rows(CgArgs a0, CgArgs a1)
{
    const CgArgs a = blockIdx.y == 0 ? a0 : a1;
    const int lane = threadIdx.x & 63;
    const int sl = lane & (LPR - 1);
    const int i = (blockIdx.x * 4 + (threadIdx.x >> 6)) * (64 / LPR) + lane / LPR;
    // the row's extent and diagonal are loaded with the stop flag
    int k0 = 0, e = 0;
    double h1 = 0.0;
    if (i < a.n) {
        k0 = a.rowptr[i];
        e = a.rowptr[i + 1];
        h1 = a.h1[i];
    }
    if (a.ctl[0]) return;
    double acc = 0.0;
#pragma unroll 4
    for (int k = k0 + sl; k < e; k += LPR) acc += a.rval[k] * a.u[a.colind[k]];
    acc = dev::group_sum<LPR>(acc);
    if (sl == 0 && i < a.n) a.y[i] = acc / (h1 + a.shift);
}

// Stop test on the r.r partials of the update kernel (same order in every
// workgroup), then w = (h2 + shift) u - sigma' y with LPR lanes per column and
// the partials of r.u and w.u per workgroup (fixed order).
template <int LPR>
__global__ __launch_bounds__(256) void k_cg_cols(CgArgs a0, CgArgs a1, int it)
{
    constexpr int CPW = 64 / LPR;   // columns per wave
    __shared__ double sm[2][4 * CPW];
    const CgArgs a = blockIdx.y == 0 ? a0 : a1;
    const int lane = threadIdx.x & 63;
    const int sl = lane & (LPR - 1);
    const int slot = (threadIdx.x >> 6) * CPW + lane / LPR;
    const int j = blockIdx.x * (4 * CPW) + slot;
    // the column's extent and entries are loaded with the stop flag and the
    // partials
    int k0 = 0, e = 0;
    double uj = 0.0, h2j = 0.0, rj = 0.0;
    if (j < a.m1) {
        k0 = a.colptr[j];
        e = a.colptr[j + 1];
        uj = a.u[j];
        h2j = a.h2[j];
        rj = a.r[j];
    }
    // every wave sums the update's r.r partials itself, in the same order
    double v = 0.0;
    for (int b = lane; b < a.nbu; b += 64) v += a.part[2 * a.nbc + b];
    if (a.ctl[0]) return;
    const double rr = dev::group_sum<64>(v);
    if (rr < a.scal[5]) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.scal[4] = rr;
            a.ctl[1] = it;   // the converged update is not counted (Eigen's convention)
            a.ctl[0] = 1;
        }
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) a.scal[4] = rr;
    double t = 0.0;
#pragma unroll 4
    for (int k = k0 + sl; k < e; k += LPR) t += a.cval[k] * a.y[a.rowind[k]];
    t = dev::group_sum<LPR>(t);
    if (sl == 0) {
        double g = 0.0, d = 0.0;
        if (j < a.m1) {
            const double wj = h2j * uj + a.shift * uj - t;
            a.w[j] = wj;
            g = rj * uj;
            d = wj * uj;
        }
        sm[0][slot] = g;
        sm[1][slot] = d;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double g = 0.0, d = 0.0;
        for (int q = 0; q < 4 * CPW; q++) {
            g += sm[0][q];
            d += sm[1][q];
        }
        a.part[2 * blockIdx.x] = g;
        a.part[2 * blockIdx.x + 1] = d;
    }
}

// Matrix-free form of k_cg_cols: sigma' y arrives as nwg FP64 partials per
// column (mpart, nwg x m1) from the pass over M, which also made the stop
// test; w = (h2 + shift) u - sigma' y and the partials of r.u, w.u per
// workgroup of 256 columns.
__global__ __launch_bounds__(256) void k_cg_cols_mf(CgArgs a, const double* __restrict__ mpart,
                                                    int nwg)
{
    __shared__ double sm[2][4];
    if (a.ctl[0]) return;
    const int m1 = a.m1;
    const int j = blockIdx.x * 256 + threadIdx.x;
    double g = 0.0, d = 0.0;
    if (j < m1) {
        double t = 0.0;
        int w = 0;
        for (; w + 8 <= nwg; w += 8) {
            double x[8];
#pragma unroll
            for (int q = 0; q < 8; q++) x[q] = mpart[static_cast<long>(w + q) * m1 + j];
#pragma unroll
            for (int q = 0; q < 8; q++) t += x[q];
        }
        for (; w < nwg; w++) t += mpart[static_cast<long>(w) * m1 + j];
        const double uj = a.u[j];
        const double wj = a.h2[j] * uj + a.shift * uj - t;
        a.w[j] = wj;
        g = a.r[j] * uj;
        d = wj * uj;
    }
    double v[2] = {g, d};
    wg_sums<2>(v, sm);
    if (threadIdx.x == 0) {
        a.part[2 * blockIdx.x] = v[0];
        a.part[2 * blockIdx.x + 1] = v[1];
    }
}

}  // namespace

void hess_ensure_transpose(regot_problem* p)
{
    if (p->dMT) return;
    const size_t esz = p->bulk == 0 ? 4 : 8;
    const long align_el = 256 / static_cast<long>(esz);
    p->ldt = (p->n + align_el - 1) / align_el * align_el;
    HIP_CHECK(hipMalloc(&p->dMT, static_cast<size_t>(p->m) * p->ldt * esz));
    launch_transpose(p->bulk, p->dM, p->ldm, p->n, p->m, p->dMT, p->ldt, p->stream);
}

SparseHessian& hess_of(regot_problem* p)
{
    if (!p->hess) {
        p->hess = std::make_unique<SparseHessian>();
        p->hess->n = p->n;
        p->hess->m = p->m;
    }
    return *p->hess;
}

static ElemCtx make_ctx(regot_problem* p, SparseHessian& H)
{
    ElemCtx c;
    c.alpha = H.vn;
    c.beta = H.vm;
    c.k = p->k;
    c.reg = p->reg;
    c.n = p->n;
    c.m = p->m;
    return c;
}

// Stored entries from the row counts in H.rcnt: scan into H.rowptr, total
// into H.nnz (a host synchronisation).
static long long count_stored(regot_problem* p, SparseHessian& H)
{
    const int n = p->n;
    hipStream_t s = p->stream;
    H.rowptr.ensure(n + 1);
    H.ltmp.ensure(4);
    launch_scan(H.rcnt, n, H.rowptr, H.ltmp.p, s);
    H.hll.ensure(4);
    HIP_CHECK(hipMemcpyAsync(H.hll.p, H.ltmp.p, sizeof(long long), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    H.nnz = H.hll.p[0];
    return H.nnz;
}

static MfArgs mf_args(regot_problem* p, const SparseHessian& H)
{
    MfArgs a;
    a.M = static_cast<const float*>(p->dM);
    a.ldm = p->ldm;
    a.n = p->n;
    a.m1 = p->m - 1;
    a.alpha = H.vn;
    a.beta = H.vm;
    a.k = p->k;
    a.inv_reg = 1.0 / p->reg;
    a.T = H.mf_t;
    a.ldt = H.mf_ldt;
    a.h1 = H.h1;
    return a;
}

// Whether a build may go matrix-free: FP32 bulk, m - 1 <= 16384, and the
// handle's storage mode (driver: the caller is a solver driver, which needs
// only products).
static bool mf_allowed(const regot_problem* p, bool driver)
{
    if (p->bulk != 0 || !mf_shape_ok(p->n, p->m) || p->hess_storage == 0) return false;
    return p->hess_storage == 1 || driver;
}

// Buffers of a matrix-free build: the kept matrix (n x mf_pitch(m) FP32) and
// the column partials of the passes.
static void mf_buffers(regot_problem* p, SparseHessian& H)
{
    H.mf_nwg = mf_blocks(p->n, p->ncu);
    H.mf_ldt = mf_pitch(p->m);
    H.mf_t.ensure(static_cast<size_t>(p->n) * H.mf_ldt);
    H.mf_part.ensure(static_cast<size_t>(H.mf_nwg) * std::max(p->m - 1, 1));
}

// Matrix-free build with every entry kept (the full Hessian).
static void build_mf_full(regot_problem* p, SparseHessian& H)
{
    mf_buffers(p, H);
    MfArgs a = mf_args(p, H);
    a.Tout = H.mf_t;
    launch_mf_full(a, H.mf_nwg, p->stream);
    H.mf = true;
}

// Shared tail of both rules: row counts (in H.rcnt) -> scan -> CSR write ->
// CSC by a counting sort of the CSR.
static void build_storage(regot_problem* p, SparseHessian& H, const ElemCtx& c,
                          const KeepRule& rule, bool counted = false)
{
    const int n = p->n, m = p->m;
    hipStream_t s = p->stream;
    const long long nnz = counted ? H.nnz : count_stored(p, H);
    if (nnz > 0x7fffffffLL) fail(E_INVALID, "sparsified Hessian has %lld entries (> 2^31)", nnz);
    H.colind.ensure(nnz);
    H.rval.ensure(nnz);
    H.rowind.ensure(nnz);
    H.cval.ensure(nnz);
    launch_row_write(p->bulk, p->dM, p->ldm, c, rule, H.rowptr, H.colind, H.rval, s);
    // CSC: the transpose of the CSR just written.  Sparse sets by a counting
    // sort of the nnz entries; dense ones by a pass over M' with the ordered
    // row writer when M' is there (threshold rule, FP64 bulk), else over M by
    // column strips.  At most 512 row blocks: the offset table stays <= 2 KiB
    // per column.
    const int rows_per_block = std::max(32, ceil_div(n, 512));
    H.ccnt.ensure(m);
    H.colptr.ensure(m);
    H.csc_work.ensure(static_cast<size_t>(ceil_div(n, rows_per_block)) * std::max(m - 1, 1));
    // Without M' (FP32 bulk) the strip pass recomputes T column-wise and
    // costs more than the sort up to the matrix-free densities (1/12).
    const double sort_max = p->dMT ? 0.02 : 0.09;
    const bool dense = static_cast<double>(nnz) > sort_max * n * static_cast<double>(m - 1);
    const int how = !dense ? 0 : (rule.mode == 0 && p->dMT) ? 2 : 1;
    launch_build_csc(p->bulk, p->dM, p->ldm, p->dMT, p->ldt, c, rule, H, how, H.csc_work,
                     rows_per_block, H.ccnt, H.colptr, H.ltmp.p + 1, H.rowind, H.cval, s);
}

static void prepare(regot_problem* p, SparseHessian& H, const double* d_gamma, const double* d_grad)
{
    const int n = p->n, m = p->m;
    hipStream_t s = p->stream;
    H.vn.ensure(n);
    H.vm.ensure(m);
    H.h1.ensure(n);
    H.h2.ensure(m);
    H.rcnt.ensure(n + 1);   // also receives an (n+1)-entry scan
    H.zcnt.ensure(n);
    H.mf = false;
    launch_split_gamma(d_gamma, n, m, H.vn, H.vm, s);
    hipLaunchKernelGGL(k_hess_diag, dim3(ceil_div(std::max(n, m), 256)), dim3(256), 0, s, d_grad,
                       p->a.p, p->b.p, n, m, p->reg, H.h1.p, H.h2.p);
    KERNEL_CHECK();
}

void hess_threshold(regot_problem* p, const double* d_gamma, const double* d_grad, double delta,
                    double density_hint, bool driver)
{
    SparseHessian& H = hess_of(p);
    prepare(p, H, d_gamma, d_grad);
    const ElemCtx c = make_ctx(p, H);
    hipStream_t s = p->stream;
    H.tau.ensure(p->m);
    H.rho.ensure(p->n);
    const int find_small = density_hint > 0.2 ? 1 : 0;
    const bool may_mf = mf_allowed(p, driver);
    if (p->bulk == 0) {
        // FP32 bulk: tau from row-major M (no M'), then the row side in one
        // read of M, with the kept-set bitmask when the build may go
        // matrix-free
        threshold_tau_rows(p->dM, p->ldm, c, delta, find_small, H, H.tau, s);
        if (mf_shape_ok(p->n, p->m)) {
            if (may_mf) mf_buffers(p, H);
            MfArgs a = mf_args(p, H);
            a.tau = H.tau;
            a.Tout = may_mf ? H.mf_t.p : nullptr;
            launch_mf_rule(a, mf_blocks(p->n, p->ncu), 2.0 * delta, H.rho, H.rcnt, H.zcnt, s);
        } else {
            launch_row_threshold(p->bulk, p->dM, p->ldm, c, H.tau, 2.0 * delta, H.rho, H.rcnt,
                                 H.zcnt, s);
        }
    } else {
        hess_ensure_transpose(p);
        launch_col_threshold(p->bulk, p->dMT, p->ldt, c, delta, find_small, H.tau, s);
        launch_row_threshold(p->bulk, p->dM, p->ldm, c, H.tau, 2.0 * delta, H.rho, H.rcnt,
                             H.zcnt, s);
    }
    KeepRule rule;
    rule.mode = 0;
    rule.tau = H.tau;
    rule.rho = H.rho;
    bool mf = false;
    if (may_mf) {
        // matrix-free once one product over the stored copies (2 x 12 bytes
        // per entry) would move more than half a read of M (4 bytes per
        // entry): the stored builds' CSR + CSC writes and the CG products'
        // gathers make them dearer per byte than the pass over M
        const long long nnz = count_stored(p, H);
        mf = p->hess_storage == 1 || 24.0 * static_cast<double>(nnz) >
                                         2.0 * static_cast<double>(p->n) * (p->m - 1);
        if (mf)
            H.mf = true;
        else
            build_storage(p, H, c, rule, true);
    } else {
        build_storage(p, H, c, rule);
    }
    // FP64-zero kept entries count towards the reported density
    launch_scan(H.zcnt, p->n, H.rcnt, H.ltmp.p + 2, s);
    HIP_CHECK(hipMemcpyAsync(H.hll.p, H.ltmp.p, 3 * sizeof(long long), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    H.nnz_zero = H.hll.p[2];
    if (!mf && H.hll.p[1] != H.nnz)
        fail(E_INTERNAL, "CSR and CSC counts differ (%lld vs %lld)", H.nnz, H.hll.p[1]);
}

void hess_topk(regot_problem* p, const double* d_gamma, const double* d_grad, double density,
               bool driver)
{
    SparseHessian& H = hess_of(p);
    prepare(p, H, d_gamma, d_grad);
    const ElemCtx c = make_ctx(p, H);
    hipStream_t s = p->stream;
    const int n = p->n, m = p->m;
    density = std::min(1.0, std::max(0.0, density));
    const long long len = static_cast<long long>(n) * (m - 1);
    const long long kk = static_cast<long long>(static_cast<double>(len) * density);
    if (kk >= len && mf_allowed(p, driver)) {
        // every nonzero entry kept: products straight from M
        build_mf_full(p, H);
        H.nnz = len;
        H.nnz_zero = 0;
        return;
    }
    KeepRule rule;
    rule.mode = 1;
    rule.preserve = 1;
    H.ltmp.ensure(std::max(4, n));
    if (kk <= 0) {
        rule.kstar = ~0ull;   // no key exceeds it; no tie can match a finite value's key
        rule.tie_take = 0;
    } else if (kk >= len) {
        rule.kstar = 0ull;    // every key > 0 kept; ties at key 0 (exact zeros) all taken
        rule.tie_take = len;
    } else {
        // the k-th key and, in the same reads of M, the per-row tallies
        unsigned long long kstar = 0;
        long long above = 0;
        topk_select_rows(p->bulk, p->dM, p->ldm, c, kk, rule.preserve, H, &kstar, &above,
                         H.rcnt, H.zcnt, s);
        rule.kstar = kstar;
        rule.tie_take = kk - above;
    }
    // per row: ties and the stored entries kept without a tie slot; the tie
    // offsets (row-major order) then give the stored count of every row
    if (kk <= 0 || kk >= len)
        launch_row_tally(p->bulk, p->dM, p->ldm, c, rule.kstar, rule.preserve, H.rcnt, H.zcnt,
                         s);
    H.ltmp.ensure(n + 4);
    launch_scan64(H.rcnt, n, H.ltmp.p + 4, s);
    rule.tie_off = H.ltmp.p + 4;
    launch_topk_count(H.rcnt, H.zcnt, rule, n, H.rcnt, s);
    H.nnz_zero = 0;
    build_storage(p, H, c, rule);
    HIP_CHECK(hipMemcpyAsync(H.hll.p, H.ltmp.p, 2 * sizeof(long long), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (H.hll.p[1] != H.nnz)
        fail(E_INTERNAL, "CSR and CSC counts differ (%lld vs %lld)", H.nnz, H.hll.p[1]);
}

// One matrix-free pass (see hess_mf.hpp for the modes).
static void mf_pass(regot_problem* p, SparseHessian& H, int mode, const double* xc,
                    const double* xr, double shift, double* rowout)
{
    MfArgs a = mf_args(p, H);
    a.xc = xc;
    a.xr = xr;
    a.shift = shift;
    a.rowout = rowout;
    a.part = H.mf_part;
    launch_mf_pass(mode, a, H.mf_nwg, p->stream);
}

// Delta x = h2 .* x + shift x - sigma' ((sigma x) ./ (h1 + shift)); x, out: m-1.
void hess_apply_delta(regot_problem* p, const double* x, double shift, double* out)
{
    SparseHessian& H = hess_of(p);
    if (H.mf) {
        mf_pass(p, H, MF_CG, x, nullptr, shift, nullptr);
        launch_mf_fold(H.mf_part, H.mf_nwg, p->m - 1, 2, x, H.h2, shift, out, p->stream);
        return;
    }
    H.cg_y_buf.ensure(p->n);
    launch_spmv_rows(H, x, H.cg_y_buf, true, shift, p->stream);
    launch_spmv_cols(H, H.cg_y_buf, x, shift, 1, out, p->stream);
}

// Hs x = [h1 .* w + sigma z ; h2 .* z + sigma' w]; x, out: n+m-1.
void hess_apply_hs(regot_problem* p, const double* x, double* out)
{
    SparseHessian& H = hess_of(p);
    const int n = p->n, m = p->m;
    if (H.mf)
        mf_pass(p, H, MF_HS, x + n, x, 0.0, out);
    else
        launch_spmv_rows(H, x + n, out, false, 0.0, p->stream);
    hipLaunchKernelGGL(k_diag_axpy, dim3(ceil_div(n, 256)), dim3(256), 0, p->stream, out, H.h1.p,
                       x, 0.0, n, 0);
    KERNEL_CHECK();
    if (H.mf)
        launch_mf_fold(H.mf_part, H.mf_nwg, m - 1, 3, x + n, H.h2, 0.0, out + n, p->stream);
    else
        launch_spmv_cols(H, x, x + n, 0.0, 2, out + n, p->stream);
}

// CG on Delta (size m-1) from zero; textbook unpreconditioned CG with the
// stopping rule ||r||^2 < max(tol^2 ||b||^2, DBL_MIN) and the iteration cap
// counted as Eigen's ConjugateGradient counts it (src/sinkhorn_cg.cpp:29-38).
// Lanes per row (column) of the CG products: about three entries per lane,
// and at least 8192 waves over `count` rows (one resident round of the chip;
// fewer waves leave the load latency exposed).
static int lanes_for(double avg, int count)
{
    int l = 4;
    while (l < 64 && (3.0 * l < avg || static_cast<long>(count) * l < 8192L * 64)) l <<= 1;
    return l;
}

// a0 / a1: the right-hand sides (ny of them, 1 or 2) in grid rows 0 / 1.
template <int LPR>
static void launch_cg_rows(const CgArgs& a0, const CgArgs& a1, int ny, hipStream_t s)
{
    const int rows_per_wg = 4 * (64 / LPR);
    hipLaunchKernelGGL(k_cg_rows<LPR>, dim3(ceil_div(a0.n, rows_per_wg), ny), dim3(256), 0, s,
                       a0, a1);
}

template <int LPR>
static void launch_cg_cols(const CgArgs& a0, const CgArgs& a1, int ny, int it, hipStream_t s)
{
    hipLaunchKernelGGL(k_cg_cols<LPR>, dim3(a0.nbc, ny), dim3(256), 0, s, a0, a1, it);
}

static void cg_rows(int lpr, const CgArgs& a0, const CgArgs& a1, int ny, hipStream_t s)
{
    switch (lpr) {
    case 4: launch_cg_rows<4>(a0, a1, ny, s); break;
    case 8: launch_cg_rows<8>(a0, a1, ny, s); break;
    case 16: launch_cg_rows<16>(a0, a1, ny, s); break;
    case 32: launch_cg_rows<32>(a0, a1, ny, s); break;
    default: launch_cg_rows<64>(a0, a1, ny, s); break;
    }
}

static void cg_cols(int lpr, const CgArgs& a0, const CgArgs& a1, int ny, int it, hipStream_t s)
{
    switch (lpr) {
    case 4: launch_cg_cols<4>(a0, a1, ny, it, s); break;
    case 8: launch_cg_cols<8>(a0, a1, ny, it, s); break;
    case 16: launch_cg_cols<16>(a0, a1, ny, it, s); break;
    case 32: launch_cg_cols<32>(a0, a1, ny, it, s); break;
    default: launch_cg_cols<64>(a0, a1, ny, it, s); break;
    }
}

// One right-hand side of the CG on Delta: its inputs, its device state and
// its outcome.
struct CgSide {
    const double* rhs = nullptr;   // device, m-1
    double* x = nullptr;           // device, m-1 (the solution)
    const double* x0 = nullptr;    // device, m-1, or null (start from 0)
    DBuf<double>*r, *pv, *q, *sv, *u, *y, *part, *scal;
    DBuf<int>* ctl;
    HBuf<int>* hctl;
    HBuf<double>* hbuf;
    std::vector<double> hpart;
    bool active = false;
    double rhs_norm2 = 0.0;
    int iters = 0;
    double tol_error = 0.0;
    CgArgs a;
};

// CG on Delta for nr (1 or 2) right-hand sides, each to its own stop rule
// ||r||^2 < max(tol^2 ||rhs||^2, DBL_MIN) and cap (Eigen's counting, the
// converged update not counted; src/sinkhorn_cg.cpp:29-38).  Two sides run
// in the same launches (grid rows 0 and 1) and each does exactly the
// arithmetic of a CG of its own, so the iterates are the same bit for bit.
static void cg_delta_sides(regot_problem* p, int nr, CgSide* S, double shift, double tol,
                           int max_iter, bool jacobi)
{
    SparseHessian& H = hess_of(p);
    const int n = p->n, m1 = p->m - 1;
    hipStream_t s = p->stream;
    if (nr < 1 || nr > 2 || (nr == 2 && H.mf))
        fail(E_INTERNAL, "CG over %d right-hand sides (matrix-free: %d)", nr, H.mf ? 1 : 0);
    if (m1 <= 0) {
        for (int k = 0; k < nr; k++) {
            S[k].iters = 0;
            S[k].tol_error = 0.0;
        }
        return;
    }
    H.vn.ensure(n);
    const int lpr_r = lanes_for(static_cast<double>(H.nnz) / std::max(n, 1), n);
    const int lpr_c = lanes_for(static_cast<double>(H.nnz) / m1, m1);
    // column workgroups: 256 columns each when the products are passes over M
    const int nbc = H.mf ? ceil_div(m1, 256) : ceil_div(m1, 4 * (64 / lpr_c));
    const int nbu = ceil_div(m1, 256);
    const int nbs = ceil_div(m1, 1024);   // set-up partials
    if (jacobi) {
        H.cg_dinv.ensure(m1);
        if (H.mf) {
            mf_pass(p, H, MF_DIAG, nullptr, nullptr, shift, nullptr);
            launch_mf_fold(H.mf_part, H.mf_nwg, m1, 4, nullptr, H.h2, shift, H.cg_dinv, s);
        } else {
            hipLaunchKernelGGL(k_delta_diag, dim3(ceil_div(m1, 4)), dim3(256), 0, s, H.colptr.p,
                               H.rowind.p, H.cval.p, H.h1.p, H.h2.p, shift, m1, H.cg_dinv.p);
            KERNEL_CHECK();
        }
    }
    const double* dinv = jacobi ? H.cg_dinv.p : nullptr;

    // per side: x = x0 (or 0), r = rhs - Delta x, u = D^-1 r (or r),
    // w = Delta u (in q), p = s = 0 (beta_0 = 0)
    for (int k = 0; k < nr; k++) {
        CgSide& c = S[k];
        c.r->ensure(m1);
        c.pv->ensure(m1);
        c.q->ensure(m1);
        c.sv->ensure(m1);
        c.u->ensure(m1);
        c.y->ensure(n);
        c.part->ensure(std::max(2L * nbc + nbu, 2L * nbs));
        c.scal->ensure(8);
        c.ctl->ensure(4);
        c.hctl->ensure(16);
        c.hbuf->ensure(8);
        if (c.x0) {
            vec_copy(c.x, c.x0, m1, s);
            hess_apply_delta(p, c.x0, shift, *c.q);
            vec_axpby(*c.r, 1.0, c.rhs, -1.0, *c.q, m1, s);
        } else {
            launch_fill(c.x, 0.0, m1, s);
            vec_copy(*c.r, c.rhs, m1, s);
        }
        hipLaunchKernelGGL(k_pcg_start, dim3(nbs), dim3(256), 0, s, c.r->p, dinv, m1, c.u->p,
                           c.part->p);
        KERNEL_CHECK();
        hess_apply_delta(p, *c.u, shift, *c.q);
        const double* xs[2] = {c.rhs, c.q->p};
        const double* ys[2] = {c.rhs, c.u->p};
        vec_dots(xs, ys, m1, 2, c.scal->p + 6, s);
        c.hpart.assign(2L * nbs, 0.0);
        HIP_CHECK(hipMemcpyAsync(c.hpart.data(), c.part->p, 2L * nbs * sizeof(double),
                                 hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipMemcpyAsync(c.hbuf->p, c.scal->p + 6, 2 * sizeof(double),
                                 hipMemcpyDeviceToHost, s));
        launch_fill(*c.pv, 0.0, m1, s);
        launch_fill(*c.sv, 0.0, m1, s);
    }
    HIP_CHECK(hipStreamSynchronize(s));
    int act[2] = {0, 0};
    int na = 0;
    for (int k = 0; k < nr; k++) {
        CgSide& c = S[k];
        c.rhs_norm2 = c.hbuf->p[0];
        const double delta0 = c.hbuf->p[1];
        double rz0 = 0.0, rr0 = 0.0;
        for (int b = 0; b < nbs; b++) {
            rz0 += c.hpart[2 * b];
            rr0 += c.hpart[2 * b + 1];
        }
        c.iters = 0;
        c.active = false;
        if (c.rhs_norm2 == 0.0) {
            if (c.x0) launch_fill(c.x, 0.0, m1, s);
            c.tol_error = 0.0;
            continue;
        }
        double thr = tol * tol * c.rhs_norm2;
        if (thr < 2.2250738585072014e-308) thr = 2.2250738585072014e-308;
        // converged at the start (tol >= 1 without a guess, or an exact guess)
        if (rr0 < thr || max_iter <= 0) {
            c.tol_error = std::sqrt(rr0 / c.rhs_norm2);
            continue;
        }
        double hs[7] = {rz0 / delta0, 0.0, rz0, 0.0, rr0, thr, c.rhs_norm2};
        HIP_CHECK(hipMemcpyAsync(c.scal->p, hs, 7 * sizeof(double), hipMemcpyHostToDevice, s));
        int hc[4] = {0, 0, 0, 0};
        HIP_CHECK(hipMemcpyAsync(c.ctl->p, hc, 4 * sizeof(int), hipMemcpyHostToDevice, s));
        CgArgs& a = c.a;
        a.rowptr = H.rowptr;
        a.colind = H.colind;
        a.rval = H.rval;
        a.colptr = H.colptr;
        a.rowind = H.rowind;
        a.cval = H.cval;
        a.h1 = H.h1;
        a.h2 = H.h2;
        a.shift = shift;
        a.n = n;
        a.m1 = m1;
        a.x = c.x;
        a.r = *c.r;
        a.p = *c.pv;
        a.s = *c.sv;
        a.w = *c.q;
        a.u = *c.u;
        a.y = *c.y;
        a.part = *c.part;
        a.scal = *c.scal;
        a.ctl = *c.ctl;
        a.nbc = nbc;
        a.nbu = nbu;
        a.dinv = dinv;
        c.active = true;
        act[na++] = k;
    }
    if (na == 0) return;
    const CgArgs& a0 = S[act[0]].a;
    const CgArgs& a1 = S[act[na - 1]].a;

    // batches of K iterations, two in flight; iteration indices run 0 ..
    // max_iter - 1 and nothing is enqueued past them
    const int K = 16;
    int k_enq = 0;
    hipEvent_t ev[2];
    HIP_CHECK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    // matrix-free: the pass over M makes the stop test and y = (sigma u) /
    // (h1 + shift) and sigma' y in one read; k_cg_cols_mf folds its partials
    MfArgs ma;
    if (H.mf) {
        ma = mf_args(p, H);
        ma.xc = a0.u;
        ma.shift = shift;
        ma.part = H.mf_part;
        ma.ctl = a0.ctl;
        ma.rr_part = a0.part + 2 * nbc;
        ma.rr_n = nbu;
        ma.scal = a0.scal;
    }
    auto launch_batch = [&](int slot) {
        for (int q = 0; q < K && k_enq < max_iter; q++, k_enq++) {
            hipLaunchKernelGGL(k_cg_update, dim3(nbu, na), dim3(256), 0, s, a0, a1, k_enq);
            if (H.mf) {
                ma.it = k_enq;
                launch_mf_pass(MF_CG, ma, H.mf_nwg, s);
                hipLaunchKernelGGL(k_cg_cols_mf, dim3(nbc), dim3(256), 0, s, a0, H.mf_part.p,
                                   H.mf_nwg);
            } else {
                cg_rows(lpr_r, a0, a1, na, s);
                cg_cols(lpr_c, a0, a1, na, k_enq, s);
            }
        }
        KERNEL_CHECK();
        for (int j = 0; j < na; j++) {
            CgSide& c = S[act[j]];
            HIP_CHECK(hipMemcpyAsync(c.hctl->p + 4 * slot, c.ctl->p, 4 * sizeof(int),
                                     hipMemcpyDeviceToHost, s));
        }
        HIP_CHECK(hipEventRecord(ev[slot], s));
    };
    auto all_done = [&](int slot) {
        for (int j = 0; j < na; j++)
            if (!S[act[j]].hctl->p[4 * slot]) return false;
        return true;
    };
    try {
        std::deque<int> qu;
        launch_batch(0);
        qu.push_back(0);
        if (k_enq < max_iter) {
            launch_batch(1);
            qu.push_back(1);
        }
        while (!qu.empty()) {
            const int slot = qu.front();
            qu.pop_front();
            HIP_CHECK(hipEventSynchronize(ev[slot]));
            if (all_done(slot)) break;
            if (k_enq < max_iter) {
                launch_batch(slot);
                qu.push_back(slot);
            }
        }
        HIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
        (void)hipStreamSynchronize(s);
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        throw;
    }
    HIP_CHECK(hipEventDestroy(ev[0]));
    HIP_CHECK(hipEventDestroy(ev[1]));
    for (int j = 0; j < na; j++) {
        CgSide& c = S[act[j]];
        HIP_CHECK(hipMemcpy(c.hctl->p + 8, c.ctl->p, 4 * sizeof(int), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(c.hbuf->p, c.scal->p, 5 * sizeof(double), hipMemcpyDeviceToHost));
        c.iters = c.hctl->p[8] ? c.hctl->p[9] : max_iter;
        c.tol_error = std::sqrt(c.hbuf->p[4] / c.rhs_norm2);
    }
}

// Side k of the Hessian's CG state (0: cg_*, 1: cg1_*).
static CgSide cg_side(SparseHessian& H, int k)
{
    CgSide c;
    if (k == 0) {
        c.r = &H.cg_r;
        c.pv = &H.cg_p;
        c.q = &H.cg_q;
        c.sv = &H.cg_s;
        c.u = &H.cg_u;
        c.y = &H.cg_y_buf;
        c.part = &H.cg_part;
        c.scal = &H.cg_scal;
        c.ctl = &H.cg_ctl;
        c.hctl = &H.cg_hctl;
        c.hbuf = &H.hbuf;
    } else {
        c.r = &H.cg1_r;
        c.pv = &H.cg1_p;
        c.q = &H.cg1_q;
        c.sv = &H.cg1_s;
        c.u = &H.cg1_u;
        c.y = &H.cg1_y;
        c.part = &H.cg1_part;
        c.scal = &H.cg1_scal;
        c.ctl = &H.cg1_ctl;
        c.hctl = &H.cg1_hctl;
        c.hbuf = &H.hbuf1;
    }
    return c;
}

int hess_cg_delta(regot_problem* p, const double* rhs, double shift, double tol, int max_iter,
                  double* x, double* tol_error, const double* x0, bool jacobi)
{
    CgSide c = cg_side(hess_of(p), 0);
    c.rhs = rhs;
    c.x = x;
    c.x0 = x0;
    cg_delta_sides(p, 1, &c, shift, tol, max_iter, jacobi);
    if (tol_error) *tol_error = c.tol_error;
    return c.iters;
}

// The Schur right-hand side z - sigma' (w / (h1 + s)) of rhs = [w; z] into
// zs (m-1), with ybuf (n) as scratch.
static void cg_schur_rhs(regot_problem* p, SparseHessian& H, const double* rhs, double shift,
                         double* ybuf, double* zs)
{
    const int n = p->n, m1 = p->m - 1;
    hipStream_t s = p->stream;
    hipLaunchKernelGGL(k_diag_axpy, dim3(ceil_div(n, 256)), dim3(256), 0, s, ybuf, H.h1.p, rhs,
                       shift, n, 1);
    KERNEL_CHECK();
    if (H.mf) {
        mf_pass(p, H, MF_COLS, nullptr, ybuf, 0.0, nullptr);
        launch_mf_fold(H.mf_part, H.mf_nwg, m1, 1, rhs + n, nullptr, 0.0, zs, s);
    } else {
        launch_spmv_cols(H, ybuf, nullptr, 0.0, 0, zs, s);
        vec_axpby(zs, -1.0, zs, 1.0, rhs + n, m1, s);
    }
}

// out[0:n] = (w - sigma r2) / (h1 + s) from the Schur solution r2 = out[n:].
static void cg_back(regot_problem* p, SparseHessian& H, const double* rhs, double shift,
                    double* ybuf, double* out)
{
    const int n = p->n;
    hipStream_t s = p->stream;
    if (H.mf)
        mf_pass(p, H, MF_ROWS, out + n, nullptr, 0.0, ybuf);
    else
        launch_spmv_rows(H, out + n, ybuf, false, 0.0, s);
    hipLaunchKernelGGL(k_sub_div, dim3(ceil_div(n, 256)), dim3(256), 0, s, rhs, ybuf, H.h1.p,
                       shift, n, out);
    KERNEL_CHECK();
}

// (Hs + shift I)^{-1} rhs by block elimination (src/sinkhorn_cg.cpp:9-50):
// v = w / (h1 + s); y = sigma' v; solve Delta r2 = z - y; r1 = (w - sigma r2) / (h1 + s).
int hess_cg(regot_problem* p, const double* rhs, double shift, double tol, int max_iter,
            double* out, double* tol_error, const double* x0, bool jacobi)
{
    SparseHessian& H = hess_of(p);
    const int n = p->n, m1 = p->m - 1;
    if (max_iter <= 0) max_iter = 2 * p->m;
    H.vn.ensure(n);
    H.vm2.ensure(std::max(m1, 1));
    H.cg_y_buf.ensure(n);
    cg_schur_rhs(p, H, rhs, shift, H.cg_y_buf, H.vm2);
    const int it = hess_cg_delta(p, H.vm2, shift, tol, max_iter, out + n, tol_error,
                                 x0 ? x0 + n : nullptr, jacobi);
    cg_back(p, H, rhs, shift, H.cg_y_buf, out);
    return it;
}

int hess_cg2(regot_problem* p, const double* rhs0, const double* rhs1, double shift, double tol,
             int max_iter, double* out0, double* out1, int* iters, double* tol_error,
             bool jacobi)
{
    SparseHessian& H = hess_of(p);
    const int n = p->n, m1 = p->m - 1;
    if (H.mf) {
        // the matrix-free CG serves one right-hand side: two solves
        double e0 = 0.0, e1 = 0.0;
        iters[0] = hess_cg(p, rhs0, shift, tol, max_iter, out0, &e0, nullptr, jacobi);
        iters[1] = hess_cg(p, rhs1, shift, tol, max_iter, out1, &e1, nullptr, jacobi);
        if (tol_error) {
            tol_error[0] = e0;
            tol_error[1] = e1;
        }
        return iters[0] + iters[1];
    }
    if (max_iter <= 0) max_iter = 2 * p->m;
    H.vn.ensure(n);
    H.vm2.ensure(std::max(m1, 1));
    H.cg1_rhs.ensure(std::max(m1, 1));
    H.cg_y_buf.ensure(n);
    H.cg1_y.ensure(n);
    cg_schur_rhs(p, H, rhs0, shift, H.cg_y_buf, H.vm2);
    cg_schur_rhs(p, H, rhs1, shift, H.cg1_y, H.cg1_rhs);
    CgSide c[2] = {cg_side(H, 0), cg_side(H, 1)};
    c[0].rhs = H.vm2;
    c[0].x = out0 + n;
    c[1].rhs = H.cg1_rhs;
    c[1].x = out1 + n;
    cg_delta_sides(p, 2, c, shift, tol, max_iter, jacobi);
    cg_back(p, H, rhs0, shift, H.cg_y_buf, out0);
    cg_back(p, H, rhs1, shift, H.cg1_y, out1);
    iters[0] = c[0].iters;
    iters[1] = c[1].iters;
    if (tol_error) {
        tol_error[0] = c[0].tol_error;
        tol_error[1] = c[1].tol_error;
    }
    return iters[0] + iters[1];
}

}  // namespace regot

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------

using namespace regot;

namespace regot {
int set_error(int code, const char* what);
}

#define CAPI_BEGIN try {
#define CAPI_END                                                                          \
    }                                                                                     \
    catch (const regot::Error& e) { return set_error(e.code, e.what()); }                 \
    catch (const std::bad_alloc&) { return set_error(E_NOMEM, "host allocation failed"); } \
    catch (const std::exception& e) { return set_error(E_INTERNAL, e.what()); }            \
    return OK;

static void hess_check(regot_problem* p, bool need_built)
{
    if (!p) fail(E_INVALID, "null problem handle");
    HIP_CHECK(hipSetDevice(p->device));
    if (need_built && (!p->hess || (!p->hess->mf && !p->hess->rowptr.p)))
        fail(E_STATE, "no sparsified Hessian: call regot_sparsified_hess* first");
}

// gamma (host) -> device gamma and grad in p->gamma / p->grad.
static void upload_and_grad(regot_problem* p, const double* gamma)
{
    const long N = p->n + p->m - 1;
    HIP_CHECK(hipMemcpyAsync(p->gamma, gamma, N * sizeof(double), hipMemcpyHostToDevice,
                             p->stream));
    (void)p->dev_obj_grad(p->gamma, p->grad, nullptr);
}

extern "C" {

int regot_sparsified_hess(regot_problem* p, const double* gamma, double delta,
                          double density_hint, long long* nnz)
{
    CAPI_BEGIN
    hess_check(p, false);
    if (!gamma) fail(E_INVALID, "null gamma");
    upload_and_grad(p, gamma);
    hess_threshold(p, p->gamma, p->grad, delta, density_hint);
    if (nnz) *nnz = p->hess->nnz + p->hess->nnz_zero;
    CAPI_END
}

int regot_sparsified_hess_with_density(regot_problem* p, const double* gamma, double density,
                                       long long* nnz)
{
    CAPI_BEGIN
    hess_check(p, false);
    if (!gamma) fail(E_INVALID, "null gamma");
    upload_and_grad(p, gamma);
    hess_topk(p, p->gamma, p->grad, density);
    if (nnz) *nnz = p->hess->nnz;
    CAPI_END
}

int regot_set_hess_storage(regot_problem* p, int mode)
{
    CAPI_BEGIN
    if (!p) fail(E_INVALID, "null problem handle");
    if (mode < -1 || mode > 1) fail(E_INVALID, "storage mode must be -1, 0 or 1 (got %d)", mode);
    p->hess_storage = mode;
    CAPI_END
}

int regot_hess_info(regot_problem* p, long long* nnz_stored, long long* nnz_zero, double* density)
{
    CAPI_BEGIN
    hess_check(p, true);
    if (nnz_stored) *nnz_stored = p->hess->nnz;
    if (nnz_zero) *nnz_zero = p->hess->nnz_zero;
    if (density) *density = p->hess->density();
    CAPI_END
}

int regot_hess_export(regot_problem* p, double* h1, double* h2, int* colptr, int* rowind,
                      double* cval, int* rowptr, int* colind, double* rval)
{
    CAPI_BEGIN
    hess_check(p, true);
    SparseHessian& H = *p->hess;
    if (H.mf)
        fail(E_STATE, "the Hessian was built matrix-free and holds no CSR / CSC "
                      "(regot_set_hess_storage(p, 0) keeps them)");
    const int n = p->n, m = p->m;
    const long long z = H.nnz;
    auto cp = [&](void* dst, const void* src, size_t bytes) {
        if (dst && bytes) HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    };
    cp(h1, H.h1.p, n * sizeof(double));
    cp(h2, H.h2.p, (m - 1) * sizeof(double));
    cp(colptr, H.colptr.p, m * sizeof(int));
    cp(rowind, H.rowind.p, z * sizeof(int));
    cp(cval, H.cval.p, z * sizeof(double));
    cp(rowptr, H.rowptr.p, (n + 1) * sizeof(int));
    cp(colind, H.colind.p, z * sizeof(int));
    cp(rval, H.rval.p, z * sizeof(double));
    CAPI_END
}

int regot_hess_apply_delta(regot_problem* p, const double* x, double shift, double* res)
{
    CAPI_BEGIN
    hess_check(p, true);
    const int m1 = p->m - 1;
    DBuf<double> dx(m1), dr(m1);
    HIP_CHECK(hipMemcpyAsync(dx, x, m1 * sizeof(double), hipMemcpyHostToDevice, p->stream));
    hess_apply_delta(p, dx, shift, dr);
    HIP_CHECK(hipMemcpyAsync(res, dr, m1 * sizeof(double), hipMemcpyDeviceToHost, p->stream));
    p->sync();
    CAPI_END
}

int regot_hess_apply_hs(regot_problem* p, const double* x, double* res)
{
    CAPI_BEGIN
    hess_check(p, true);
    const long N = p->n + p->m - 1;
    DBuf<double> dx(N), dr(N);
    HIP_CHECK(hipMemcpyAsync(dx, x, N * sizeof(double), hipMemcpyHostToDevice, p->stream));
    hess_apply_hs(p, dx, dr);
    HIP_CHECK(hipMemcpyAsync(res, dr, N * sizeof(double), hipMemcpyDeviceToHost, p->stream));
    p->sync();
    CAPI_END
}

int regot_hess_cg2(regot_problem* p, const double* rhs0, const double* rhs1, double shift,
                   double tol, int max_iter, double* res0, double* res1, int* iters,
                   double* tol_error)
{
    CAPI_BEGIN
    hess_check(p, true);
    if (!rhs0 || !rhs1 || !res0 || !res1 || !iters) fail(E_INVALID, "null argument");
    const long N = p->n + p->m - 1;
    DBuf<double> d0(N), d1(N), r0(N), r1(N);
    HIP_CHECK(hipMemcpyAsync(d0, rhs0, N * sizeof(double), hipMemcpyHostToDevice, p->stream));
    HIP_CHECK(hipMemcpyAsync(d1, rhs1, N * sizeof(double), hipMemcpyHostToDevice, p->stream));
    double te[2] = {0.0, 0.0};
    hess_cg2(p, d0, d1, shift, tol, max_iter, r0, r1, iters, te);
    HIP_CHECK(hipMemcpyAsync(res0, r0, N * sizeof(double), hipMemcpyDeviceToHost, p->stream));
    HIP_CHECK(hipMemcpyAsync(res1, r1, N * sizeof(double), hipMemcpyDeviceToHost, p->stream));
    p->sync();
    if (tol_error) {
        tol_error[0] = te[0];
        tol_error[1] = te[1];
    }
    CAPI_END
}

int regot_hess_cg(regot_problem* p, const double* rhs, double shift, double tol, int max_iter,
                  double* res, int* iters, double* tol_error)
{
    CAPI_BEGIN
    hess_check(p, true);
    const long N = p->n + p->m - 1;
    DBuf<double> dx(N), dr(N);
    HIP_CHECK(hipMemcpyAsync(dx, rhs, N * sizeof(double), hipMemcpyHostToDevice, p->stream));
    double te = 0.0;
    const int it = hess_cg(p, dx, shift, tol, max_iter, dr, &te);
    HIP_CHECK(hipMemcpyAsync(res, dr, N * sizeof(double), hipMemcpyDeviceToHost, p->stream));
    p->sync();
    if (iters) *iters = it;
    if (tol_error) *tol_error = te;
    CAPI_END
}

}  // extern "C"
