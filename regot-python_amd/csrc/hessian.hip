// hessian.hip -- building the sparsified Hessian from the current dual point,
// its products, and the block-elimination CG solve, plus their C-ABI.
#include <cmath>
#include <cstring>
#include <deque>
#include <vector>

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
// Conjugate gradients, optionally with the Jacobi preconditioner z = D^-1 r
// (dinv = 1 / diag(Delta); null: plain CG, z = r).
// ctl: [0] done, [1] iterations.  scal: [0] rho = r.z (current), [1] stop
// threshold on r.r, [2] rho (next, published by the direction kernel),
// [3] rhs norm^2, [4] r.r of the last update.
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
    double* q;
    double* y;      // n
    double* part;   // per column-workgroup partials (2 arrays)
    double* scal;
    int* ctl;
    int max_iter;
    int nbc;        // column workgroups (4 columns each)
    int nbu;        // update workgroups (1024 entries each)
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

// z = dinv .* r (or r), p = z, partials of r.z and r.r per 1024 entries.
__global__ __launch_bounds__(256) void k_pcg_start(const double* __restrict__ r,
                                                   const double* __restrict__ dinv, int m1,
                                                   double* __restrict__ p,
                                                   double* __restrict__ part)
{
    __shared__ double sm[2][256];
    double rz = 0.0, rr = 0.0;
    for (int j = blockIdx.x * 1024 + threadIdx.x; j < (blockIdx.x + 1) * 1024 && j < m1;
         j += 256) {
        const double rj = r[j];
        const double zj = dinv ? rj * dinv[j] : rj;
        p[j] = zj;
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

// y = (sigma p) / (h1 + shift), one wave per row, the waves of the grid
// striding over the rows (the grid is one resident round); workgroup 0 also
// moves the next rho into place.
__global__ __launch_bounds__(256) void k_cg_rows(CgArgs a)
{
    if (a.ctl[0]) return;
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.ctl[1] > 0) a.scal[0] = a.scal[2];
    const int lane = threadIdx.x & 63;
    const int nw = gridDim.x * 4;
    for (int i = blockIdx.x * 4 + (threadIdx.x >> 6); i < a.n; i += nw) {
        double s = 0.0;
        for (int k = a.rowptr[i] + lane; k < a.rowptr[i + 1]; k += 64)
            s += a.rval[k] * a.p[a.colind[k]];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (lane == 0) a.y[i] = s / (a.h1[i] + a.shift);
    }
}

__device__ __forceinline__ double wg_sum4(double v, double* sm)
{
    // lane 0 of each wave holds a value; fixed-order sum of the 4
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sm[wave] = v;
    __syncthreads();
    const double r = ((sm[0] + sm[1]) + sm[2]) + sm[3];
    __syncthreads();
    return r;
}

// q = (h2 + shift) p - sigma' y, one wave per column, the waves of the grid
// striding over the columns; partial p.q per workgroup (fixed order).
__global__ __launch_bounds__(256) void k_cg_cols(CgArgs a)
{
    __shared__ double sm[4];
    if (a.ctl[0]) return;
    const int lane = threadIdx.x & 63;
    const int nw = gridDim.x * 4;
    double pq = 0.0;
    for (int j = blockIdx.x * 4 + (threadIdx.x >> 6); j < a.m1; j += nw) {
        double s = 0.0;
        for (int k = a.colptr[j] + lane; k < a.colptr[j + 1]; k += 64)
            s += a.cval[k] * a.y[a.rowind[k]];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        const double qj = a.h2[j] * a.p[j] + a.shift * a.p[j] - s;
        if (lane == 0) a.q[j] = qj;
        pq += a.p[j] * qj;
    }
    pq = wg_sum4(pq, sm);
    if (threadIdx.x == 0) a.part[blockIdx.x] = pq;
}

// Every workgroup reduces the p.q partials (same order, same value), then
// x += alpha p, r -= alpha q over its 1024 entries; partials of r.z and r.r.
__global__ __launch_bounds__(256) void k_cg_update(CgArgs a)
{
    __shared__ double sm[2][256];
    if (a.ctl[0]) return;
    double v = 0.0;
    for (int k = threadIdx.x; k < a.nbc; k += 256) v += a.part[k];
    sm[0][threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sm[0][threadIdx.x] += sm[0][threadIdx.x + o];
        __syncthreads();
    }
    const double pq = sm[0][0];
    const double alpha = a.scal[0] / pq;
    __syncthreads();
    double rr = 0.0, rz = 0.0;
    for (int j = blockIdx.x * 1024 + threadIdx.x; j < (blockIdx.x + 1) * 1024 && j < a.m1; j += 256) {
        a.x[j] += alpha * a.p[j];
        const double rj = a.r[j] - alpha * a.q[j];
        a.r[j] = rj;
        rr += rj * rj;
        rz += a.dinv ? rj * rj * a.dinv[j] : 0.0;
    }
    sm[0][threadIdx.x] = rr;
    sm[1][threadIdx.x] = rz;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + o];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.part[a.nbc + 2 * blockIdx.x] = sm[0][0];
        a.part[a.nbc + 2 * blockIdx.x + 1] = sm[1][0];
    }
}

// Every workgroup reduces r.r and r.z; stop (rr < threshold, or the
// iteration cap after this update), else p = z + (rho_new / rho) p.
__global__ __launch_bounds__(256) void k_cg_dir(CgArgs a)
{
    __shared__ double sm[2][256];
    if (a.ctl[0]) return;
    double v = 0.0, w = 0.0;
    for (int k = threadIdx.x; k < a.nbu; k += 256) {
        v += a.part[a.nbc + 2 * k];
        w += a.part[a.nbc + 2 * k + 1];
    }
    sm[0][threadIdx.x] = v;
    sm[1][threadIdx.x] = w;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + o];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    const double rr_new = sm[0][0];
    const double rho_new = a.dinv ? sm[1][0] : rr_new;
    const bool conv = rr_new < a.scal[1];
    if (!conv) {
        const double beta = rho_new / a.scal[0];
        for (int j = blockIdx.x * 1024 + threadIdx.x; j < (blockIdx.x + 1) * 1024 && j < a.m1;
             j += 256)
            a.p[j] = (a.dinv ? a.r[j] * a.dinv[j] : a.r[j]) + beta * a.p[j];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.scal[2] = rho_new;
        a.scal[4] = rr_new;
        if (conv) {
            a.ctl[0] = 1;
        } else {
            const int it = a.ctl[1] + 1;
            a.ctl[1] = it;
            if (it >= a.max_iter) a.ctl[0] = 1;
        }
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

// Shared tail of both rules: row counts -> scan -> CSR write -> CSC by a
// counting sort of the CSR.
static void build_storage(regot_problem* p, SparseHessian& H, const ElemCtx& c,
                          const KeepRule& rule, bool row_counts_ready)
{
    const int n = p->n, m = p->m;
    hipStream_t s = p->stream;
    if (!row_counts_ready) launch_row_count(p->bulk, p->dM, p->ldm, c, rule, H.rcnt, s);
    H.rowptr.ensure(n + 1);
    H.ltmp.ensure(4);
    launch_scan(H.rcnt, n, H.rowptr, H.ltmp.p, s);
    H.hll.ensure(4);
    HIP_CHECK(hipMemcpyAsync(H.hll.p, H.ltmp.p, sizeof(long long), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    const long long nnz = H.hll.p[0];
    if (nnz > 0x7fffffffLL) fail(E_INVALID, "sparsified Hessian has %lld entries (> 2^31)", nnz);
    H.nnz = nnz;
    H.colind.ensure(nnz);
    H.rval.ensure(nnz);
    H.rowind.ensure(nnz);
    H.cval.ensure(nnz);
    launch_row_write(p->bulk, p->dM, p->ldm, c, rule, H.rowptr, H.colind, H.rval, s);
    // CSC: the transpose of the CSR just written (a counting sort of nnz
    // entries instead of two more passes over M)
    // at most 512 row blocks: the offset table stays <= 2 KiB per column
    const int rows_per_block = std::max(32, ceil_div(n, 512));
    H.ccnt.ensure(m);
    H.colptr.ensure(m);
    H.csc_work.ensure(static_cast<size_t>(ceil_div(n, rows_per_block)) * std::max(m - 1, 1));
    launch_csr_to_csc(H, H.csc_work, rows_per_block, H.ccnt, H.colptr, H.ltmp.p + 1, H.rowind,
                      H.cval, s);
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
    launch_split_gamma(d_gamma, n, m, H.vn, H.vm, s);
    hipLaunchKernelGGL(k_hess_diag, dim3(ceil_div(std::max(n, m), 256)), dim3(256), 0, s, d_grad,
                       p->a.p, p->b.p, n, m, p->reg, H.h1.p, H.h2.p);
    KERNEL_CHECK();
}

void hess_threshold(regot_problem* p, const double* d_gamma, const double* d_grad, double delta,
                    double density_hint)
{
    SparseHessian& H = hess_of(p);
    prepare(p, H, d_gamma, d_grad);
    const ElemCtx c = make_ctx(p, H);
    hipStream_t s = p->stream;
    H.tau.ensure(p->m);
    H.rho.ensure(p->n);
    hess_ensure_transpose(p);
    launch_col_threshold(p->bulk, p->dMT, p->ldt, c, delta, density_hint > 0.2 ? 1 : 0, H.tau, s);
    launch_row_threshold(p->bulk, p->dM, p->ldm, c, H.tau, 2.0 * delta, H.rho, H.rcnt, H.zcnt, s);
    KeepRule rule;
    rule.mode = 0;
    rule.tau = H.tau;
    rule.rho = H.rho;
    build_storage(p, H, c, rule, true);
    // FP64-zero kept entries count towards the reported density
    launch_scan(H.zcnt, p->n, H.rcnt, H.ltmp.p + 2, s);
    HIP_CHECK(hipMemcpyAsync(H.hll.p, H.ltmp.p, 3 * sizeof(long long), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    H.nnz_zero = H.hll.p[2];
    if (H.hll.p[1] != H.nnz)
        fail(E_INTERNAL, "CSR and CSC counts differ (%lld vs %lld)", H.nnz, H.hll.p[1]);
}

void hess_topk(regot_problem* p, const double* d_gamma, const double* d_grad, double density)
{
    SparseHessian& H = hess_of(p);
    prepare(p, H, d_gamma, d_grad);
    const ElemCtx c = make_ctx(p, H);
    hipStream_t s = p->stream;
    const int n = p->n, m = p->m;
    density = std::min(1.0, std::max(0.0, density));
    const long long len = static_cast<long long>(n) * (m - 1);
    const long long kk = static_cast<long long>(static_cast<double>(len) * density);
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
        unsigned long long kstar = 0;
        long long above = 0;
        topk_select(p->bulk, p->dM, p->ldm, c, kk, H, &kstar, &above, s);
        rule.kstar = kstar;
        rule.tie_take = kk - above;
    }
    // tie offsets per row (row-major order)
    launch_row_ties(p->bulk, p->dM, p->ldm, c, rule.kstar, H.rcnt, s);
    H.ltmp.ensure(n + 4);
    launch_scan64(H.rcnt, n, H.ltmp.p + 4, s);
    rule.tie_off = H.ltmp.p + 4;
    H.nnz_zero = 0;
    build_storage(p, H, c, rule, false);
    HIP_CHECK(hipMemcpyAsync(H.hll.p, H.ltmp.p, 2 * sizeof(long long), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (H.hll.p[1] != H.nnz)
        fail(E_INTERNAL, "CSR and CSC counts differ (%lld vs %lld)", H.nnz, H.hll.p[1]);
}

// Delta x = h2 .* x + shift x - sigma' ((sigma x) ./ (h1 + shift)); x, out: m-1.
void hess_apply_delta(regot_problem* p, const double* x, double shift, double* out)
{
    SparseHessian& H = hess_of(p);
    H.cg_y_buf.ensure(p->n);
    launch_spmv_rows(H, x, H.cg_y_buf, true, shift, p->stream);
    launch_spmv_cols(H, H.cg_y_buf, x, shift, 1, out, p->stream);
}

// Hs x = [h1 .* w + sigma z ; h2 .* z + sigma' w]; x, out: n+m-1.
void hess_apply_hs(regot_problem* p, const double* x, double* out)
{
    SparseHessian& H = hess_of(p);
    const int n = p->n, m = p->m;
    launch_spmv_rows(H, x + n, out, false, 0.0, p->stream);
    hipLaunchKernelGGL(k_diag_axpy, dim3(ceil_div(n, 256)), dim3(256), 0, p->stream, out, H.h1.p,
                       x, 0.0, n, 0);
    launch_spmv_cols(H, x, x + n, 0.0, 2, out + n, p->stream);
    (void)m;
    KERNEL_CHECK();
}

// CG on Delta (size m-1) from zero; textbook unpreconditioned CG with the
// stopping rule ||r||^2 < max(tol^2 ||b||^2, DBL_MIN) and the iteration cap
// counted as Eigen's ConjugateGradient counts it (src/sinkhorn_cg.cpp:29-38).
int hess_cg_delta(regot_problem* p, const double* rhs, double shift, double tol, int max_iter,
                  double* x, double* tol_error, const double* x0, bool jacobi)
{
    SparseHessian& H = hess_of(p);
    const int n = p->n, m1 = p->m - 1;
    hipStream_t s = p->stream;
    if (m1 <= 0) {
        if (tol_error) *tol_error = 0.0;
        return 0;
    }
    H.cg_r.ensure(m1);
    H.cg_p.ensure(m1);
    H.cg_q.ensure(m1);
    H.vn.ensure(n);
    H.cg_y_buf.ensure(n);
    // one resident round of workgroups for the SpMV kernels (8 per CU)
    const int nbr = std::min(ceil_div(n, 4), 2048);
    const int nbc = std::min(ceil_div(m1, 4), 2048);
    const int nbu = ceil_div(m1, 1024);
    H.cg_part.ensure(nbc + 2L * nbu);
    H.cg_scal.ensure(8);
    if (jacobi) {
        H.cg_dinv.ensure(m1);
        hipLaunchKernelGGL(k_delta_diag, dim3(ceil_div(m1, 4)), dim3(256), 0, s, H.colptr.p,
                           H.rowind.p, H.cval.p, H.h1.p, H.h2.p, shift, m1, H.cg_dinv.p);
        KERNEL_CHECK();
    }
    const double* dinv = jacobi ? H.cg_dinv.p : nullptr;
    H.cg_ctl.ensure(4);
    H.cg_hctl.ensure(16);
    H.hbuf.ensure(8);

    // x = x0 (or 0), r = rhs - Delta x, p = z = D^-1 r (or r)
    if (x0) {
        vec_copy(x, x0, m1, s);
        hess_apply_delta(p, x0, shift, H.cg_q);
        vec_axpby(H.cg_r, 1.0, rhs, -1.0, H.cg_q, m1, s);
    } else {
        launch_fill(x, 0.0, m1, s);
        vec_copy(H.cg_r, rhs, m1, s);
    }
    hipLaunchKernelGGL(k_pcg_start, dim3(nbu), dim3(256), 0, s, H.cg_r.p, dinv, m1, H.cg_p.p,
                       H.cg_part.p);
    KERNEL_CHECK();
    const double* xs[1] = {rhs};
    vec_dots(xs, xs, m1, 1, H.cg_scal.p + 3, s);
    std::vector<double> hpart(2L * nbu);
    HIP_CHECK(hipMemcpyAsync(hpart.data(), H.cg_part.p, 2L * nbu * sizeof(double),
                             hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(H.hbuf.p, H.cg_scal.p + 3, sizeof(double), hipMemcpyDeviceToHost,
                             s));
    HIP_CHECK(hipStreamSynchronize(s));
    const double rhs_norm2 = H.hbuf.p[0];
    double rz0 = 0.0, rr0 = 0.0;
    for (int b = 0; b < nbu; b++) {
        rz0 += hpart[2 * b];
        rr0 += hpart[2 * b + 1];
    }
    if (rhs_norm2 == 0.0) {
        if (x0) launch_fill(x, 0.0, m1, s);
        if (tol_error) *tol_error = 0.0;
        return 0;
    }
    double thr = tol * tol * rhs_norm2;
    if (thr < 2.2250738585072014e-308) thr = 2.2250738585072014e-308;
    // converged at the start (tol >= 1 without a guess, or an exact guess)
    if (rr0 < thr) {
        if (tol_error) *tol_error = std::sqrt(rr0 / rhs_norm2);
        return 0;
    }
    double hs[5] = {rz0, thr, rz0, rhs_norm2, rr0};
    HIP_CHECK(hipMemcpyAsync(H.cg_scal.p, hs, 5 * sizeof(double), hipMemcpyHostToDevice, s));
    int hc[4] = {0, 0, 0, 0};
    HIP_CHECK(hipMemcpyAsync(H.cg_ctl.p, hc, 4 * sizeof(int), hipMemcpyHostToDevice, s));
    if (max_iter <= 0) {
        if (tol_error) *tol_error = 1.0;
        HIP_CHECK(hipStreamSynchronize(s));
        return 0;
    }

    CgArgs a;
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
    a.x = x;
    a.r = H.cg_r;
    a.p = H.cg_p;
    a.q = H.cg_q;
    a.y = H.cg_y_buf;
    a.part = H.cg_part;
    a.scal = H.cg_scal;
    a.ctl = H.cg_ctl;
    a.max_iter = max_iter;
    a.nbc = nbc;
    a.nbu = nbu;
    a.dinv = dinv;

    const int K = 16;
    hipEvent_t ev[2];
    HIP_CHECK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    auto launch_batch = [&](int slot) {
        for (int it = 0; it < K; it++) {
            hipLaunchKernelGGL(k_cg_rows, dim3(nbr), dim3(256), 0, s, a);
            hipLaunchKernelGGL(k_cg_cols, dim3(nbc), dim3(256), 0, s, a);
            hipLaunchKernelGGL(k_cg_update, dim3(nbu), dim3(256), 0, s, a);
            hipLaunchKernelGGL(k_cg_dir, dim3(nbu), dim3(256), 0, s, a);
        }
        KERNEL_CHECK();
        HIP_CHECK(hipMemcpyAsync(H.cg_hctl.p + 4 * slot, H.cg_ctl.p, 4 * sizeof(int),
                                 hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipEventRecord(ev[slot], s));
    };
    int iters = 0;
    try {
        std::deque<int> q;
        launch_batch(0);
        q.push_back(0);
        launch_batch(1);
        q.push_back(1);
        for (;;) {
            const int slot = q.front();
            q.pop_front();
            HIP_CHECK(hipEventSynchronize(ev[slot]));
            if (H.cg_hctl.p[4 * slot]) {
                break;
            }
            launch_batch(slot);
            q.push_back(slot);
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
    HIP_CHECK(hipMemcpy(H.cg_hctl.p + 8, H.cg_ctl.p, 4 * sizeof(int), hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(H.hbuf.p, H.cg_scal.p, 5 * sizeof(double), hipMemcpyDeviceToHost));
    iters = H.cg_hctl.p[9];
    if (tol_error) *tol_error = std::sqrt(H.hbuf.p[4] / rhs_norm2);
    return iters;
}

// (Hs + shift I)^{-1} rhs by block elimination (src/sinkhorn_cg.cpp:9-50):
// v = w / (h1 + s); y = sigma' v; solve Delta r2 = z - y; r1 = (w - sigma r2) / (h1 + s).
int hess_cg(regot_problem* p, const double* rhs, double shift, double tol, int max_iter,
            double* out, double* tol_error, const double* x0, bool jacobi)
{
    SparseHessian& H = hess_of(p);
    const int n = p->n, m1 = p->m - 1;
    hipStream_t s = p->stream;
    if (max_iter <= 0) max_iter = 2 * p->m;
    H.vn.ensure(n);
    H.vm2.ensure(std::max(m1, 1));
    H.cg_y_buf.ensure(n);
    hipLaunchKernelGGL(k_diag_axpy, dim3(ceil_div(n, 256)), dim3(256), 0, s, H.cg_y_buf.p, H.h1.p,
                       rhs, shift, n, 1);
    KERNEL_CHECK();
    launch_spmv_cols(H, H.cg_y_buf, nullptr, 0.0, 0, H.vm2, s);
    vec_axpby(H.vm2, -1.0, H.vm2, 1.0, rhs + n, m1, s);
    const int it = hess_cg_delta(p, H.vm2, shift, tol, max_iter, out + n, tol_error,
                                 x0 ? x0 + n : nullptr, jacobi);
    launch_spmv_rows(H, out + n, H.cg_y_buf, false, 0.0, s);
    hipLaunchKernelGGL(k_sub_div, dim3(ceil_div(n, 256)), dim3(256), 0, s, rhs, H.cg_y_buf.p,
                       H.h1.p, shift, n, out);
    KERNEL_CHECK();
    return it;
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
    if (need_built && (!p->hess || !p->hess->rowptr.p))
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
