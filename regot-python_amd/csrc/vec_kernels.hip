// vec_kernels.hip -- FP64 vector helpers for the host drivers.  Reductions use
// one 1024-thread workgroup with a fixed tree, so every call is reproducible.
#include "problem.hpp"

namespace regot {
namespace {

__device__ __forceinline__ double block_reduce(double v, double* sm)
{
    const int t = threadIdx.x;
    sm[t] = v;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (t < o) sm[t] += sm[t + o];
        __syncthreads();
    }
    const double r = sm[0];
    __syncthreads();
    return r;
}

__global__ void k_axpby(double* out, double alpha, const double* x, double beta, const double* y,
                        long count)
{
    for (long i = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; i < count;
         i += static_cast<long>(gridDim.x) * 256)
        out[i] = alpha * x[i] + (y ? beta * y[i] : 0.0);
}

__global__ void k_axpy(double* y, double alpha, const double* x, long count)
{
    for (long i = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; i < count;
         i += static_cast<long>(gridDim.x) * 256)
        y[i] += alpha * x[i];
}

struct DotArgs {
    const double* x[8];
    const double* y[8];
};

__global__ __launch_bounds__(1024) void k_dots(DotArgs d, long count, int npairs, double* out)
{
    __shared__ double sm[1024];
    for (int p = 0; p < npairs; p++) {
        double s = 0.0;
        for (long i = threadIdx.x; i < count; i += 1024) s += d.x[p][i] * d.y[p][i];
        s = block_reduce(s, sm);
        if (threadIdx.x == 0) out[p] = s;
    }
}

__global__ __launch_bounds__(1024) void k_objgrad_finish(const double* r, const double* c,
                                                         const double* a, const double* b,
                                                         const double* gamma, int n, int m,
                                                         double reg, double* grad, double* out)
{
    __shared__ double sm[1024];
    double tsum = 0.0, gn2 = 0.0, da = 0.0, db = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) {
        const double g = r[i] - a[i];
        if (grad) grad[i] = g;
        tsum += r[i];
        gn2 += g * g;
        da += gamma[i] * a[i];
    }
    for (int j = threadIdx.x; j < m - 1; j += 1024) {
        if (c) {
            const double g = c[j] - b[j];
            if (grad) grad[n + j] = g;
            gn2 += g * g;
        }
        db += gamma[n + j] * b[j];
    }
    tsum = block_reduce(tsum, sm);
    gn2 = block_reduce(gn2, sm);
    da = block_reduce(da, sm);
    db = block_reduce(db, sm);
    if (threadIdx.x == 0) {
        out[0] = reg * tsum - da - db;
        out[1] = gn2;
    }
}

__global__ void k_normalise(const double* alpha, const double* beta, int n, int m, double* gamma)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    const double bm = beta[m - 1];
    if (t < n) gamma[t] = alpha[t] + bm;
    if (t < m - 1) gamma[n + t] = beta[t] - bm;
}

// L-BFGS two-loop product res = H (a v) with the inverse-Hessian operator of
// the last ncorr pairs (s_j, y_j) in a ring of mm slots (slot j at S + j N),
// H0 = I / theta.  Order of operations as BFGSMat::apply_Hv,
// src/LBFGSpp/LBFGSpp/BFGSMat.h:276-302.  One 1024-thread workgroup: every
// thread owns the indices i = t mod 1024, so only the dot products need the
// workgroup (fixed-order tree, reproducible).
struct LbfgsArgs {
    const double* S;
    const double* Y;
    long N;
    int mm, ncorr, ptr;
    double theta;
    double ys[8];
};

__global__ __launch_bounds__(1024) void k_lbfgs_apply(LbfgsArgs L, const double* __restrict__ v,
                                                      double a, double* __restrict__ res)
{
    __shared__ double sm[1024];
    __shared__ double al[8];
    const long N = L.N;
    for (long i = threadIdx.x; i < N; i += 1024) res[i] = a * v[i];
    int j = L.ptr % L.mm;
    for (int k = 0; k < L.ncorr; k++) {
        j = (j + L.mm - 1) % L.mm;
        const double* sj = L.S + static_cast<long>(j) * N;
        const double* yj = L.Y + static_cast<long>(j) * N;
        double d = 0.0;
        for (long i = threadIdx.x; i < N; i += 1024) d += sj[i] * res[i];
        const double aj = block_reduce(d, sm) / L.ys[j];
        if (threadIdx.x == 0) al[j] = aj;
        for (long i = threadIdx.x; i < N; i += 1024) res[i] -= aj * yj[i];
    }
    for (long i = threadIdx.x; i < N; i += 1024) res[i] /= L.theta;
    __syncthreads();
    for (int k = 0; k < L.ncorr; k++) {
        const double* sj = L.S + static_cast<long>(j) * N;
        const double* yj = L.Y + static_cast<long>(j) * N;
        double d = 0.0;
        for (long i = threadIdx.x; i < N; i += 1024) d += yj[i] * res[i];
        const double bet = block_reduce(d, sm) / L.ys[j];
        const double c = al[j] - bet;
        for (long i = threadIdx.x; i < N; i += 1024) res[i] += c * sj[i];
        j = (j + 1) % L.mm;
    }
}

int grid_for(long count)
{
    long g = (count + 255) / 256;
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    return static_cast<int>(g);
}

}  // namespace

void vec_axpy(double* y, double alpha, const double* x, long count, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_axpy, dim3(grid_for(count)), dim3(256), 0, s, y, alpha, x, count);
    KERNEL_CHECK();
}

void vec_axpby(double* out, double alpha, const double* x, double beta, const double* y,
               long count, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(count)), dim3(256), 0, s, out, alpha, x, beta, y,
                       count);
    KERNEL_CHECK();
}

void vec_normalise(const double* alpha, const double* beta, int n, int m, double* gamma,
                   hipStream_t s)
{
    const int c = n > m ? n : m;
    hipLaunchKernelGGL(k_normalise, dim3((c + 255) / 256), dim3(256), 0, s, alpha, beta, n, m,
                       gamma);
    KERNEL_CHECK();
}

void vec_copy(double* dst, const double* src, long count, hipStream_t s)
{
    if (count <= 0) return;
    HIP_CHECK(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToDevice, s));
}

void vec_dots(const double* const* xs, const double* const* ys, long count, int npairs,
              double* out, hipStream_t s)
{
    if (npairs < 1 || npairs > 8) fail(E_INTERNAL, "vec_dots: %d pairs", npairs);
    DotArgs d{};
    for (int p = 0; p < npairs; p++) {
        d.x[p] = xs[p];
        d.y[p] = ys[p];
    }
    hipLaunchKernelGGL(k_dots, dim3(1), dim3(1024), 0, s, d, count, npairs, out);
    KERNEL_CHECK();
}

void vec_objgrad_finish(const double* r, const double* c, const double* a, const double* b,
                        const double* gamma, int n, int m, double reg, double* grad,
                        double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_objgrad_finish, dim3(1), dim3(1024), 0, s, r, c, a, b, gamma, n, m, reg,
                       grad, out);
    KERNEL_CHECK();
}

void vec_obj_finish(const double* r, const double* a, const double* b, const double* gamma,
                    int n, int m, double reg, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_objgrad_finish, dim3(1), dim3(1024), 0, s, r, nullptr, a, b, gamma, n,
                       m, reg, nullptr, out);
    KERNEL_CHECK();
}

void vec_lbfgs_apply(const double* S, const double* Y, const double* ys, int mm, int ncorr,
                     int ptr, double theta, const double* v, double a, double* res, long count,
                     hipStream_t s)
{
    if (mm < 1 || mm > 8 || ncorr < 0 || ncorr > mm) fail(E_INTERNAL, "L-BFGS memory %d/%d", ncorr, mm);
    LbfgsArgs L{};
    L.S = S;
    L.Y = Y;
    L.N = count;
    L.mm = mm;
    L.ncorr = ncorr;
    L.ptr = ptr;
    L.theta = theta;
    for (int q = 0; q < mm; q++) L.ys[q] = ys[q];
    hipLaunchKernelGGL(k_lbfgs_apply, dim3(1), dim3(1024), 0, s, L, v, a, res);
    KERNEL_CHECK();
}

}  // namespace regot
