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

// out[p] = x_p . y_p, one 1024-thread workgroup per pair: thread t sums its
// elements t, t + 1024, ... in order (sixteen loads of each vector in
// flight), then the 1024 partials go through a halving tree in LDS.
__global__ __launch_bounds__(1024) void k_dots(DotArgs d, long count, double* out)
{
    __shared__ double sm[1024];
    const int t = threadIdx.x, p = blockIdx.x;
    const double* __restrict__ x = d.x[p];
    const double* __restrict__ y = d.y[p];
    double s = 0.0;
    for (long i0 = t; i0 < count; i0 += 16 * 1024) {
        double xv[16], yv[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const long i = i0 + 1024L * u;
            xv[u] = (i < count) ? x[i] : 0.0;
            yv[u] = (i < count) ? y[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 16; u++)
            if (i0 + 1024L * u < count) s += xv[u] * yv[u];
    }
    sm[t] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (t < o) sm[t] += sm[t + o];
        __syncthreads();
    }
    if (t == 0) out[p] = sm[0];
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

// Many dot products over one index range in one pass: workgroup g sums its
// contiguous chunk for every pair (all loads of a thread independent, fixed
// order), then one workgroup folds the chunks.  Used where a single-workgroup reduction chain would be latency
// bound (the L-BFGS update reads up to 31 dot products at once).
struct MDotArgs {
    const double* x[32];
    const double* y[32];
};

__global__ __launch_bounds__(256) void k_mdots(MDotArgs d, int npairs, long count, long chunk,
                                               double* __restrict__ part)
{
    __shared__ double sm[4][32];
    const long i0 = static_cast<long>(blockIdx.x) * chunk;
    const long i1 = (i0 + chunk < count) ? i0 + chunk : count;
    double v[32];
#pragma unroll
    for (int q = 0; q < 32; q++) v[q] = 0.0;
    for (long i = i0 + threadIdx.x; i < i1; i += 256) {
#pragma unroll
        for (int q = 0; q < 32; q++)
            if (q < npairs) v[q] += d.x[q][i] * d.y[q][i];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 32; q++) {
        if (q < npairs) {
            double t = v[q];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
            if (lane == 0) sm[wave][q] = t;
        }
    }
    __syncthreads();
    if (threadIdx.x < npairs) {
        const int q = threadIdx.x;
        part[static_cast<long>(blockIdx.x) * 32 + q] = ((sm[0][q] + sm[1][q]) + sm[2][q]) + sm[3][q];
    }
}

// One workgroup of 256: 8 threads per pair sum strided chunks, then the 8
// partials in order.
__global__ __launch_bounds__(256) void k_mdots_fold(const double* __restrict__ part, int nchunks,
                                                    int npairs, double* __restrict__ out)
{
    __shared__ double sm[256];
    const int q = threadIdx.x >> 3, sub = threadIdx.x & 7;
    double v = 0.0;
    if (q < npairs)
        for (int g = sub; g < nchunks; g += 8) v += part[static_cast<long>(g) * 32 + q];
    sm[threadIdx.x] = v;
    __syncthreads();
    if (sub == 0 && q < npairs) {
        double t = 0.0;
        for (int u = 0; u < 8; u++) t += sm[threadIdx.x + u];
        out[q] = t;
    }
}

// out = sum_k c_k v_k over up to 16 vectors (fixed order per element).
struct LinCombArgs {
    const double* v[16];
    double c[16];
};

__global__ void k_lincomb(LinCombArgs L, int nv, long count, double* __restrict__ out)
{
    for (long i = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; i < count;
         i += static_cast<long>(gridDim.x) * 256) {
        double acc = 0.0;
        for (int k = 0; k < nv; k++) acc += L.c[k] * L.v[k][i];
        out[i] = acc;
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
    hipLaunchKernelGGL(k_dots, dim3(npairs), dim3(1024), 0, s, d, count, out);
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

// Chunks of at least 256 entries (one per thread), at most 512 workgroups.
static int mdot_chunks(long count, long* chunk)
{
    long g = (count + 255) / 256;
    g = g < 1 ? 1 : (g > 512 ? 512 : g);
    *chunk = (count + g - 1) / g;
    return static_cast<int>(g);
}

void vec_mdots(const double* const* xs, const double* const* ys, int npairs, long count,
               double* part, double* out, hipStream_t s)
{
    if (npairs < 1 || npairs > 32) fail(E_INTERNAL, "vec_mdots: %d pairs", npairs);
    MDotArgs d{};
    for (int q = 0; q < npairs; q++) {
        d.x[q] = xs[q];
        d.y[q] = ys[q];
    }
    long chunk = 0;
    const int g = mdot_chunks(count, &chunk);
    hipLaunchKernelGGL(k_mdots, dim3(g), dim3(256), 0, s, d, npairs, count, chunk, part);
    hipLaunchKernelGGL(k_mdots_fold, dim3(1), dim3(256), 0, s, part, g, npairs, out);
    KERNEL_CHECK();
}

long vec_mdots_scratch(long count)
{
    long chunk = 0;
    return 32L * mdot_chunks(count, &chunk);
}

void vec_lincomb(const double* const* vs, const double* coef, int nv, long count, double* out,
                 hipStream_t s)
{
    if (nv < 1 || nv > 16) fail(E_INTERNAL, "vec_lincomb: %d vectors", nv);
    LinCombArgs L{};
    for (int k = 0; k < nv; k++) {
        L.v[k] = vs[k];
        L.c[k] = coef[k];
    }
    hipLaunchKernelGGL(k_lincomb, dim3(grid_for(count)), dim3(256), 0, s, L, nv, count, out);
    KERNEL_CHECK();
}

}  // namespace regot
