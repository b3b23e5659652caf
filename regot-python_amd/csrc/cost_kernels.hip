// cost_kernels.hip -- the squared-Euclidean cost M_ij = |x_i|^2 + |y_j|^2 -
// 2 x_i . y_j formed from BF16 point clouds on the matrix cores.
//
// X (n x d) and Y (m x d) are row-major BF16.  The contraction X Y' runs on
// v_mfma_f32_32x32x16_bf16 with FP32 accumulation (exact BF16 products), one
// 32 x 32 output tile per wave, four waves per 64 x 64 workgroup tile.  The
// squared norms are summed in FP64 from the same BF16 values and the three
// terms combine in FP64 before the FP32 store.  At d = 64 the kernel moves
// 4 n m bytes (the M store) against 2 n m d flops: it is bound by the store,
// not by the matrix cores.
//
// MFMA 32x32x16 lane layout (checked on gfx950): lane l supplies A row
// l % 32, k = 8 (l / 32) .. +7 and B column l % 32, the same k; accumulator
// register r of lane l holds C(8 (r / 4) + 4 (l / 32) + r % 4, l % 32).
#include <algorithm>

#include "common.hpp"
#include "regot_hip.h"

namespace regot {

int set_error(int code, const char* what);

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ double bf16_to_double(unsigned short h)
{
    return static_cast<double>(__uint_as_float(static_cast<unsigned>(h) << 16));
}

// out[i] = sum_k X[i][k]^2 in FP64 (one thread per row).
__global__ void k_sqnorm_bf16(const unsigned short* __restrict__ X, long ldx, int n, int d,
                              double* __restrict__ out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const unsigned short* r = X + static_cast<long>(i) * ldx;
    double s = 0.0;
    for (int k = 0; k < d; k++) {
        const double v = bf16_to_double(r[k]);
        s += v * v;
    }
    out[i] = s;
}

// 8 BF16 values of row `row` from column kk (zero past d); vec: 16-byte aligned rows.
__device__ __forceinline__ bf16x8 load8(const unsigned short* __restrict__ P, long ld, int row,
                                        int kk, int d, bool vec)
{
    const unsigned short* p = P + static_cast<long>(row) * ld + kk;
    u16x8 u;
    if (vec && kk + 8 <= d) {
        u = *reinterpret_cast<const u16x8*>(p);
    } else {
#pragma unroll
        for (int q = 0; q < 8; q++) u[q] = (kk + q < d) ? p[q] : static_cast<unsigned short>(0);
    }
    return __builtin_bit_cast(bf16x8, u);
}

__global__ __launch_bounds__(256) void k_sqeuclid_bf16(
    const unsigned short* __restrict__ X, long ldx, const unsigned short* __restrict__ Y,
    long ldy, int n, int m, int d, const double* __restrict__ xx, const double* __restrict__ yy,
    float* __restrict__ M, long ldm, int vec)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int i0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const int j0 = blockIdx.x * 64 + (wave & 1) * 32;
    // rows / columns past the edge load a valid row and are not stored
    const int ia = min(i0 + (lane & 31), n - 1);
    const int jb = min(j0 + (lane & 31), m - 1);
    const int kq = 8 * (lane >> 5);
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.0f;
    for (int k0 = 0; k0 < d; k0 += 16) {
        const bf16x8 a = load8(X, ldx, ia, k0 + kq, d, vec != 0);
        const bf16x8 b = load8(Y, ldy, jb, k0 + kq, d, vec != 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    const int j = j0 + (lane & 31);
    if (j >= m) return;
    const double yj = yy[j];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int i = i0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (i < n)
            M[static_cast<long>(i) * ldm + j] =
                static_cast<float>((xx[i] + yj) - 2.0 * static_cast<double>(acc[r]));
    }
}

}  // namespace
}  // namespace regot

using namespace regot;

extern "C" {

int regot_sqeuclid_bf16(const void* X, long long ldx, const void* Y, long long ldy, int n, int m,
                        int d, float* M, long long ldm, int device)
{
    try {
        if (!X || !Y || !M) fail(E_INVALID, "null pointer");
        if (n < 1 || m < 1 || d < 1 || ldx < d || ldy < d || ldm < m)
            fail(E_INVALID, "bad sizes: n %d, m %d, d %d, ldx %lld, ldy %lld, ldm %lld", n, m, d,
                 ldx, ldy, ldm);
        HIP_CHECK(hipSetDevice(device));
        hipStream_t s = nullptr;
        HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        try {
            DBuf<double> xx(n), yy(m);
            const auto* Xp = static_cast<const unsigned short*>(X);
            const auto* Yp = static_cast<const unsigned short*>(Y);
            hipLaunchKernelGGL(k_sqnorm_bf16, dim3(ceil_div(n, 256)), dim3(256), 0, s, Xp,
                               static_cast<long>(ldx), n, d, xx.p);
            hipLaunchKernelGGL(k_sqnorm_bf16, dim3(ceil_div(m, 256)), dim3(256), 0, s, Yp,
                               static_cast<long>(ldy), m, d, yy.p);
            KERNEL_CHECK();
            const bool vec = (ldx % 8 == 0) && (ldy % 8 == 0) &&
                             (reinterpret_cast<uintptr_t>(X) % 16 == 0) &&
                             (reinterpret_cast<uintptr_t>(Y) % 16 == 0);
            hipLaunchKernelGGL(k_sqeuclid_bf16, dim3(ceil_div(m, 64), ceil_div(n, 64)), dim3(256),
                               0, s, Xp, static_cast<long>(ldx), Yp, static_cast<long>(ldy), n, m,
                               d, xx.p, yy.p, M, static_cast<long>(ldm), vec ? 1 : 0);
            KERNEL_CHECK();
            HIP_CHECK(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
            throw;
        }
        HIP_CHECK(hipStreamDestroy(s));
    } catch (const regot::Error& e) {
        return set_error(e.code, e.what());
    } catch (const std::exception& e) {
        return set_error(E_INTERNAL, e.what());
    }
    return OK;
}

}  // extern "C"
