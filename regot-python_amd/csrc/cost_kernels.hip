// cost_kernels.hip -- the squared-Euclidean cost M_ij = |x_i|^2 + |y_j|^2 -
// 2 x_i . y_j formed from BF16 point clouds on the matrix cores.
//
// X (n x d) and Y (m x d) are row-major BF16.  The contraction X Y' runs on
// v_mfma_f32_32x32x16_bf16 with FP32 accumulation (exact BF16 products), one
// 32 x 32 output tile per wave, four waves per 64 x 64 workgroup tile.  The
// squared norms come from the same BF16 values the lanes load for the
// contraction (FP64 squares; each half-wave holds half of every 16-wide k
// step, so the two halves' sums are added), and the three terms combine in
// FP64 before the FP32 store.  For d <= 64 every k step's operands are loaded
// before the first MFMA (the loads are in flight together).  At d = 64 the
// kernel moves 4 n m bytes (the M store) against 2 n m d flops: it is bound
// by the store, not by the matrix cores.
//
// MFMA 32x32x16 lane layout (checked on gfx950): lane l supplies A row
// l % 32, k = 8 (l / 32) .. +7 and B column l % 32, the same k; accumulator
// register r of lane l holds C(8 (r / 4) + 4 (l / 32) + r % 4, l % 32).
#include <algorithm>
#include <cstdint>
#include <mutex>

#include "common.hpp"
#include "regot_hip.h"

namespace regot {

int set_error(int code, const char* what);

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// sum of squares of 8 BF16 values in FP64 (BF16 -> FP32 is exact)
__device__ __forceinline__ double sq8(bf16x8 v)
{
    const u16x8 u = __builtin_bit_cast(u16x8, v);
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const double x = static_cast<double>(__uint_as_float(static_cast<unsigned>(u[q]) << 16));
        s = fma(x, x, s);
    }
    return s;
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

// KS > 0: d <= 16 KS, every k step's operands in registers at once; KS = 0:
// any d, one k step at a time.
template <int KS>
__global__ __launch_bounds__(256) void k_sqeuclid_bf16(
    const unsigned short* __restrict__ X, long ldx, const unsigned short* __restrict__ Y,
    long ldy, int n, int m, int d, float* __restrict__ M, long ldm, int vec)
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
    double na = 0.0, nb = 0.0;
    if (KS > 0) {
        bf16x8 av[KS > 0 ? KS : 1], bv[KS > 0 ? KS : 1];
#pragma unroll
        for (int s = 0; s < KS; s++) {
            av[s] = load8(X, ldx, ia, 16 * s + kq, d, vec != 0);
            bv[s] = load8(Y, ldy, jb, 16 * s + kq, d, vec != 0);
        }
#pragma unroll
        for (int s = 0; s < KS; s++) {
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[s], bv[s], acc, 0, 0, 0);
            na += sq8(av[s]);
            nb += sq8(bv[s]);
        }
    } else {
        for (int k0 = 0; k0 < d; k0 += 16) {
            const bf16x8 a = load8(X, ldx, ia, k0 + kq, d, vec != 0);
            const bf16x8 b = load8(Y, ldy, jb, k0 + kq, d, vec != 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
            na += sq8(a);
            nb += sq8(b);
        }
    }
    // |x_ia|^2 and |y_jb|^2: the other half-wave holds the other half of every
    // k step (the same sum either way round)
    na += __shfl_xor(na, 32);
    nb += __shfl_xor(nb, 32);
    // the workgroup's 64 x 64 tile through LDS (padded rows), then written
    // as 256-byte row segments of float4 stores
    __shared__ float tile[64][68];
    const int tr0 = (wave >> 1) * 32, tc = (wave & 1) * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int ri = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const double xi = __shfl(na, ri);   // lane ri loaded row i0 + ri
        tile[tr0 + ri][tc] = static_cast<float>((xi + nb) - 2.0 * static_cast<double>(acc[r]));
    }
    __syncthreads();
    const int bi = blockIdx.y * 64, bj = blockIdx.x * 64;
    const int cq = 4 * (threadIdx.x & 15);   // column quad within the tile
    const bool full = bj + 64 <= m && (ldm % 4 == 0) &&
                      (reinterpret_cast<uintptr_t>(M) % 16 == 0);   // workgroup-uniform
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int tr = q * 16 + (threadIdx.x >> 4);
        const int i = bi + tr;
        if (i >= n) continue;
        float* dst = M + static_cast<long>(i) * ldm + bj + cq;
        if (full) {
            *reinterpret_cast<float4*>(dst) =
                make_float4(tile[tr][cq], tile[tr][cq + 1], tile[tr][cq + 2], tile[tr][cq + 3]);
        } else {
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (bj + cq + u < m) dst[u] = tile[tr][cq + u];
        }
    }
}

// One non-blocking stream per device for the cost kernel, kept for the process.
hipStream_t cost_stream(int device)
{
    static std::mutex mu;
    static hipStream_t st[64] = {};
    if (device < 0 || device >= 64) fail(E_INVALID, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(mu);
    if (!st[device]) HIP_CHECK(hipStreamCreateWithFlags(&st[device], hipStreamNonBlocking));
    return st[device];
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
        hipStream_t s = cost_stream(device);
        const auto* Xp = static_cast<const unsigned short*>(X);
        const auto* Yp = static_cast<const unsigned short*>(Y);
        const bool vec = (ldx % 8 == 0) && (ldy % 8 == 0) &&
                         (reinterpret_cast<uintptr_t>(X) % 16 == 0) &&
                         (reinterpret_cast<uintptr_t>(Y) % 16 == 0);
        const dim3 grid(ceil_div(m, 64), ceil_div(n, 64)), blk(256);
        const long lx = ldx, ly = ldy, lm = ldm;
        const int vf = vec ? 1 : 0;
        switch (ceil_div(d, 16)) {
        case 1:
            hipLaunchKernelGGL(k_sqeuclid_bf16<1>, grid, blk, 0, s, Xp, lx, Yp, ly, n, m, d, M, lm,
                               vf);
            break;
        case 2:
            hipLaunchKernelGGL(k_sqeuclid_bf16<2>, grid, blk, 0, s, Xp, lx, Yp, ly, n, m, d, M, lm,
                               vf);
            break;
        case 3:
            hipLaunchKernelGGL(k_sqeuclid_bf16<3>, grid, blk, 0, s, Xp, lx, Yp, ly, n, m, d, M, lm,
                               vf);
            break;
        case 4:
            hipLaunchKernelGGL(k_sqeuclid_bf16<4>, grid, blk, 0, s, Xp, lx, Yp, ly, n, m, d, M, lm,
                               vf);
            break;
        default:
            hipLaunchKernelGGL(k_sqeuclid_bf16<0>, grid, blk, 0, s, Xp, lx, Yp, ly, n, m, d, M, lm,
                               vf);
            break;
        }
        KERNEL_CHECK();
        HIP_CHECK(hipStreamSynchronize(s));
    } catch (const regot::Error& e) {
        return set_error(e.code, e.what());
    } catch (const std::exception& e) {
        return set_error(E_INTERNAL, e.what());
    }
    return OK;
}

}  // extern "C"
