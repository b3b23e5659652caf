// noalloc.hip -- the dual objective and gradient on caller-provided device
// buffers without any device allocation, for callers that cannot allocate
// (code under HIP graph capture, tight serving loops).  The kernels are those
// of regot_dual_obj_grad (the FP32 row and column passes, or the one-pass FP64
// kernel for FP64 bulk); their scratch -- split potentials, log marginals,
// row / column partial sums -- lives in the file-scope device array
// `scratch`, one copy per GPU, held until the process ends; f, ||g||^2 and the
// gradient go straight to the caller's buffers.  Every call on a GPU shares
// that GPU's copy, so calls that overlap on one GPU race.  Every scratch word
// a call reads is written earlier in the same call, so nothing needs clearing
// between calls.
#include <cstdint>
#include <mutex>

#include "dense.hpp"
#include "problem.hpp"
#include "regot_hip.h"

namespace regot {

int set_error(int code, const char* what);

namespace {

// Maximum scratch: 32 MiB per GPU (n = m = 16384 in FP64 bulk uses ~16 MiB).
constexpr long kScratchDoubles = 4L << 20;

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (the array's name in
// host code is a host-side stand-in, never a device address).
double* scratch_on(int device)
{
    if (device < 0 || device >= kMaxDevices) fail(E_INVALID, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(E_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                 hipGetErrorString(e));
        g_ptr[device] = static_cast<double*>(p);
    }
    return g_ptr[device];
}

__global__ void k_log_nalloc(const double* x, double* y, int count)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) y[i] = log(x[i]);
}

}  // namespace
}  // namespace regot

using namespace regot;

extern "C" int regot_dual_obj_grad_noalloc(const void* M, int bulk, long long ldm, int n, int m,
                                           const double* d_a, const double* d_b, double reg,
                                           const double* d_gamma, double* d_out, double* d_grad,
                                           void* stream, int device)
{
    try {
        if (!M || !d_a || !d_b || !d_gamma || !d_out || !d_grad)
            fail(E_INVALID, "null pointer");
        if (bulk != REGOT_F32 && bulk != REGOT_F64) fail(E_INVALID, "bulk must be F32 or F64");
        if (n < 1 || m < 2) fail(E_INVALID, "need n >= 1 and m >= 2 (got %d, %d)", n, m);
        if (!(reg > 0.0)) fail(E_INVALID, "reg must be positive");
        const long esz = bulk == REGOT_F32 ? 4 : 8;
        if (ldm == 0) ldm = m;
        if (ldm < m) fail(E_INVALID, "ldm (%lld) < m (%d)", ldm, m);
        if (reinterpret_cast<std::uintptr_t>(M) % 16 != 0 || (ldm * esz) % 16 != 0)
            fail(E_INVALID, "M and its row pitch must be 16-byte aligned");
        // scratch layout (doubles): alpha n | beta m | loga n | logb m | then
        // the kernels' partials
        const long base = 2L * n + 2L * m;
        long nr = 0, nc = 0, np = 0, need = 0;
        int nblk = 0;
        if (bulk == REGOT_F64) {
            sums_scratch_sizes(n, m, &nr, &nc, &np);
            need = base + nr + nc + np;
        } else {
            nblk = col_pass_blocks(n, m);
            need = base + 2L * n + 2L * m + static_cast<long>(nblk) * m;
        }
        if (need > kScratchDoubles)
            fail(E_INVALID,
                 "n = %d, m = %d needs %ld bytes of scratch; the no-allocation entry point "
                 "holds at most %ld",
                 n, m, need * 8, kScratchDoubles * 8);

        int prev = 0;
        HIP_CHECK(hipGetDevice(&prev));
        HIP_CHECK(hipSetDevice(device));
        struct Restore {
            int d;
            ~Restore() { (void)hipSetDevice(d); }
        } restore{prev};
        double* sc = scratch_on(device);
        hipStream_t s = static_cast<hipStream_t>(stream);
        double* alpha = sc;
        double* beta = alpha + n;
        double* loga = beta + m;
        double* logb = loga + n;
        double* rest = logb + m;
        launch_split_gamma(d_gamma, n, m, alpha, beta, s);
        if (bulk == REGOT_F64) {
            launch_objgrad_fused(bulk, M, static_cast<long>(ldm), n, m, alpha, beta, d_a, d_b,
                                 reg, true, d_grad, rest, rest + nr, rest + nr + nc, d_out, s);
        } else {
            // the FP32 row and column passes of regot_dual_obj_grad
            double* rsum = rest;
            double* csum = rsum + n;
            double* part = csum + m;
            hipLaunchKernelGGL(k_log_nalloc, dim3(ceil_div(n, 256)), dim3(256), 0, s, d_a, loga,
                               n);
            hipLaunchKernelGGL(k_log_nalloc, dim3(ceil_div(m, 256)), dim3(256), 0, s, d_b, logb,
                               m);
            KERNEL_CHECK();
            const double k = 1.4426950408889634074 / reg;
            RowEpi e;
            e.loga = loga;
            e.rowpot = alpha;
            e.rowsum = rsum;
            e.reg = reg;
            launch_row_pass(bulk, M, static_cast<long>(ldm), n, m, beta, k, e, nullptr, s);
            ColEpi c;
            c.colpot = beta;
            c.colsum = csum;
            c.reg = reg;
            c.logb = logb;
            c.b = d_b;
            launch_col_pass(bulk, M, static_cast<long>(ldm), n, m, alpha, beta, logb, k, part,
                            nblk, c, nullptr, nullptr, s);
            vec_objgrad_finish(rsum, csum, d_a, d_b, d_gamma, n, m, reg, d_grad, d_out, s);
        }
        return OK;
    } catch (const regot::Error& e) {
        return set_error(e.code, e.what());
    } catch (const std::exception& e) {
        return set_error(E_INTERNAL, e.what());
    }
}
