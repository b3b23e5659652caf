// problem.hip -- the EROT problem handle: M resident in HBM, the two streaming
// passes, and the operator-level C-ABI entry points.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "problem.hpp"
#include "qrot.hpp"
#include "regot_hip.h"
#include "sparse.hpp"

namespace regot {

thread_local std::string g_last_error;

namespace {

template <typename Tin, typename Tout>
__global__ void k_convert_pitch(const Tin* __restrict__ src, long lds, Tout* __restrict__ dst,
                                long ldd, int n, int m)
{
    const long total = static_cast<long>(n) * ldd;
    for (long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; t < total;
         t += static_cast<long>(gridDim.x) * 256) {
        const long i = t / ldd, j = t % ldd;
        dst[t] = (j < m) ? static_cast<Tout>(src[i * lds + j]) : Tout(0);
    }
}

__global__ void k_log(const double* x, double* y, int count)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < count) y[i] = log(x[i]);
}

}  // namespace
}  // namespace regot

using namespace regot;

regot_problem::~regot_problem()
{
    hess.reset();
    bcd.reset();
    qrot.reset();
    for (auto& slot : prof_slot) {
        for (auto& e : slot.pending) {
            (void)hipEventDestroy(e.start);
            (void)hipEventDestroy(e.stop);
        }
        for (auto& e : slot.pool) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    }
    if (own_M && dM) (void)hipFree(dM);
    if (dMT) (void)hipFree(dMT);
    if (stream) (void)hipStreamDestroy(stream);
}

void regot_problem::setup(int n_, int m_, const void* M, int M_dtype, int bulk_, int M_device,
                          long long ldm_in, const double* a_h, const double* b_h, double reg_,
                          int device_)
{
    if (n_ < 1 || m_ < 1) fail(E_INVALID, "n and m must be >= 1 (got %d, %d)", n_, m_);
    if (!M || !a_h || !b_h) fail(E_INVALID, "M, a and b must be non-null");
    if ((M_dtype != REGOT_F32 && M_dtype != REGOT_F64) ||
        (bulk_ != REGOT_F32 && bulk_ != REGOT_F64))
        fail(E_INVALID, "dtype codes must be REGOT_F32 or REGOT_F64");
    if (!(reg_ > 0.0)) fail(E_INVALID, "reg must be positive (got %g)", reg_);
    const long lds = ldm_in > 0 ? static_cast<long>(ldm_in) : m_;
    if (lds < m_) fail(E_INVALID, "ldm (%ld) < m (%d)", lds, m_);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        fail(E_HIP, "no GPU visible to the HIP runtime");
    if (device_ < 0 || device_ >= ndev) fail(E_INVALID, "device %d out of range", device_);
    device = device_;
    HIP_CHECK(hipSetDevice(device));
    HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    n = n_;
    m = m_;
    bulk = bulk_;
    reg = reg_;
    k = 1.4426950408889634074 / reg;
    const size_t esz = bulk == REGOT_F32 ? 4 : 8;
    const size_t in_esz = M_dtype == REGOT_F32 ? 4 : 8;

    const bool aligned = (reinterpret_cast<uintptr_t>(M) % 16 == 0) && ((lds * esz) % 16 == 0);
    if (M_device && M_dtype == bulk && aligned) {
        dM = const_cast<void*>(M);
        ldm = lds;
        own_M = false;
    } else {
        const long align_el = 256 / static_cast<long>(esz);
        ldm = (m + align_el - 1) / align_el * align_el;
        HIP_CHECK(hipMalloc(&dM, static_cast<size_t>(n) * ldm * esz));
        own_M = true;
        if (M_device) {
            const int g = 4096;
            if (M_dtype == REGOT_F32 && bulk == REGOT_F32)
                hipLaunchKernelGGL((k_convert_pitch<float, float>), dim3(g), dim3(256), 0, stream,
                                   static_cast<const float*>(M), lds, static_cast<float*>(dM),
                                   ldm, n, m);
            else if (M_dtype == REGOT_F32)
                hipLaunchKernelGGL((k_convert_pitch<float, double>), dim3(g), dim3(256), 0,
                                   stream, static_cast<const float*>(M), lds,
                                   static_cast<double*>(dM), ldm, n, m);
            else if (bulk == REGOT_F32)
                hipLaunchKernelGGL((k_convert_pitch<double, float>), dim3(g), dim3(256), 0,
                                   stream, static_cast<const double*>(M), lds,
                                   static_cast<float*>(dM), ldm, n, m);
            else
                hipLaunchKernelGGL((k_convert_pitch<double, double>), dim3(g), dim3(256), 0,
                                   stream, static_cast<const double*>(M), lds,
                                   static_cast<double*>(dM), ldm, n, m);
            KERNEL_CHECK();
        } else {
            // Host source: convert/pad row chunks in pinned staging, then copy.
            const long rows_per = std::max<long>(1, (32L << 20) / (ldm * esz));
            HBuf<unsigned char> stage;
            stage.ensure(static_cast<size_t>(rows_per) * ldm * esz);
            for (long r0 = 0; r0 < n; r0 += rows_per) {
                const long nr = std::min<long>(rows_per, n - r0);
                HIP_CHECK(hipStreamSynchronize(stream));
                for (long r = 0; r < nr; r++) {
                    const unsigned char* src =
                        static_cast<const unsigned char*>(M) + (r0 + r) * lds * in_esz;
                    unsigned char* dst = stage.p + r * ldm * esz;
                    if (M_dtype == bulk) {
                        std::memcpy(dst, src, m * esz);
                    } else if (bulk == REGOT_F32) {
                        const double* s = reinterpret_cast<const double*>(src);
                        float* d = reinterpret_cast<float*>(dst);
                        for (int j = 0; j < m; j++) d[j] = static_cast<float>(s[j]);
                    } else {
                        const float* s = reinterpret_cast<const float*>(src);
                        double* d = reinterpret_cast<double*>(dst);
                        for (int j = 0; j < m; j++) d[j] = static_cast<double>(s[j]);
                    }
                    std::memset(dst + m * esz, 0, (ldm - m) * esz);
                }
                HIP_CHECK(hipMemcpyAsync(static_cast<unsigned char*>(dM) + r0 * ldm * esz,
                                         stage.p, nr * ldm * esz, hipMemcpyHostToDevice, stream));
            }
            HIP_CHECK(hipStreamSynchronize(stream));
        }
    }

    // sum a and sum_{j<m-1} b (host FP64, fixed order) for the objective
    sum_a = 0.0;
    sum_b1 = 0.0;
    for (int i = 0; i < n; i++) sum_a += a_h[i];
    for (int j = 0; j < m - 1; j++) sum_b1 += b_h[j];
    sum_b = sum_b1 + b_h[m - 1];
    min_a = a_h[0];
    min_b = b_h[0];
    for (int i = 1; i < n; i++) min_a = std::min(min_a, a_h[i]);
    for (int j = 1; j < m; j++) min_b = std::min(min_b, b_h[j]);
    sumsq_a = 0.0;
    sumsq_b = 0.0;
    for (int i = 0; i < n; i++) sumsq_a += a_h[i] * a_h[i];
    for (int j = 0; j < m; j++) sumsq_b += b_h[j] * b_h[j];
    HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
    HIP_CHECK(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device));

    a.alloc(n);
    b.alloc(m);
    loga.alloc(n);
    logb.alloc(m);
    HIP_CHECK(hipMemcpyAsync(a, a_h, n * sizeof(double), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(b, b_h, m * sizeof(double), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_log, dim3(ceil_div(n, 256)), dim3(256), 0, stream, a.p, loga.p, n);
    hipLaunchKernelGGL(k_log, dim3(ceil_div(m, 256)), dim3(256), 0, stream, b.p, logb.p, m);
    KERNEL_CHECK();

    alpha.alloc(n);
    beta.alloc(m);
    rsum.alloc(n);
    csum.alloc(m);
    gamma.alloc(n + m);
    grad.alloc(n + m);
    scal.alloc(64);
    ctl.alloc(16);
    nblk = col_pass_blocks(n, m);
    part.alloc(static_cast<size_t>(nblk) * m);
    ntiles = col_combine_blocks(m);
    hscal.ensure(64);
    hctl.ensure(16);
    HIP_CHECK(hipStreamSynchronize(stream));
}

void regot_problem::sync() { HIP_CHECK(hipStreamSynchronize(stream)); }

void regot_problem::prof_begin(int kind, hipEvent_t* stop_out)
{
    *stop_out = nullptr;
    if (!prof) return;
    auto& slot = prof_slot[kind];
    std::pair<hipEvent_t, hipEvent_t> ev;
    if (!slot.pool.empty()) {
        ev = slot.pool.back();
        slot.pool.pop_back();
    } else {
        HIP_CHECK(hipEventCreate(&ev.first));
        HIP_CHECK(hipEventCreate(&ev.second));
    }
    HIP_CHECK(hipEventRecord(ev.first, stream));
    slot.pending.push_back({ev.first, ev.second, prof_tag});
    *stop_out = ev.second;
}

void regot_problem::prof_end(int kind, hipEvent_t stop)
{
    if (!prof || !stop) return;
    HIP_CHECK(hipEventRecord(stop, stream));
    if (prof_slot[kind].pending.size() >= 2048) prof_collect();
}

void regot_problem::prof_collect()
{
    for (auto& slot : prof_slot) {
        for (auto& e : slot.pending) {
            HIP_CHECK(hipEventSynchronize(e.stop));
            float ms = 0.f;
            HIP_CHECK(hipEventElapsedTime(&ms, e.start, e.stop));
            if (e.tag < 0) {
                slot.total_ms += ms;
                slot.launches += 1;
            } else {
                slot.tagged.emplace_back(e.tag, static_cast<double>(ms));
            }
            slot.pool.emplace_back(e.start, e.stop);
        }
        slot.pending.clear();
    }
}

void regot_problem::prof_drop_after(long last)
{
    prof_collect();
    for (auto& slot : prof_slot) {
        auto& t = slot.tagged;
        t.erase(std::remove_if(t.begin(), t.end(),
                               [last](const std::pair<long, double>& e) { return e.first > last; }),
                t.end());
    }
}

void regot_problem::prof_commit(long last)
{
    prof_collect();
    for (auto& slot : prof_slot) {
        for (const auto& t : slot.tagged) {
            if (t.first > last) continue;
            slot.total_ms += t.second;
            slot.launches += 1;
        }
        slot.tagged.clear();
    }
}

void regot_problem::row_pass(const double* beta_full, const RowEpi& epi, const int* ctl_p)
{
    hipEvent_t stop;
    prof_begin(0, &stop);
    RowEpi e = epi;
    e.reg = reg;
    launch_row_pass(bulk, dM, ldm, n, m, beta_full, k, e, ctl_p, stream, precise);
    prof_end(0, stop);
}

void regot_problem::col_pass(const double* alpha_v, const double* colpot_guess,
                             const ColEpi& epi, const int* ctl_p, const BcdState* st)
{
    hipEvent_t stop;
    prof_begin(1, &stop);
    ColEpi e = epi;
    e.reg = reg;
    e.logb = logb;
    e.b = b;
    launch_col_pass(bulk, dM, ldm, n, m, alpha_v, colpot_guess, logb, k, part, nblk, e, st,
                    ctl_p, stream, st == nullptr ? precise : bcd_precise);
    prof_end(1, stop);
}

void regot_problem::row_pass_bcd(const double* beta_full, const RowEpi& epi, const BcdState& st)
{
    hipEvent_t stop;
    prof_begin(0, &stop);
    RowEpi e = epi;
    e.reg = reg;
    launch_row_pass_bcd(bulk, dM, ldm, n, m, beta_full, k, e, st, stream, bcd_precise);
    prof_end(0, stop);
}

// The single-read pass forms column sums of the plan at (alpha_{k+1}, beta_k)
// from row-shifted FP32 terms; they keep full relative accuracy when every
// column sum is far above the 2^-106 sum a that flushing can drop, which holds
// when beta_k came out of an exact beta update and the marginals are bounded
// away from zero as below (see bcd_fused.hip).
bool regot_problem::bcd_fused_ok() const
{
    return bulk == REGOT_F32 && bcd_fused_shape_ok(n, m, ncu) && min_a > 0.0 && min_b > 0.0 &&
           min_a * min_b >= 0x1p-60 * sum_a * sum_b;
}

void regot_problem::bcd_fused_iteration(const double* alpha_k, double* alpha_out, double* beta_v,
                                        const ColEpi& epi, const BcdState& st, int nblk_f,
                                        double* fpart)
{
    hipEvent_t stop;
    prof_begin(2, &stop);
    launch_bcd_fused(dM, ldm, n, m, beta_v, k, reg, loga, a, alpha_k, alpha_out, nblk_f, fpart,
                     st.rpart, st.bm, st.ctl, stream);
    prof_end(2, stop);
    prof_begin(3, &stop);
    ColEpi e = epi;
    e.reg = reg;
    e.logb = logb;
    e.b = b;
    e.colpot = beta_v;
    launch_col_combine_sums(fpart, nblk_f, m, k, e, st, stream);
    prof_end(3, stop);
}

// One-pass FP64 objective (and gradient): the precise mode of FP32 bulk and
// every FP64-bulk evaluation.  Returns f; *gnorm2 = ||g||^2 when with_grad.
double regot_problem::dev_obj_fused(const double* d_gamma, double* d_grad, double* gnorm2)
{
    long nr, nc, np;
    sums_scratch_sizes(n, m, &nr, &nc, &np);
    s_row.ensure(nr);
    s_col.ensure(nc);
    s_part.ensure(np);
    launch_split_gamma(d_gamma, n, m, alpha, beta, stream);
    const bool with_grad = d_grad != nullptr;
    launch_objgrad_fused(bulk, dM, ldm, n, m, alpha, beta, a, b, reg, with_grad, d_grad, s_row,
                         s_col, s_part, scal, stream);
    HIP_CHECK(hipMemcpyAsync(hscal.p, scal.p, 2 * sizeof(double), hipMemcpyDeviceToHost, stream));
    sync();
    if (gnorm2) *gnorm2 = hscal.p[1];
    return hscal.p[0];
}

double regot_problem::dev_obj(const double* d_gamma)
{
    if (precise || bulk == 1) return dev_obj_fused(d_gamma, nullptr, nullptr);
    launch_split_gamma(d_gamma, n, m, alpha, beta, stream);
    RowEpi e;
    e.loga = loga;
    e.rowpot = alpha;
    e.rowsum = rsum;
    row_pass(beta, e);
    vec_obj_finish(rsum, a, b, d_gamma, n, m, reg, scal, stream);
    HIP_CHECK(hipMemcpyAsync(hscal.p, scal.p, sizeof(double), hipMemcpyDeviceToHost, stream));
    sync();
    return hscal.p[0];
}

double regot_problem::dev_obj_grad(const double* d_gamma, double* d_grad, double* gnorm2)
{
    if (precise || bulk == 1) return dev_obj_fused(d_gamma, d_grad, gnorm2);
    launch_split_gamma(d_gamma, n, m, alpha, beta, stream);
    RowEpi e;
    e.loga = loga;
    e.rowpot = alpha;
    e.rowsum = rsum;
    row_pass(beta, e);
    ColEpi c;
    c.colpot = beta;
    c.colsum = csum;
    col_pass(alpha, beta, c);
    vec_objgrad_finish(rsum, csum, a, b, d_gamma, n, m, reg, d_grad, scal, stream);
    HIP_CHECK(hipMemcpyAsync(hscal.p, scal.p, 2 * sizeof(double), hipMemcpyDeviceToHost, stream));
    sync();
    if (gnorm2) *gnorm2 = hscal.p[1];
    return hscal.p[0];
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------

namespace regot {
int set_error(int code, const char* what)
{
    g_last_error = what;
    return code;
}
}  // namespace regot

#define CAPI_BEGIN try {
#define CAPI_END                                                   \
    }                                                              \
    catch (const regot::Error& e) { return set_error(e.code, e.what()); } \
    catch (const std::bad_alloc&) { return set_error(E_NOMEM, "host allocation failed"); } \
    catch (const std::exception& e) { return set_error(E_INTERNAL, e.what()); }            \
    return OK;

static void check_handle(const regot_problem* p)
{
    if (!p) fail(E_INVALID, "null problem handle");
    HIP_CHECK(hipSetDevice(p->device));
}

extern "C" {

const char* regot_last_error(void) { return g_last_error.c_str(); }

int regot_version(void) { return 100; }

int regot_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int regot_check_exp2(const double* x, double* y, long long count, int device)
{
    CAPI_BEGIN
    if (count < 0 || (count > 0 && (!x || !y))) fail(E_INVALID, "bad arguments");
    if (count == 0) return OK;
    HIP_CHECK(hipSetDevice(device));
    DBuf<double> dx(count), dy(count);
    HIP_CHECK(hipMemcpy(dx.p, x, count * sizeof(double), hipMemcpyHostToDevice));
    launch_exp2_check(dx, dy, count, nullptr);
    HIP_CHECK(hipMemcpy(y, dy.p, count * sizeof(double), hipMemcpyDeviceToHost));
    CAPI_END
}

int regot_check_log(const double* x, double* y, long long count, int device)
{
    CAPI_BEGIN
    if (count < 0 || (count > 0 && (!x || !y))) fail(E_INVALID, "bad arguments");
    if (count == 0) return OK;
    HIP_CHECK(hipSetDevice(device));
    DBuf<double> dx(count), dy(count);
    HIP_CHECK(hipMemcpy(dx.p, x, count * sizeof(double), hipMemcpyHostToDevice));
    launch_log_check(dx, dy, count, nullptr);
    HIP_CHECK(hipMemcpy(y, dy.p, count * sizeof(double), hipMemcpyDeviceToHost));
    CAPI_END
}

int regot_problem_create(regot_problem** out, int n, int m, const void* M, int M_dtype,
                         int bulk, int M_device, long long ldm, const double* a,
                         const double* b, double reg, int device)
{
    CAPI_BEGIN
    if (!out) fail(E_INVALID, "out is null");
    *out = nullptr;
    auto p = std::make_unique<regot_problem>();
    p->setup(n, m, M, M_dtype, bulk, M_device, ldm, a, b, reg, device);
    *out = p.release();
    CAPI_END
}

int regot_problem_destroy(regot_problem* p)
{
    CAPI_BEGIN
    if (p) {
        (void)hipSetDevice(p->device);
        (void)hipStreamSynchronize(p->stream);
        delete p;
    }
    CAPI_END
}

int regot_problem_dims(const regot_problem* p, int* n, int* m, int* bulk)
{
    CAPI_BEGIN
    if (!p) fail(E_INVALID, "null problem handle");
    if (n) *n = p->n;
    if (m) *m = p->m;
    if (bulk) *bulk = p->bulk;
    CAPI_END
}

int regot_optimal_alpha(regot_problem* p, const double* beta_h, double* alpha_h)
{
    CAPI_BEGIN
    check_handle(p);
    if (!beta_h || !alpha_h) fail(E_INVALID, "null vector");
    HIP_CHECK(hipMemcpyAsync(p->beta, beta_h, p->m * sizeof(double), hipMemcpyHostToDevice,
                             p->stream));
    RowEpi e;
    e.loga = p->loga;
    e.alpha_out = p->alpha;
    p->row_pass(p->beta, e);
    HIP_CHECK(hipMemcpyAsync(alpha_h, p->alpha, p->n * sizeof(double), hipMemcpyDeviceToHost,
                             p->stream));
    p->sync();
    CAPI_END
}

int regot_optimal_beta(regot_problem* p, const double* alpha_h, double* beta_h)
{
    CAPI_BEGIN
    check_handle(p);
    if (!beta_h || !alpha_h) fail(E_INVALID, "null vector");
    HIP_CHECK(hipMemcpyAsync(p->alpha, alpha_h, p->n * sizeof(double), hipMemcpyHostToDevice,
                             p->stream));
    ColEpi c;
    c.beta_out = p->beta;
    p->col_pass(p->alpha, nullptr, c);
    HIP_CHECK(hipMemcpyAsync(beta_h, p->beta, p->m * sizeof(double), hipMemcpyDeviceToHost,
                             p->stream));
    p->sync();
    CAPI_END
}

int regot_dual_obj(regot_problem* p, const double* gamma_h, double* obj)
{
    CAPI_BEGIN
    check_handle(p);
    if (!gamma_h || !obj) fail(E_INVALID, "null pointer");
    const long N = p->n + p->m - 1;
    HIP_CHECK(hipMemcpyAsync(p->gamma, gamma_h, N * sizeof(double), hipMemcpyHostToDevice,
                             p->stream));
    *obj = p->dev_obj(p->gamma);
    CAPI_END
}

int regot_dual_obj_grad(regot_problem* p, const double* gamma_h, double* obj, double* grad_h)
{
    CAPI_BEGIN
    check_handle(p);
    if (!gamma_h || !obj || !grad_h) fail(E_INVALID, "null pointer");
    const long N = p->n + p->m - 1;
    HIP_CHECK(hipMemcpyAsync(p->gamma, gamma_h, N * sizeof(double), hipMemcpyHostToDevice,
                             p->stream));
    *obj = p->dev_obj_grad(p->gamma, p->grad, nullptr);
    HIP_CHECK(
        hipMemcpyAsync(grad_h, p->grad, N * sizeof(double), hipMemcpyDeviceToHost, p->stream));
    p->sync();
    CAPI_END
}

int regot_get_plan(regot_problem* p, const double* gamma_h, double* plan)
{
    CAPI_BEGIN
    check_handle(p);
    if (!gamma_h || !plan) fail(E_INVALID, "null pointer");
    const long N = p->n + p->m - 1;
    HIP_CHECK(hipMemcpyAsync(p->gamma, gamma_h, N * sizeof(double), hipMemcpyHostToDevice,
                             p->stream));
    launch_split_gamma(p->gamma, p->n, p->m, p->alpha, p->beta, p->stream);
    const int rows = std::max(1, std::min(p->n, static_cast<int>((64L << 20) / (8L * p->m))));
    DBuf<double> dplan(static_cast<size_t>(rows) * p->m);
    for (long r0 = 0; r0 < p->n; r0 += rows) {
        const int nr = static_cast<int>(std::min<long>(rows, p->n - r0));
        launch_plan(p->bulk, p->dM, p->ldm, p->n, p->m, p->alpha, p->beta, p->reg, dplan, r0, nr,
                    p->stream);
        HIP_CHECK(hipMemcpyAsync(plan + r0 * p->m, dplan, static_cast<size_t>(nr) * p->m * 8,
                                 hipMemcpyDeviceToHost, p->stream));
        p->sync();
    }
    CAPI_END
}

int regot_profile_enable(regot_problem* p, int on)
{
    CAPI_BEGIN
    check_handle(p);
    p->prof = (on != 0);
    CAPI_END
}

int regot_profile_get(regot_problem* p, int kind, long long* launches, double* total_ms)
{
    CAPI_BEGIN
    check_handle(p);
    if (kind < 0 || kind > 3) fail(E_INVALID, "profile kind %d", kind);
    p->prof_collect();
    if (launches) *launches = p->prof_slot[kind].launches;
    if (total_ms) *total_ms = p->prof_slot[kind].total_ms;
    CAPI_END
}

int regot_profile_reset(regot_problem* p)
{
    CAPI_BEGIN
    check_handle(p);
    p->prof_collect();
    for (auto& s : p->prof_slot) {
        s.launches = 0;
        s.total_ms = 0.0;
    }
    CAPI_END
}

}  // extern "C"
