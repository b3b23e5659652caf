// problem.hpp -- the device-resident EROT problem behind the C-ABI handle.
#pragma once

#include <memory>
#include <vector>

#include "common.hpp"
#include "dense.hpp"

struct regot_problem;

namespace regot {

struct SparseHessian;  // sparse.hpp

// FP64 BLAS-1 helpers on device vectors (vec_kernels.hip).  Reductions run in
// one workgroup in a fixed order, so they are reproducible.
void vec_axpy(double* y, double alpha, const double* x, long count, hipStream_t s);
void vec_axpby(double* out, double alpha, const double* x, double beta, const double* y,
               long count, hipStream_t s);
void vec_copy(double* dst, const double* src, long count, hipStream_t s);
// gamma = [alpha + beta_{m-1}; beta_{0..m-2} - beta_{m-1}]
void vec_normalise(const double* alpha, const double* beta, int n, int m, double* gamma,
                   hipStream_t s);
// out[0..npairs) = dot(x_i, y_i) for the listed pairs, one launch.
void vec_dots(const double* const* xs, const double* const* ys, long count, int npairs,
              double* out, hipStream_t s);
// grad = [r - a; c[0:m-1] - b[0:m-1]], out[0] = reg sum r - alpha'a - betat'bt,
// out[1] = ||grad||^2.
void vec_objgrad_finish(const double* r, const double* c, const double* a, const double* b,
                        const double* gamma, int n, int m, double reg, double* grad,
                        double* out, hipStream_t s);
// out[0] = reg sum r - alpha'a - betat'bt.
void vec_obj_finish(const double* r, const double* a, const double* b, const double* gamma,
                    int n, int m, double reg, double* out, hipStream_t s);

// out[q] = x_q . y_q for up to 32 pairs over count entries, one pass of many
// workgroups (part: vec_mdots_scratch(count) doubles), result on the device.
void vec_mdots(const double* const* xs, const double* const* ys, int npairs, long count,
               double* part, double* out, hipStream_t s);
long vec_mdots_scratch(long count);
// out = sum_k coef[k] vs[k] (host coefficients, up to 16 device vectors).
void vec_lincomb(const double* const* vs, const double* coef, int nv, long count, double* out,
                 hipStream_t s);

// APDAGD primal-plan pass (apdagd_kernels.hip): P <- (1 - tau) P + tau T
// (P <- T when init), T_ij = exp((alpha_i + beta_j - M_ij) / reg) in FP64,
// P row-major with pitch ldp (even, >= m); out[0..2] = squared full marginal
// error, <M, P>, sum P (1 - log P).  Scratch sizes from apd_scratch_sizes.
void apd_scratch_sizes(int n, int m, long* rowpart, long* colpart, long* scal, long* part);
void launch_apd_plan(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                     const double* beta, const double* a, const double* b, double reg,
                     double tau, bool init, double* P, long ldp, double* rowpart,
                     double* colpart, double* scal, double* part, double* out, hipStream_t s);

// One-pass FP64 objective / gradient (objgrad_kernels.hip): T formed once per
// element, row sums per 256-column strip and column sums per row block in
// scratch (sizes from sums_scratch_sizes); out[0] = f, out[1] = ||g||^2;
// grad written when with_grad (else only f).
void sums_scratch_sizes(int n, int m, long* rowpart, long* colpart, long* part);
void launch_objgrad_fused(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                          const double* beta, const double* a, const double* b, double reg,
                          bool with_grad, double* grad, double* rowpart, double* colpart,
                          double* part, double* out, hipStream_t s);

// Row sums (per column strip of sums_strip_width() columns:
// rowpart[strip * n + i]) and, when colpart is set, column sums (per row
// block: colpart[block * m + j]) of w_ij = x_i T_ij y_j in FP64 (x, y
// nullable = 1); block sizes as the one-pass objective kernel (sums_rows /
// sums_scratch_sizes).
int sums_rows(int n, int m);
int sums_strip_width();
void launch_wsums(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                  const double* beta, double reg, const double* x, const double* y,
                  double* rowpart, double* colpart, hipStream_t s);

// y = 2^x by the device FP64 exponential of the one-pass kernels (dexp2.hpp).
void launch_exp2_check(const double* x, double* y, long count, hipStream_t s);
void launch_log_check(const double* x, double* y, long count, hipStream_t s);

// Buffers of the on-device BCD loop, kept across solves on one handle.
struct BcdWork {
    DBuf<double> ap0, ap1, beta, gcol, rpart, cpart, bm, stall;
    DBuf<double> fpart;         // column sums of the single-read pass (blocks x m)
    DBuf<int> ctl;
    MappedRing<double> hf, hg;
    MappedRing<long long> hts;
    HBuf<int> hc;
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~BcdWork()
    {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// Event pairs of one kind of pass.  A pair recorded with a tag >= 0 (the BCD
// iteration it belongs to) is held in `tagged` until the driver knows which
// iterations did work (prof_commit); untagged pairs count at once.
struct QrotScratch;

struct ProfilePair {
    hipEvent_t start, stop;
    long tag;
};
struct ProfileSlot {
    std::vector<ProfilePair> pending;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
    std::vector<std::pair<long, double>> tagged;
    long long launches = 0;
    double total_ms = 0.0;
};

}  // namespace regot

struct regot_problem {
    int device = 0;
    hipStream_t stream = nullptr;
    int n = 0, m = 0;
    long ldm = 0;
    int bulk = 0;               // 0 = FP32, 1 = FP64
    void* dM = nullptr;
    bool own_M = false;
    void* dMT = nullptr;        // M transposed (m x n, pitch ldt), built on first use
    long ldt = 0;
    double reg = 1.0;
    double k = 1.0;             // log2(e) / reg
    double sum_a = 0.0;         // sum_i a_i
    double sum_b1 = 0.0;        // sum_{j < m-1} b_j
    double sum_b = 0.0;         // sum_j b_j
    double min_a = 0.0, min_b = 0.0;
    double sumsq_a = 0.0, sumsq_b = 0.0;   // sum a_i^2, sum b_j^2
    int ncu = 0;                // compute units of the device
    int wall_khz = 0;           // device wall clock rate (timestamps)
    regot::DBuf<double> a, b, loga, logb;

    // scratch for the operator entry points and the drivers
    regot::DBuf<double> alpha, beta, rsum, csum, gamma, grad, part, scal;
    regot::DBuf<int> ctl;
    int nblk = 1;
    int ntiles = 1;
    std::unique_ptr<regot::BcdWork> bcd;
    regot::HBuf<double> hscal;
    regot::HBuf<int> hctl;

    // sparsified Hessian of the Newton-type drivers (sparse.hpp)
    std::unique_ptr<regot::SparseHessian> hess;
    // scratch of the QROT passes (qrot.hpp)
    std::unique_ptr<regot::QrotScratch> qrot;

    // FP32 bulk: run the general row / column passes with FP64 exp and sums
    // (objective and gradient evaluations of the Newton drivers, verification).
    bool precise = false;
    // FP32 bulk: the BCD loop's row and column passes with FP64 exp and sums
    // (the precise finish of solve_bcd).
    bool bcd_precise = false;
    // Storage of the threshold-rule Hessian: -1 the Newton drivers go
    // matrix-free (products straight from M, FP32 bulk) when the kept set is
    // dense enough, the Hessian C-ABI stores CSR / CSC; 0 always stored; 1
    // matrix-free wherever the bulk and shape allow.
    int hess_storage = -1;

    // profiling of the hot passes: [0] row pass, [1] column pass (partial +
    // combine), [2] single-read BCD pass, [3] its column-sum combine
    bool prof = false;
    long prof_tag = -1;         // tag of the passes being enqueued (-1: none)
    regot::ProfileSlot prof_slot[4];

    regot_problem() = default;
    regot_problem(const regot_problem&) = delete;
    regot_problem& operator=(const regot_problem&) = delete;
    ~regot_problem();

    void setup(int n_, int m_, const void* M, int M_dtype, int bulk_, int M_device,
               long long ldm_, const double* a_h, const double* b_h, double reg_, int device_);

    // --- streaming passes (device pointers) ---
    // Row pass over full-length column potential beta (m); see RowEpi.
    void row_pass(const double* beta_full, const regot::RowEpi& epi, const int* ctl_p = nullptr);
    // Column pass (partial + combine) at row potential alpha (n); see ColEpi.
    void col_pass(const double* alpha_v, const double* colpot_guess, const regot::ColEpi& epi,
                  const int* ctl_p = nullptr, const regot::BcdState* st = nullptr);
    // Row pass with the BCD bookkeeping fused (see BcdState).
    void row_pass_bcd(const double* beta_full, const regot::RowEpi& epi,
                      const regot::BcdState& st);
    // Single-read BCD iteration k (bcd_fused.hip): alpha_k -> alpha_out and the
    // column sums into the BCD workspace, then the stop decision and beta_{k+1}
    // (beta is both the input beta_k and the output).
    bool bcd_fused_ok() const;
    void bcd_fused_iteration(const double* alpha_k, double* alpha_out, double* beta,
                             const regot::ColEpi& epi, const regot::BcdState& st, int nblk_f,
                             double* fpart);

    // f(gamma) for device gamma (n+m-1); row pass only (FP32 bulk), or the
    // one-pass FP64 kernel (precise mode, FP64 bulk).
    double dev_obj(const double* d_gamma);
    double dev_obj_fused(const double* d_gamma, double* d_grad, double* gnorm2);
    regot::DBuf<double> s_row, s_col, s_part;   // scratch of the one-pass kernel
    // f and grad (device, n+m-1) for device gamma; row and column passes.
    // Row sums land in rsum, column sums in csum.  Returns f; *gnorm2 = ||g||^2.
    double dev_obj_grad(const double* d_gamma, double* d_grad, double* gnorm2);

    void sync();
    void prof_begin(int kind, hipEvent_t* stop_out);
    void prof_end(int kind, hipEvent_t stop);
    void prof_collect();
    // Count the tagged pairs with tag <= last, drop the others (launches
    // enqueued past a stop decision return at once and are not passes).
    void prof_commit(long last);
    // Drop the tagged pairs with tag > last (launches of a loop that stopped
    // at iterate `last`, which returned at once), keeping the others pending
    // a later commit.
    void prof_drop_after(long last);
};
