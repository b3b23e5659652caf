// dense.hpp -- launchers of the dense (streaming over M) EROT kernels.
//
// Conventions shared by every kernel here:
//   * M is row-major n x m with row pitch ldm (elements), 16-byte aligned rows.
//   * Potentials are FP64.  A "column potential" array always has m entries
//     (the gamma form's beta_{m-1} = 0 is materialised by the caller).
//   * Log-sums are returned in log2 units: lse2 = log2 sum 2^(x).
//   * k = log2(e) / reg converts (potential - M) / reg into log2 units.
//   * ctl (nullable) is the device control block of a running BCD solve:
//     ctl[0] = done flag (kernels return at once when set), ctl[1] = k.
#pragma once

#include "common.hpp"

namespace regot {

// Per-row epilogue of the row pass.  With parity set, the row potential and
// the output potential come from apair[k & 1] and apair[(k + 1) & 1], where k
// is ctl[1]: the double buffer of the on-device BCD loop.
struct RowEpi {
    double* lse2 = nullptr;          // log2 sum_j 2^((beta_j - M_ij) k)
    double* alpha_out = nullptr;     // reg * (log a_i - ln2 * lse2)
    const double* loga = nullptr;    // log a (needed by alpha_out and the shift guess)
    const double* rowpot = nullptr;  // alpha used for row sums / the shift guess
    double* rowsum = nullptr;        // sum_j exp((alpha_i + beta_j - M_ij)/reg)
    double reg = 1.0;
    bool parity = false;
    double* apair[2] = {nullptr, nullptr};
};

// Per-column epilogue of the column combine.
struct ColEpi {
    double* lse2 = nullptr;          // log2 sum_i 2^((alpha_i - M_ij) k)
    double* beta_out = nullptr;      // reg * (log b_j - ln2 * lse2)
    const double* logb = nullptr;
    const double* b = nullptr;
    const double* colpot = nullptr;  // beta used for column sums / the shift guess
    double* colsum = nullptr;        // sum_i exp((alpha_i + beta_j - M_ij)/reg)
    double* resid = nullptr;         // b_j * expm1(beta_out_j/reg + ln2 lse2 - log b_j)
    double reg = 1.0;
};

// Row pass: one wave per R rows, streaming whole rows.
void launch_row_pass(int bulk, const void* M, long ldm, int n, int m, const double* beta,
                     double k, const RowEpi& epi, const int* ctl, hipStream_t s);

// Column pass, stage 1: partial log2-sums per (row block, column) into part.
// Returns the number of row blocks used (size of part must be nblk * m).
int col_pass_blocks(int n, int m);
void launch_col_partial(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                        const double* alpha_pair0, const double* alpha_pair1,
                        const double* colpot_guess, const double* logb, double k,
                        double* part, int nblk, const int* ctl, hipStream_t s);
// Column pass, stage 2: combine the partials and run the epilogue.
void launch_col_combine(const double* part, int nblk, int m, double k, const ColEpi& epi,
                        const int* ctl, hipStream_t s);

// FP64 plan: plan_ij = exp((alpha_i + beta_j - M_ij) / reg) for rows [r0, r0+nr).
void launch_plan(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                 const double* beta, double reg, double* plan, long r0, int nr, hipStream_t s);

// On-device BCD bookkeeping (single workgroup): objective, ||grad||, history,
// convergence test.  See dense_kernels.hip for the exact update order.
struct BcdState {
    const double* a;
    const double* b;
    double* apair[2];
    const double* beta;
    const double* rowsum;
    const double* gcol;
    int* ctl;          // [0] done, [1] k, [2] niter
    double* hist_f;
    double* hist_g;
    long long* hist_ts;
    int hist_cap;      // history arrays are rings of this many entries
    double tol;
    double reg;
    int max_iter;
    int n, m;
};
void launch_bcd_finalize(const BcdState& st, hipStream_t s);

// FP64-exp row and column sums of T at (alpha, beta full-length), for the
// verification of a result independent of the bulk precision.
void launch_verify_sums(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                        const double* beta, double reg, double* rowsum, double* colsum,
                        hipStream_t s);

void launch_fill(double* p, double v, long count, hipStream_t s);
void launch_fill_int(int* p, int v, long count, hipStream_t s);
void launch_stamp(long long* dst, hipStream_t s);
// gamma -> (alpha, beta) with beta_{m-1} = 0 and the given scale: out = x * k.
void launch_split_gamma(const double* gamma, int n, int m, double* alpha, double* beta,
                        hipStream_t s);

}  // namespace regot
