// dense.hpp -- launchers of the dense (streaming over M) EROT kernels.
//
// Conventions shared by every kernel here:
//   * M is row-major n x m with row pitch ldm (elements), 16-byte aligned rows.
//   * Potentials are FP64.  A "column potential" array always has m entries
//     (the gamma form's beta_{m-1} = 0 is materialised by the caller).
//   * Log-sums are returned in log2 units: lse2 = log2 sum 2^(x).
//   * k = log2(e) / reg converts (potential - M) / reg into log2 units.
//   * ctl (nullable) is the device control block of a running BCD solve:
//     ctl[0] = done flag (kernels return at once when set).
#pragma once

#include "common.hpp"

namespace regot {

// Per-row epilogue of the row pass.
struct RowEpi {
    double* lse2 = nullptr;          // log2 sum_j 2^((beta_j - M_ij) k)
    double* alpha_out = nullptr;     // reg * (log a_i - ln2 * lse2)
    const double* loga = nullptr;    // log a (needed by alpha_out and the shift guess)
    const double* rowpot = nullptr;  // alpha used for row sums / the shift guess
    double* rowsum = nullptr;        // sum_j exp((alpha_i + beta_j - M_ij)/reg)
    double reg = 1.0;
};

// Per-column epilogue of the column pass, run by the combine kernel.
struct ColEpi {
    double* lse2 = nullptr;          // log2 sum_i 2^((alpha_i - M_ij) k)
    double* beta_out = nullptr;      // reg * (log b_j - ln2 * lse2)
    const double* logb = nullptr;
    const double* b = nullptr;
    const double* colpot = nullptr;  // beta used for column sums / the shift guess
    double* colsum = nullptr;        // sum_i exp((alpha_i + beta_j - M_ij)/reg)
    // Column gradient: with beta_out, b_j * expm1(beta_out_j/reg + ln2 lse2 - log b_j)
    // (the FP64 rounding residual of the update); otherwise colsum_j - b_j.
    double* resid = nullptr;
    // Per combine workgroup t (j < m-1 only): [2t] = sum resid_j^2,
    // [2t+1] = sum beta_j b_j with beta = beta_out if set, else colpot.
    // In BCD mode the half selected by the parity of k+1 is written.
    double* tilepart = nullptr;
    double reg = 1.0;
};

// Row pass: one wave per R rows, streaming whole rows.  precise (FP32 bulk
// only): t, exp2 and the sums in FP64 over the FP32 M, for objectives and
// gradients that line searches difference at the 1e-16 relative level.
void launch_row_pass(int bulk, const void* M, long ldm, int n, int m, const double* beta,
                     double k, const RowEpi& epi, const int* ctl, hipStream_t s,
                     bool precise = false);

struct BcdState;

// Column pass: partial log2-sums per (row block, column) into part (nblk * m
// doubles), then a combine kernel (4 threads per column) with the epilogue.
// With st, the combine also takes the BCD stop decision (see BcdState).
int col_pass_blocks(int n, int m);
int col_combine_blocks(int m);
void launch_col_pass(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                     const double* colpot_guess, const double* logb, double k, double* part,
                     int nblk, const ColEpi& epi, const BcdState* st, const int* ctl,
                     hipStream_t s, bool precise = false);

// FP64 plan: plan_ij = exp((alpha_i + beta_j - M_ij) / reg) for rows [r0, r0+nr).
void launch_plan(int bulk, const void* M, long ldm, int n, int m, const double* alpha,
                 const double* beta, double reg, double* plan, long r0, int nr, hipStream_t s);

// On-device BCD bookkeeping.  The row pass (FIN) publishes per-workgroup sums
// of (g^2, r, alpha a); every combine workgroup of the same iteration then
// reduces them (with the column-tile sums of the previous update,
// double-buffered by the parity of k) to the same f_k and ||g_k|| and takes
// the same stop decision.  k is the iteration the launch belongs to (the host
// enqueues iterations in order; launches past the stop return at once).
struct BcdState {
    const double* a;
    const double* beta;
    int k;                  // iteration of this launch
    int* ctl;               // [0] done, [1] iterations completed, [2] niter
    double* hist_f;         // rings of hist_cap entries (device-visible)
    double* hist_g;
    long long* hist_ts;
    int hist_cap;
    double* rpart;          // per row-pass workgroup: sum g^2, sum r, sum alpha a
    double* bm;             // beta_k[m-1], stored by the row pass of iteration k
    const double* cpart;    // 2 x ntiles x (sum resid^2, sum beta_j b_j), j < m-1
    int nrb;                // row-pass workgroups
    int ntiles;             // combine workgroups
    double sum_a;           // sum_i a_i
    double sum_b1;          // sum_{j < m-1} b_j
    double tol;
    double reg;
    int max_iter;
    int n, m;
    // Stall test (stall_window > 0): the loop also stops, with ctl[0] = 2,
    // when ||g|| has not reached a new minimum for stall_window iterations
    // while still >= tol -- the FP32-bulk sums have hit their resolution.
    // stall holds two slots {min ||g|| so far, its iteration} by the parity
    // of k: iteration k reads slot k & 1, workgroup 0 writes the other.
    double* stall;
    int stall_window;
};
// Row pass publishing the per-workgroup BCD sums (FIN variant); precise (FP32
// bulk): FP64 exp2 and sums.
void launch_row_pass_bcd(int bulk, const void* M, long ldm, int n, int m, const double* beta,
                         double k, const RowEpi& epi, const BcdState& st, hipStream_t s,
                         bool precise = false);
int row_pass_blocks(int bulk, int n);

// Single-read BCD iteration (FP32 bulk; bcd_fused.hip): per workgroup of whole
// rows, alpha_out = optimal_alpha(beta) for its rows, the row partials of
// BcdState (g^2, r, alpha'a at iterate k) and the FP64 column sums of
// exp((alpha_out_i + beta_j - M_ij) / reg) over its rows into part[blk * m + j].
// launch_col_combine_sums then finishes the iteration like the two-pass
// combine (epi.colpot = beta).
bool bcd_fused_shape_ok(int n, int m, int ncu);
int bcd_fused_blocks(int n, int ncu);
void launch_bcd_fused(const void* M, long ldm, int n, int m, const double* beta, double k,
                      double reg, const double* loga, const double* a, const double* alpha,
                      double* alpha_out, int nblk, double* part, double* rpart, double* bm,
                      const int* ctl, hipStream_t s);
void launch_col_combine_sums(const double* part, int nblk, int m, double k, const ColEpi& epi,
                             const BcdState& st, hipStream_t s);

void launch_fill(double* p, double v, long count, hipStream_t s);
void launch_fill_int(int* p, int v, long count, hipStream_t s);
void launch_stamp(long long* dst, hipStream_t s);
// gamma -> (alpha, beta) with beta_{m-1} = 0 and the given scale: out = x * k.
void launch_split_gamma(const double* gamma, int n, int m, double* alpha, double* beta,
                        hipStream_t s);

}  // namespace regot
