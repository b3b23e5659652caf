// sparse.hpp -- the sparsified dual Hessian on the GPU.
//
//   Hs = [ diag(h1)  sigma    ]     h1 = r / reg, h2 = ct / reg,
//        [ sigma'    diag(h2) ]     sigma = kept entries of T[:, 0:m-1] / reg
//
// sigma is held twice, as CSR (rows, for sigma x) and CSC (columns, for
// sigma' y), so both products are deterministic gathers.  Both copies are
// built atomic-free: count kept entries per row / column, exclusive scan,
// write in index order.  The kept set is decided element-wise from values
// recomputed from M (row kernels) or its transpose MT (column kernels) with
// one shared element function, so both copies agree bit for bit.
#pragma once

#include "common.hpp"

namespace regot {

// What decides whether entry (i, j), j < m-1, of T is kept.
struct KeepRule {
    int mode = 0;                        // 0 = threshold rule, 1 = top-k rule
    // threshold rule: kept <=> T_ij >= min(tau_j, rho_i), or T_ij == 0 in FP64
    const double* tau = nullptr;         // per column
    const double* rho = nullptr;         // per row
    // top-k rule: kept <=> key > kstar, or key == kstar and the tie is among the
    // first tie_take ties in row-major order, or (preserve) i == 0 or j == 0
    unsigned long long kstar = 0;
    long long tie_take = 0;
    const long long* tie_off = nullptr;  // per row: ties in earlier rows
    int preserve = 1;
};

// Everything the element function needs: T_ij from alpha_i, beta_j, M_ij.
struct ElemCtx {
    const double* alpha = nullptr;   // n
    const double* beta = nullptr;    // m (beta[m-1] = 0)
    double k = 0.0;                  // log2(e) / reg
    double reg = 1.0;
    int n = 0, m = 0;
};

// An entry of the top-k boundary bucket: its key (the bit pattern of T) and
// position.
struct TopkCand {
    unsigned long long key;
    int i, j;
};

struct SparseHessian {
    int n = 0, m = 0;
    long long nnz = 0;        // stored entries of sigma
    long long nnz_zero = 0;   // threshold rule: kept entries that are exactly 0 in FP64
    DBuf<double> h1, h2;
    DBuf<int> rowptr, colind, colptr, rowind;
    DBuf<double> rval, cval;
    // work
    DBuf<double> tau, rho;
    DBuf<int> rcnt, ccnt, zcnt, csc_work;
    DBuf<long long> ltmp;
    DBuf<unsigned long long> hist;   // global count histograms of the top-k radix passes
    DBuf<TopkCand> topk_cand;        // top-k: entries of the boundary bucket
    DBuf<unsigned long long> topk_ncand;
    DBuf<double> vn, vm, vm2;        // alpha, beta (full) of the current build; scratch
    // tau from row-major M (FP32 bulk): level-1 partials, boundary buckets,
    // candidate offsets and the candidates
    DBuf<double> tau_pm, tau_base;
    DBuf<unsigned int> tau_pc, tau_px;
    DBuf<int> tau_bkt, tau_cnt, tau_off;
    DBuf<float> tau_cand;
    DBuf<double> cg_r, cg_p, cg_q, cg_s, cg_u, cg_y_buf, cg_part, cg_scal, cg_dinv;
    DBuf<int> cg_ctl;
    // the second right-hand side of a paired CG (hess_cg2): the same state
    DBuf<double> cg1_r, cg1_p, cg1_q, cg1_s, cg1_u, cg1_y, cg1_part, cg1_scal, cg1_rhs;
    DBuf<int> cg1_ctl;
    HBuf<int> cg1_hctl;
    HBuf<double> hbuf1;
    // Matrix-free build (FP32 bulk): no CSR / CSC; the kept entries as the
    // dense FP32 matrix mf_t (K_ij = T_ij where kept, else 0; row pitch
    // mf_ldt) and every product is a pass over it by mf_nwg workgroups,
    // whose column partials go to mf_part.
    bool mf = false;
    int mf_nwg = 0;
    long mf_ldt = 0;
    DBuf<float> mf_t;
    DBuf<double> mf_part;
    HBuf<int> cg_hctl;
    HBuf<double> hbuf;
    HBuf<long long> hll;
    double density() const
    {
        const double N = static_cast<double>(n) + m - 1;
        return (2.0 * static_cast<double>(nnz + nnz_zero) + n + m - 1) / N / N;
    }
};

// M^T (m x n, row pitch ldt) from M (n x m, pitch ldm); bulk 0 = FP32, 1 = FP64.
void launch_transpose(int bulk, const void* M, long ldm, int n, int m, void* MT, long ldt,
                      hipStream_t s);

// Threshold rule, step 1: tau_j for every column j < m-1 (one workgroup per
// column over MT).  find_small: columns whose whole mass is <= delta keep
// nothing (tau = +inf); otherwise they keep their largest entry.
void launch_col_threshold(int bulk, const void* MT, long ldt, const ElemCtx& c, double delta,
                          int find_small, double* tau, hipStream_t s);
// Threshold rule, step 1 for FP32 bulk from row-major M (no M'): the same
// tau_j by a histogram pass over 64 exponent buckets per column, a gather of
// each column's boundary bucket and radix passes over those candidates.
// Synchronises the stream once (the candidate count).
void threshold_tau_rows(const void* M, long ldm, const ElemCtx& c, double delta, int find_small,
                        SparseHessian& H, double* tau, hipStream_t s);
// Threshold rule, step 2: rho_i per row (dropped-mass correction to 2 delta),
// stored-entry count per row and FP64-zero kept count per row.
void launch_row_threshold(int bulk, const void* M, long ldm, const ElemCtx& c,
                          const double* tau, double delta2, double* rho, int* cnt, int* zcnt,
                          hipStream_t s);

// Top-k rule: radix passes over all entries of T[:, 0:m-1] to the key of the
// k-th largest value.  Fills *kstar and *n_greater (entries with key > kstar).
void topk_select(int bulk, const void* M, long ldm, const ElemCtx& c, long long k,
                 SparseHessian& H, unsigned long long* kstar, long long* n_greater,
                 hipStream_t s);
// The same selection with two reads of M (the top digit's histogram, then a
// pass that gathers the boundary bucket's entries and tallies every row
// outside it), the remaining digits over those candidates; also fills the
// row tallies of launch_row_tally (tcnt, gcnt) for 0 < k < n (m-1).
void topk_select_rows(int bulk, const void* M, long ldm, const ElemCtx& c, long long k,
                      int preserve, SparseHessian& H, unsigned long long* kstar,
                      long long* n_greater, int* tcnt, int* gcnt, hipStream_t s);
// Top-k tallies per row: ties (key == kstar) into tcnt; 2 x (stored entries
// kept without a tie slot) + (column 0 is a preserved tie of a row > 0) into gcnt.
void launch_row_tally(int bulk, const void* M, long ldm, const ElemCtx& c,
                      unsigned long long kstar, int preserve, int* tcnt, int* gcnt,
                      hipStream_t s);
// Stored entries per row under the top-k rule from the tallies and r.tie_off.
void launch_topk_count(const int* tcnt, const int* gcnt, const KeepRule& r, int n, int* cnt,
                       hipStream_t s);
// Write the CSR (values T / reg) at the given row pointers.
void launch_row_write(int bulk, const void* M, long ldm, const ElemCtx& c, const KeepRule& r,
                      const int* rowptr, int* colind, double* val, hipStream_t s);

// CSC (colptr, rowind, cval) of the kept set whose CSR H holds: column
// counts per block of rows_per_block rows from the CSR (work = (n /
// rows_per_block) x (m-1) ints, colcnt = m-1 ints), then, by `how`: 0 a
// stable counting sort of the CSR entries (sparse sets); 1 a pass over M by
// column strips under the same rule r (dense sets); 2 the ordered row writer
// on M' (threshold rule, dense sets, M' present).  The entries of a column
// come out in row order every way.
void launch_build_csc(int bulk, const void* M, long ldm, const void* MT, long ldt,
                      const ElemCtx& c, const KeepRule& r, const SparseHessian& H, int how,
                      int* work, int rows_per_block, int* colcnt, int* colptr, long long* total,
                      int* rowind, double* cval, hipStream_t s);

// Exclusive scan of count (int) into out (count + 1 entries, int64 total check).
void launch_scan(const int* cnt, int count, int* out, long long* total, hipStream_t s);
// Exclusive scan of int counts into int64 offsets (count entries).
void launch_scan64(const int* cnt, int count, long long* out, hipStream_t s);

// y = sigma x (x: m-1, y: n), optionally y /= (h1 + shift) when div.
void launch_spmv_rows(const SparseHessian& H, const double* x, double* y, bool div,
                      double shift, hipStream_t s);
// Column products (y: n; z, out: m-1):
//   mode 0: out = sigma' y
//   mode 1: out = h2 .* z + shift z - sigma' y   (Delta form)
//   mode 2: out = h2 .* z + sigma' y             (lower block of Hs)
void launch_spmv_cols(const SparseHessian& H, const double* y, const double* z, double shift,
                      int mode, double* out, hipStream_t s);

}  // namespace regot
