// hess_mf.hpp -- products with the FP32-bulk sparsified Hessian held as the
// dense matrix of its kept entries (hess_mf.hip).
#pragma once

#include "common.hpp"

namespace regot {

// What one pass over the kept matrix K computes (sigma = K / reg, j < m-1):
//   MF_CG    y_i = (sigma xc)_i / (h1_i + shift); column partials of sigma' y
//   MF_HS    rowout = sigma xc; column partials of sigma' xr
//   MF_ROWS  rowout = sigma xc
//   MF_COLS  column partials of sigma' xr
//   MF_DIAG  column partials of sum_i sigma_ij^2 / (h1_i + shift)
enum { MF_CG = 0, MF_HS = 1, MF_ROWS = 2, MF_COLS = 3, MF_DIAG = 4 };

struct MfArgs {
    // builds: T from M
    const float* M = nullptr;
    long ldm = 0;
    int n = 0, m1 = 0;             // rows, columns of sigma (m - 1)
    const double* alpha = nullptr;  // n
    const double* beta = nullptr;   // m (beta[m-1] = 0)
    double k = 0.0;                 // log2(e) / reg
    const double* tau = nullptr;    // per column (threshold rows)
    float* Tout = nullptr;          // the kept matrix written by a build (pitch ldt), or null
    // products
    const float* T = nullptr;       // the kept matrix K (K_ij = T_ij where kept, else 0)
    long ldt = 0;                   // row pitch of K (floats)
    double inv_reg = 0.0;
    const double* xc = nullptr;     // m1: input of the row half
    const double* xr = nullptr;     // n: input of the column half (HS, COLS)
    const double* h1 = nullptr;
    double shift = 0.0;
    double* rowout = nullptr;       // n (HS, ROWS)
    double* part = nullptr;         // nwg x m1 column partials
    int* ctl = nullptr;             // nullable: ctl[0] != 0 -> return at once
    // MF_CG inside the device-controlled CG (nullable): stop test on the sum
    // of rr_n partials of r.r against scal[5]; on stop workgroup 0 records
    // scal[4] = r.r, ctl[1] = it, ctl[0] = 1 and no workgroup does the pass
    const double* rr_part = nullptr;
    int rr_n = 0;
    double* scal = nullptr;
    int it = 0;
    int rpw = 0;                    // rows per workgroup (set by the launcher)
    int nt = 0;                     // nontemporal loads (set by the launcher)
};

bool mf_shape_ok(int n, int m);
// Row pitch (floats) of the kept matrix: m - 1 rounded up to 64 (256 bytes).
long mf_pitch(int m);
int mf_blocks(int n, int ncu);
// The full Hessian's kept matrix (every entry of T[:, 0:m-1]) into a.Tout.
void launch_mf_full(MfArgs a, int nwg, hipStream_t s);
// Row side of the threshold rule from tau (a.tau) in one read of M: rho,
// stored and FP64-zero counts per row (k_row_threshold's outputs) and, when
// a.Tout is not null, the kept matrix.
void launch_mf_rule(MfArgs a, int nwg, double delta2, double* rho, int* cnt, int* zcnt,
                    hipStream_t s);
void launch_mf_pass(int mode, MfArgs a, int nwg, hipStream_t s);
// out (m1) from the nwg x m1 partials: epi 0 sum; 1 z - sum; 2 (h2 + shift) z -
// sum; 3 h2 z + sum; 4 the Jacobi inverse diagonal 1 / (h2 + shift - sum).
void launch_mf_fold(const double* part, int nwg, int m1, int epi, const double* z,
                    const double* h2, double shift, double* out, hipStream_t s);

}  // namespace regot
