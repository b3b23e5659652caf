// hessian.hpp -- sparsified Hessian build, products and CG on a problem handle.
#pragma once

#include "problem.hpp"
#include "sparse.hpp"

namespace regot {

SparseHessian& hess_of(regot_problem* p);
void hess_ensure_transpose(regot_problem* p);
// Threshold rule (SSNS): sparsify_mat3 semantics at (gamma, grad), device vectors.
// driver: the caller needs only products, so under the handle's default
// storage mode a dense kept set (FP32 bulk, m - 1 <= 16384) is held as a
// bitmask and every product is a pass over M (SparseHessian::mf).
void hess_threshold(regot_problem* p, const double* d_gamma, const double* d_grad, double delta,
                    double density_hint, bool driver = false);
// Top-k rule (sparse Newton, SPLR): sparsify_mat_with_density semantics, preserve = true.
// At density 1 (every entry kept) the build may go matrix-free as above.
void hess_topk(regot_problem* p, const double* d_gamma, const double* d_grad, double density,
               bool driver = false);
void hess_apply_delta(regot_problem* p, const double* x, double shift, double* out);
void hess_apply_hs(regot_problem* p, const double* x, double* out);
// CG on the Schur complement Delta (m-1) from x0 (device, m-1) or zero;
// relative residual tol against rhs; jacobi: preconditioned by diag(Delta);
// returns iterations.
int hess_cg_delta(regot_problem* p, const double* rhs, double shift, double tol, int max_iter,
                  double* x, double* tol_error, const double* x0 = nullptr,
                  bool jacobi = false);
// (Hs + shift I)^{-1} rhs (n+m-1) by block elimination; x0 (device, n+m-1,
// only its last m-1 entries are used) warm-starts the CG; returns CG iterations.
int hess_cg(regot_problem* p, const double* rhs, double shift, double tol, int max_iter,
            double* out, double* tol_error, const double* x0 = nullptr, bool jacobi = false);
// Two right-hand sides of the same system, their CGs run in the same
// launches (stored Hessian; two hess_cg solves when matrix-free); each gives
// the iterates of its own hess_cg bit for bit.  iters[2], tol_error[2]
// (nullable); returns the sum of the iterations.
int hess_cg2(regot_problem* p, const double* rhs0, const double* rhs1, double shift, double tol,
             int max_iter, double* out0, double* out1, int* iters, double* tol_error,
             bool jacobi = false);

}  // namespace regot
