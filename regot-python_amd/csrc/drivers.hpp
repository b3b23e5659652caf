// drivers.hpp -- solver drivers behind regot_solve() and their result object.
#pragma once

#include <vector>

#include "problem.hpp"
#include "regot_hip.h"

struct regot_result {
    int solver = 0;
    int n = 0, m = 0;
    int niter = 0;
    std::vector<double> dual, obj_vals, mar_errs, run_times, densities;
    // extras, not in the reference result (Newton drivers): host wall time and
    // counts per phase -- [0] objective / gradient passes, [1] Hessian
    // sparsification, [2] CG solves -- and the total CG iterations
    double phase_ms[3] = {0.0, 0.0, 0.0};
    long long phase_n[3] = {0, 0, 0};
    long long cg_iters = 0;
    double alg_bytes = 0.0;   // algorithmic HBM bytes of the solve (Newton-type drivers)
    // BCD in FP32 bulk: the iterate from which the loop ran in FP64 arithmetic
    // (0: from the start, -1: never)
    int precise_from = -1;
    // APDAGD's primal plan (device, row pitch plan_ld); empty for the others,
    // whose plan is formed from the dual on request
    regot::DBuf<double> plan;
    long plan_ld = 0;
    int plan_device = 0;
    // QROT results (regot_qrot_solve): dual = [alpha; beta] (n + m), the primal
    // value of the rounded plan per history entry, plan (alpha (+) beta - M)+ / reg
    bool qrot = false;
    std::vector<double> prim_vals;
};

namespace regot {

void solve_ssns(regot_problem* p, double tol, int max_iter, int verbose, const regot_opts& o,
                regot_result* res);
// splr = false: sparse Newton; true: SPLR.
void solve_newton_family(regot_problem* p, bool splr, double tol, int max_iter, int verbose,
                         const regot_opts& o, regot_result* res);
// Newton's method with the full dual Hessian (newton.cpp).
void solve_newton_dense(regot_problem* p, double tol, int max_iter, int verbose,
                        const regot_opts& o, regot_result* res);
// L-BFGS on the dual and APDAGD (first_order.cpp).
void solve_lbfgs_dual(regot_problem* p, double tol, int max_iter, int verbose,
                      const regot_opts& o, regot_result* res);
void solve_apdagd(regot_problem* p, double tol, int max_iter, int verbose, const regot_opts& o,
                  regot_result* res);

// QROT drivers (qrot.cpp): solver 0 = BCD, 1 = gradient descent, 2 = L-BFGS,
// 3 = L-BFGS on the semi-dual, 4 = APDAGD (keeps its primal plan in res->plan).
void solve_qrot(regot_problem* p, int solver, double tol, int max_iter, int verbose,
                const regot_opts& o, regot_result* res);

}  // namespace regot
