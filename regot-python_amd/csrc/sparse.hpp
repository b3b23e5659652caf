// sparse.hpp -- sparsified dual Hessian on the GPU (threshold and top-k rules,
// CSR + CSC storage, Hessian-vector products and CG).
#pragma once

#include "common.hpp"

namespace regot {

struct SparseHessian {
    int n = 0, m = 0;
    long long nnz = 0;
};

}  // namespace regot
