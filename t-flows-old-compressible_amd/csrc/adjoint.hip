// adjoint.hip -- transposed solves and sparse outer-product gradients.
//
// For x = A^-1 b and a scalar loss L(x), the adjoint state lambda solves
// A^T lambda = dL/dx; then dL/db = lambda and dL/dA_ij = -lambda_i x_j for
// every stored entry (i, j).  The reference never forms A^T; its pattern is
// structurally symmetric (TopolM.f90:65-66 inserts (c1,c2) and (c2,c1) for
// every face), so A^T lives on the same pattern: the value of (A^T)_ij is the
// value stored at the mirror position of (i, j), the role SidAij(1:2,s)
// plays for faces (TopolM.f90:81-94).  The transposed solve runs the same
// CG / BiCG / CGS kernels on those values.
#include <algorithm>

#include "system.h"

namespace tflows {

// a_t[p] = a[mirror[p]]
__global__ __launch_bounds__(kBlock) void k_mirror_gather(int64_t nnz, const int* __restrict__ mirror,
                                                          const double* __restrict__ a, double* __restrict__ a_t) {
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p < nnz) a_t[p] = a[mirror[p]];
}

// grad[p] = -lambda[row(p)] * x[col(p)] for the entries of each row; one
// thread per row; slot NONZ (the trailing unused one) is zeroed by the host.
__global__ __launch_bounds__(kBlock) void k_outer_grad(int n, const int* __restrict__ rowptr,
                                                       const int* __restrict__ col,
                                                       const double* __restrict__ lambda,
                                                       const double* __restrict__ x, double* __restrict__ grad) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double li = lambda[i];
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) grad[p] = -(li * x[col[p]]);
}

// Mirror positions, built on the host once per pattern: for entry p = (i, j)
// the position of (j, i), found by binary search in row j's sorted columns.
const int* System::transpose_map() {
    if (tpos_ready_) return tpos_.p;
    std::vector<int> m(nnz_);
    for (int i = 0; i < n_; ++i) {
        for (int p = h_rowptr_[i]; p < h_rowptr_[i + 1]; ++p) {
            const int j = h_col_[p];
            const int* b = h_col_.data() + h_rowptr_[j];
            const int* e = h_col_.data() + h_rowptr_[j + 1];
            const int* f = std::lower_bound(b, e, i);
            if (f == e || *f != i)
                throw std::invalid_argument("the pattern is not structurally symmetric: A^T needs (j,i) for (i,j)");
            m[p] = (int)(f - h_col_.data());
        }
    }
    tpos_.ensure(nnz_);
    TF_CHECK(hipMemcpy(tpos_.p, m.data(), sizeof(int) * nnz_, hipMemcpyHostToDevice));
    tpos_ready_ = true;
    return tpos_.p;
}

System& System::transposed(hipStream_t s) {
    const int* mir = transpose_map();
    if (!transposed_) {
        // same pattern (the constructor validates and uploads it again)
        std::vector<int> acol(n_ + 1), arow(nnz_), adia(n_);
        for (int i = 0; i <= n_; ++i) acol[i] = h_rowptr_[i] + 1;
        for (int p = 0; p < nnz_; ++p) arow[p] = h_col_[p] + 1;
        for (int i = 0; i < n_; ++i) adia[i] = h_dia_[i] + 1;
        transposed_.reset(new System(n_, nnz_ + 1, acol.data(), arow.data(), adia.data(), false, s));
        transposed_valid_ = false;
    }
    if (!transposed_valid_) {
        // A^T values in the caller's entry order, then stored the way the
        // transposed system keeps its own values
        const int64_t nnz = nnz_;
        tvals_.ensure(nnz_);
        hipLaunchKernelGGL(k_mirror_gather, dim3((unsigned)((nnz + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, nnz,
                           mir, values_original(), tvals_.p);
        TF_CHECK(hipGetLastError());
        transposed_->set_values(tvals_.p, true, s);
        transposed_valid_ = true;
    }
    return *transposed_;
}

void System::outer_grad(const double* lambda, const double* x, double* grad, hipStream_t s) {
    // over the caller's pattern and numbering (lambda, x, grad all in it)
    const int* rp = reordered_ ? orowptr_.p : rowptr_.p;
    const int* cl = reordered_ ? ocol_.p : col_.p;
    hipLaunchKernelGGL(k_outer_grad, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n_, rp, cl,
                       lambda, x, grad);
    TF_CHECK(hipMemsetAsync(grad + nnz_, 0, sizeof(double), s));   // the NONZ-th slot
    TF_CHECK(hipGetLastError());
}

}  // namespace tflows
