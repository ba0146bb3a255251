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

// grad[p] = -lambda[row(p)] * x[col(p)] for every stored entry.  One block
// per tile of 256 rows: the tile's row pointers go to LDS, and thread t
// takes entries e0 + t, e0 + t + 256, ... of the tile, finding each entry's
// row by binary search there, so the loads of col and the stores of grad
// are contiguous across the block.  Slot NONZ (the trailing unused one) is
// zeroed by the host.
__global__ __launch_bounds__(kBlock) void k_outer_grad(int n, const int* __restrict__ rowptr,
                                                       const int* __restrict__ col,
                                                       const double* __restrict__ lambda,
                                                       const double* __restrict__ x, double* __restrict__ grad) {
    __shared__ int s_rp[kTileRows + 1];
    const int ntiles = (n + kTileRows - 1) / kTileRows;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int r0 = t * kTileRows, nr = min(kTileRows, n - r0);
        __syncthreads();
        for (int i = threadIdx.x; i <= nr; i += kBlock) s_rp[i] = rowptr[r0 + i];
        __syncthreads();
        const int e0 = s_rp[0], e1 = s_rp[nr];
        for (int e = e0 + (int)threadIdx.x; e < e1; e += kBlock) {
            int lo = 0, hi = nr;                // the row: s_rp[lo] <= e < s_rp[lo + 1]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_rp[mid] <= e) lo = mid;
                else hi = mid;
            }
            grad[e] = -(lambda[r0 + lo] * x[col[e]]);
        }
    }
}

// 1 in *flag when some a[p] differs from a[mirror[p]] (A is not symmetric)
__global__ __launch_bounds__(kBlock) void k_asymmetry(int64_t nnz, const int* __restrict__ mirror,
                                                      const double* __restrict__ a, int* __restrict__ flag) {
    for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * kBlock)
        if (!(a[p] == a[mirror[p]])) *flag = 1;
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
    if (!transposed_valid_) {
        // a symmetric A (the pressure-correction matrix) is its own
        // transpose: the adjoint solve then runs on this system
        const int64_t nnz = nnz_;
        sym_flag_.ensure(1);
        TF_CHECK(hipMemsetAsync(sym_flag_.p, 0, sizeof(int), s));
        hipLaunchKernelGGL(k_asymmetry, dim3((unsigned)std::min<int64_t>(2048, (nnz + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, s, nnz, mir, values_original(), sym_flag_.p);
        TF_CHECK(hipGetLastError());
        int asym = 0;
        TF_CHECK(hipMemcpyAsync(&asym, sym_flag_.p, sizeof(int), hipMemcpyDeviceToHost, s));
        TF_CHECK(hipStreamSynchronize(s));
        symmetric_ = asym == 0;
        if (symmetric_) {
            transposed_valid_ = true;
            return *this;
        }
    } else if (symmetric_) {
        return *this;
    }
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
    const int tiles = (n_ + kTileRows - 1) / kTileRows;
    hipLaunchKernelGGL(k_outer_grad, dim3(std::min(tiles, 8192)), dim3(kBlock), 0, s, n_, rp, cl, lambda, x,
                       grad);
    TF_CHECK(hipMemsetAsync(grad + nnz_, 0, sizeof(double), s));   // the NONZ-th slot
    TF_CHECK(hipGetLastError());
}

}  // namespace tflows
