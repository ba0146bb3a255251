// reorder.hip -- locality reordering of a system's internal storage.
//
// A randomly renumbered mesh (the momentum and adjoint systems the
// reference solves after Divide/renumbering, or any mesh written by an
// unstructured generator) scatters the column indices of every row over
// the whole vector, so each gather of the SpMV touches a different cache
// line.  The System then stores its matrix in reverse Cuthill-McKee order:
// rows are permuted symmetrically, and within each row the entries keep
// their ORIGINAL order (column indices are renamed, not re-sorted).  Every
// row sum therefore adds the same products in the same order as the
// reference loop (Process/CG.f90:193-199), and the lower / upper split at
// Adia is unchanged, so the SpMV, Resid, Prec1 and the DIC sweeps give the
// same bits as without the permutation.  Vectors cross the boundary in the
// caller's numbering; the System gathers them in and scatters them out.
#include <algorithm>
#include <cstdlib>
#include <numeric>

#include "system.h"

namespace tflows {

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather(int n, const int* __restrict__ perm, const T* __restrict__ src,
                                                   T* __restrict__ dst) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scatter(int n, const int* __restrict__ perm, const T* __restrict__ src,
                                                    T* __restrict__ dst) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dst[perm[i]] = src[i];
}

template <typename T>
void System::to_internal(const T* ext, T* in, hipStream_t s) const {
    hipLaunchKernelGGL((k_gather<T>), dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n_,
                       (const int*)perm_.p, ext, in);
    TF_CHECK(hipGetLastError());
}

template <typename T>
void System::to_external(const T* in, T* ext, hipStream_t s) const {
    hipLaunchKernelGGL((k_scatter<T>), dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n_,
                       (const int*)perm_.p, in, ext);
    TF_CHECK(hipGetLastError());
}

template void System::to_internal<double>(const double*, double*, hipStream_t) const;
template void System::to_internal<float>(const float*, float*, hipStream_t) const;
template void System::to_external<double>(const double*, double*, hipStream_t) const;
template void System::to_external<float>(const float*, float*, hipStream_t) const;

// Entries [e0, e1) of old row o hold columns col[e]; the graph of the
// pattern (structurally symmetric in T-FlowS, TopolM.f90:65-66) is walked
// through its off-diagonal entries.
namespace {

struct Graph {
    const std::vector<int>& rowptr;
    const std::vector<int>& col;
    int deg(int v) const { return rowptr[v + 1] - rowptr[v] - 1; }
};

// Breadth-first level structure from `root` over unvisited vertices;
// returns the vertices in visit order (the last is in the deepest level).
void bfs(const Graph& g, int root, const std::vector<char>& done, std::vector<int>& stamp, int tag,
         std::vector<int>& out) {
    out.clear();
    out.push_back(root);
    stamp[root] = tag;
    for (size_t h = 0; h < out.size(); ++h) {
        const int v = out[h];
        for (int e = g.rowptr[v]; e < g.rowptr[v + 1]; ++e) {
            const int w = g.col[e];
            if (w == v || done[w] || stamp[w] == tag) continue;
            stamp[w] = tag;
            out.push_back(w);
        }
    }
}

}  // namespace

// Reverse Cuthill-McKee order (new -> old), component by component, each
// started from a pseudo-peripheral vertex (two BFS sweeps from a
// minimum-degree vertex); neighbours are visited by increasing degree.
std::vector<int> rcm_order(int n, const std::vector<int>& rowptr, const std::vector<int>& col) {
    const Graph g{rowptr, col};
    std::vector<char> done(n, 0);
    std::vector<int> stamp(n, -1), order, level, nb;
    order.reserve(n);
    int tag = 0, scan = 0;
    while ((int)order.size() < n) {
        while (done[scan]) ++scan;
        // pseudo-peripheral start inside this component
        int root = scan;
        bfs(g, root, done, stamp, tag++, level);
        int best = root;
        for (int v : level) best = (g.deg(v) < g.deg(best)) ? v : best;
        root = best;
        for (int pass = 0; pass < 2; ++pass) {
            bfs(g, root, done, stamp, tag++, level);
            root = level.back();
        }
        // Cuthill-McKee from root
        const size_t first = order.size();
        order.push_back(root);
        done[root] = 1;
        for (size_t h = first; h < order.size(); ++h) {
            const int v = order[h];
            nb.clear();
            for (int e = rowptr[v]; e < rowptr[v + 1]; ++e) {
                const int w = col[e];
                if (w != v && !done[w]) {
                    done[w] = 1;
                    nb.push_back(w);
                }
            }
            std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return g.deg(a) < g.deg(b); });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

// Mean |column - row| over the off-diagonal entries under the numbering
// `num` (old -> new; identity when null).
static double mean_offset(int n, const std::vector<int>& rowptr, const std::vector<int>& col, const int* num) {
    double sum = 0.0;
    int64_t cnt = 0;
    for (int r = 0; r < n; ++r)
        for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) {
            const int c = col[e];
            if (c == r) continue;
            const int a = num ? num[r] : r, b = num ? num[c] : c;
            sum += std::abs(a - b);
            ++cnt;
        }
    return cnt ? sum / (double)cnt : 0.0;
}

// TFLOWS_REORDER: 0 never, 1 always, anything else (default) when the
// reordering shortens the mean column offset at least 4x.
static int reorder_mode() {
    const char* e = std::getenv("TFLOWS_REORDER");
    if (!e) return 2;
    return std::atoi(e) == 0 ? 0 : (std::atoi(e) == 1 ? 1 : 2);
}

std::vector<int> locality_order(int n, const std::vector<int>& rowptr, const std::vector<int>& col, bool allow) {
    const int mode = allow ? reorder_mode() : 0;
    if (mode == 0 || n <= 1) return {};
    std::vector<int> perm = rcm_order(n, rowptr, col);
    std::vector<int> inv(n);
    for (int i = 0; i < n; ++i) inv[perm[i]] = i;
    const double before = mean_offset(n, rowptr, col, nullptr);
    const double after = mean_offset(n, rowptr, col, inv.data());
    if (mode == 1 || before > 4.0 * after) return perm;
    return {};
}

void launch_gather(int n, const int* perm, const double* src, double* dst, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL((k_gather<double>), dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, perm, src, dst);
    TF_CHECK(hipGetLastError());
}
void launch_scatter(int n, const int* perm, const double* src, double* dst, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL((k_scatter<double>), dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n, perm, src, dst);
    TF_CHECK(hipGetLastError());
}

// Decide on and build the internal (reordered) storage from the original
// 0-based host pattern.  Fills the device pattern arrays either way.
void System::build_storage(bool allow_reorder, hipStream_t s) {
    std::vector<int> perm = locality_order(n_, h_rowptr_, h_col_, allow_reorder);
    reordered_ = !perm.empty();
    if (reordered_) {
        iperm_h_.assign(n_, 0);
        for (int i = 0; i < n_; ++i) iperm_h_[perm[i]] = i;
    }
    if (!reordered_) {
        TF_CHECK(hipMemcpyAsync(rowptr_.p, h_rowptr_.data(), sizeof(int) * (n_ + 1), hipMemcpyHostToDevice, s));
        TF_CHECK(hipMemcpyAsync(col_.p, h_col_.data(), sizeof(int) * nnz_, hipMemcpyHostToDevice, s));
        TF_CHECK(hipMemcpyAsync(dia_.p, h_dia_.data(), sizeof(int) * n_, hipMemcpyHostToDevice, s));
        TF_CHECK(hipStreamSynchronize(s));
        return;
    }
    std::vector<int> rp(n_ + 1), cl(nnz_), dg(n_), ep(nnz_);
    rp[0] = 0;
    for (int i = 0; i < n_; ++i) {
        const int o = perm[i];
        rp[i + 1] = rp[i] + (h_rowptr_[o + 1] - h_rowptr_[o]);
    }
    for (int i = 0; i < n_; ++i) {
        const int o = perm[i];
        int k = rp[i];
        for (int e = h_rowptr_[o]; e < h_rowptr_[o + 1]; ++e, ++k) {
            cl[k] = iperm_h_[h_col_[e]];
            ep[k] = e;
        }
        dg[i] = rp[i] + (h_dia_[o] - h_rowptr_[o]);
    }
    perm_.ensure(n_);
    eperm_.ensure(nnz_);
    orowptr_.ensure(n_ + 1);
    ocol_.ensure(nnz_);
    val_orig_.ensure(nnz_);
    TF_CHECK(hipMemcpyAsync(rowptr_.p, rp.data(), sizeof(int) * (n_ + 1), hipMemcpyHostToDevice, s));
    TF_CHECK(hipMemcpyAsync(col_.p, cl.data(), sizeof(int) * nnz_, hipMemcpyHostToDevice, s));
    TF_CHECK(hipMemcpyAsync(dia_.p, dg.data(), sizeof(int) * n_, hipMemcpyHostToDevice, s));
    TF_CHECK(hipMemcpyAsync(perm_.p, perm.data(), sizeof(int) * n_, hipMemcpyHostToDevice, s));
    TF_CHECK(hipMemcpyAsync(eperm_.p, ep.data(), sizeof(int) * nnz_, hipMemcpyHostToDevice, s));
    TF_CHECK(hipMemcpyAsync(orowptr_.p, h_rowptr_.data(), sizeof(int) * (n_ + 1), hipMemcpyHostToDevice, s));
    TF_CHECK(hipMemcpyAsync(ocol_.p, h_col_.data(), sizeof(int) * nnz_, hipMemcpyHostToDevice, s));
    TF_CHECK(hipMemsetAsync(val_orig_.p, 0, sizeof(double) * nnz_, s));
    TF_CHECK(hipStreamSynchronize(s));
}

// val (internal order) = aval (original order) through the entry map.
void System::permute_values(hipStream_t s) {
    const int64_t nnz = nnz_;
    hipLaunchKernelGGL((k_gather<double>), dim3((unsigned)((nnz + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       (int)nnz, (const int*)eperm_.p, (const double*)val_orig_.p, val_.p);
    TF_CHECK(hipGetLastError());
}

}  // namespace tflows
