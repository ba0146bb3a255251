// system.hip -- orchestration of the CG / BiCG / CGS solves on one device
// matrix: pattern upload, preconditioner set-up, level analysis for the
// DIC sweeps, and the host loop that keeps the stream full of iterations.
#include "system.h"
#include "tile_plan.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <thread>
#include <type_traits>

namespace tflows {

// Upper bound on the blocks of the row-streaming kernels (default 2048 =
// 8 per CU); TFLOWS_MAX_BLOCKS overrides it for tuning.
// Jacobi CG kernel form of one system's solve: 3 (default, fastest
// measured) = p update, SpMV, update with device iteration counters (the
// form the mixed-precision path uses); 0 = the same three kernels indexed
// by the host (the batch kernels); 2 = p formed inside the SpMV gathers.
// TFLOWS_CG_FORM selects (read as iterations are issued: set it between
// solves).
int cg_form() {
    const char* e = std::getenv("TFLOWS_CG_FORM");
    return e ? std::atoi(e) : 3;
}
bool fused_cg_enabled() { return cg_form() != 0; }
// Jacobi CG of one system in forms 0 and 3 (and the mixed-precision inner
// CG): x += alfa p moved from the update kernel into the next p kernel,
// which reads p anyway (1, default); TFLOWS_CG_CARRY=0 keeps it in the
// update kernel (read as iterations are issued: set it between solves).
// Both give the same bits.
bool cg_carry() {
    const char* e = std::getenv("TFLOWS_CG_CARRY");
    return !e || std::atoi(e) != 0;
}
bool carries_x(const BatchCtx& c, int method, bool jac) {
    return c.blk_sys == nullptr && c.kbase == nullptr && method == 0 && jac && cg_form() != 2 && cg_carry();
}

// Lanes per row of the SpMV kernels (1, 2, 4, 8); TFLOWS_LANES sets the
// default of new systems.
static int lanes_from_env() {
    const char* e = std::getenv("TFLOWS_LANES");
    const int v = e ? std::atoi(e) : 1;
    return (v == 2 || v == 4 || v == 8) ? v : 1;
}

// Layout of an operator from the environment: "csr" 0, "sliced" 1,
// otherwise 2 (time both on first use).
static int op_mode_from_env(const char* name) {
    const char* e = std::getenv(name);
    if (e && std::string(e) == "csr") return 0;
    if (e && std::string(e) == "sliced") return 1;
    return 2;
}

// Mean time of `reps` launches of f on stream s (after two warm-up ones).
template <class F>
static float time_launches(hipStream_t s, int reps, F&& f) {
    hipEvent_t e0, e1;
    TF_CHECK(hipEventCreate(&e0));
    TF_CHECK(hipEventCreate(&e1));
    f();
    f();
    TF_CHECK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) f();
    TF_CHECK(hipEventRecord(e1, s));
    TF_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    TF_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / reps;
}

// True when launches of fa take no longer than launches of fb: three
// alternating rounds of 8 launches each, each form judged by its best
// round, so that one disturbed round (another process's work, a profiler's
// per-dispatch overhead) does not decide the choice.
template <class FA, class FB>
static bool first_is_faster(hipStream_t s, FA&& fa, FB&& fb) {
    float ta = 1e30f, tb = 1e30f;
    for (int round = 0; round < 3; ++round) {
        ta = std::min(ta, time_launches(s, 8, fa));
        tb = std::min(tb, time_launches(s, 8, fb));
    }
    return ta <= tb;
}

bool System::op64_sliced(hipStream_t s) {
    if (op64_mode_ == 2) {
        ws_.ensure((size_t)V_COUNT * n_);
        double* x = ws_.p + (size_t)V_P * n_;
        double* y = ws_.p + (size_t)V_Q * n_;
        TF_CHECK(hipMemsetAsync(x, 0, sizeof(double) * n_, s));
        const Sell<double> S = sell64(s);
        const Csr<double> A = csr();
        op64_mode_ = first_is_faster(s, [&] { launch_spmv(S, (const double*)x, y, grid_, s, ctx()); },
                                     [&] { launch_spmv(A, (const double*)x, y, grid_, s, ctx()); })
                         ? 1
                         : 0;
    }
    return op64_mode_ == 1;
}

bool System::op32_sliced(hipStream_t s) {
    if (op32_mode_ == 2) {
        extbuf32_.ensure(2 * (size_t)n_);
        float* x = extbuf32_.p;
        float* y = extbuf32_.p + n_;
        TF_CHECK(hipMemsetAsync(x, 0, sizeof(float) * n_, s));
        const DiffSell32 S = sell32(s);
        const DiffCsr32 A = diff32(s);
        op32_mode_ = first_is_faster(s, [&] { launch_spmv(S, (const float*)x, y, grid_, s, ctx()); },
                                     [&] { launch_spmv(A, (const float*)x, y, grid_, s, ctx()); })
                         ? 1
                         : 0;
    }
    return op32_mode_ == 1;
}

void slice_layout(int n, const std::vector<int>& rp, const std::vector<int>& cl, bool pad_own, std::vector<int>& off,
                  std::vector<int>& col, std::vector<int>& map) {
    const int nsl = (n + kSlice - 1) / kSlice;
    off.assign(nsl + 1, 0);
    for (int sl = 0; sl < nsl; ++sl) {
        int w = 0;
        for (int i = sl * kSlice; i < std::min(n, (sl + 1) * kSlice); ++i) w = std::max(w, rp[i + 1] - rp[i]);
        if ((int64_t)off[sl] + (int64_t)w * kSlice > INT32_MAX)
            throw std::invalid_argument("sliced layout exceeds 2^31 slots");
        off[sl + 1] = off[sl] + w * kSlice;
    }
    col.assign(off[nsl], 0);
    map.assign(off[nsl], -1);
    for (int sl = 0; sl < nsl; ++sl) {
        const int w = (off[sl + 1] - off[sl]) / kSlice;
        for (int l = 0; l < kSlice; ++l) {
            const int i = sl * kSlice + l;
            const int len = i < n ? rp[i + 1] - rp[i] : 0;
            for (int j = 0; j < w; ++j) {
                const int e = off[sl] + j * kSlice + l;
                if (j < len) {
                    col[e] = cl[rp[i] + j];
                    map[e] = rp[i] + j;
                } else {
                    col[e] = pad_own ? (i < n ? i : 0) : -1;
                }
            }
        }
    }
}

bool shared_slots_from_env() {
    const char* e = std::getenv("TFLOWS_SHARED_SLOTS");
    return !(e && e[0] == '0');
}

static_assert(kSlice == kSliceRows && kMaxSharedSlots == kMaxSliceSlots,
              "the host layouts and the kernels agree on the slice shape");
static_assert(sizeof(SlotPair) == sizeof(int2) && alignof(int2) <= 8, "mirror bases read as int2");

// The FP32 shared-offset SpMV loop over all slots at once (TFLOWS_SPMV_UNROLL,
// default 1) or two at a time (0); read when the operator is handed out.
bool spmv_unroll_from_env() {
    const char* e = std::getenv("TFLOWS_SPMV_UNROLL");
    return !(e && e[0] == '0');
}

// The symmetric shared-offset form: FP64 operator unless TFLOWS_SYM_SLOTS=0;
// FP32 operator when TFLOWS_SYM_SLOTS32=1 (its matrix at 128^3 sits in the
// MALL, where halving its values saves no time and the mirror reads cost
// some; at 256^3 the two forms measured the same).
bool sym_slots_from_env(bool fp32) {
    const char* e = std::getenv(fp32 ? "TFLOWS_SYM_SLOTS32" : "TFLOWS_SYM_SLOTS");
    return fp32 ? (e && e[0] == '1') : !(e && e[0] == '0');
}

// 1 in *flag when some pair of entries (a lower one and its mirror) differ
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pair_asymmetry(int64_t npairs, const int* __restrict__ pairs,
                                                           const T* __restrict__ a, int* __restrict__ flag) {
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < npairs; q += (int64_t)gridDim.x * kBlock)
        if (!(a[pairs[2 * q]] == a[pairs[2 * q + 1]])) *flag = 1;
}

// true when the values a (CSR order) equal their mirrors at every pair
template <typename T>
bool pairs_symmetric(const DevArray<int>& pairs, int64_t npairs, const T* a, DevArray<int>& flag, hipStream_t s) {
    if (npairs == 0) return true;
    flag.ensure(1);
    TF_CHECK(hipMemsetAsync(flag.p, 0, sizeof(int), s));
    hipLaunchKernelGGL((k_pair_asymmetry<T>), dim3((unsigned)std::min<int64_t>(2048, (npairs + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, s, npairs, (const int*)pairs.p, a, flag.p);
    TF_CHECK(hipGetLastError());
    int asym = 0;
    TF_CHECK(hipMemcpyAsync(&asym, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    TF_CHECK(hipStreamSynchronize(s));
    return asym == 0;
}
template bool pairs_symmetric<double>(const DevArray<int>&, int64_t, const double*, DevArray<int>&, hipStream_t);
template bool pairs_symmetric<float>(const DevArray<int>&, int64_t, const float*, DevArray<int>&, hipStream_t);

template <typename T>
__global__ __launch_bounds__(kBlock) void k_slot_values(int64_t count, const int* __restrict__ map,
                                                        const T* __restrict__ src, T* __restrict__ dst) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e < count) dst[e] = map[e] < 0 ? T(0) : src[map[e]];
}

Sell<double> System::sell64(hipStream_t s) {
    if (!sell64_built_) {
        std::vector<int> rp(n_ + 1), cl(nnz_), off, col, map;
        TF_CHECK(hipMemcpyAsync(rp.data(), rowptr_.p, sizeof(int) * (n_ + 1), hipMemcpyDeviceToHost, s));
        TF_CHECK(hipMemcpyAsync(cl.data(), col_.p, sizeof(int) * nnz_, hipMemcpyDeviceToHost, s));
        TF_CHECK(hipStreamSynchronize(s));
        std::vector<int> soff;
        std::vector<unsigned char> mask;
        sell64_sidx_.release();
        sell64_lbase_.release();
        sell64_pairs_.release();
        sell64_npairs_ = 0;
        if (shared_slots_from_env() && shared_slot_layout(n_, rp, cl, off, soff, mask, map)) {
            // one column offset per slot of a slice, a mask per row
            sell64_soff_.ensure(soff.size());
            sell64_mask_.ensure(mask.size());
            TF_CHECK(hipMemcpy(sell64_soff_.p, soff.data(), sizeof(int) * soff.size(), hipMemcpyHostToDevice));
            TF_CHECK(hipMemcpy(sell64_mask_.p, mask.data(), mask.size(), hipMemcpyHostToDevice));
            sell64_col_.release();
            sell64_dc_.release();
            // the symmetric form where the pattern allows it; the values are
            // checked at every change (below)
            SymSlots sym;
            if (sym_slots_from_env(false) && !sym64_refused_ &&
                symmetric_slot_layout(n_, rp, cl, off, soff, mask, map, sym)) {
                off.swap(sym.off);
                map.swap(sym.map);
                sell64_sidx_.ensure(sym.sidx.size());
                sell64_lbase_.ensure(sym.lbase.size());
                sell64_pairs_.ensure(sym.pairs.size());
                TF_CHECK(hipMemcpy(sell64_sidx_.p, sym.sidx.data(), sizeof(int) * sym.sidx.size(),
                                   hipMemcpyHostToDevice));
                TF_CHECK(hipMemcpy(sell64_lbase_.p, sym.lbase.data(), sizeof(SlotPair) * sym.lbase.size(),
                                   hipMemcpyHostToDevice));
                if (!sym.pairs.empty())
                    TF_CHECK(hipMemcpy(sell64_pairs_.p, sym.pairs.data(), sizeof(int) * sym.pairs.size(),
                                       hipMemcpyHostToDevice));
                sell64_npairs_ = (int64_t)sym.pairs.size() / 2;
            }
        } else {
            sell64_soff_.release();
            sell64_mask_.release();
            slice_layout(n_, rp, cl, false, off, col, map);
            sell64_col_.ensure(col.size());
            TF_CHECK(hipMemcpy(sell64_col_.p, col.data(), sizeof(int) * col.size(), hipMemcpyHostToDevice));
            // 16-bit column offsets where every one fits (TFLOWS_FP64_COL16=0
            // keeps 32-bit columns); padding slots (column -1) become kPad16
            std::vector<short> dc;
            if (col16_from_env("TFLOWS_FP64_COL16") && slot_offsets16(n_, off, col, false, dc)) {
                sell64_dc_.ensure(dc.size());
                TF_CHECK(hipMemcpy(sell64_dc_.p, dc.data(), sizeof(short) * dc.size(), hipMemcpyHostToDevice));
            } else {
                sell64_dc_.release();
            }
        }
        sell64_off_.ensure(off.size());
        sell64_map_.ensure(map.size());
        sell64_val_.ensure(map.size());
        TF_CHECK(hipMemcpy(sell64_off_.p, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice));
        TF_CHECK(hipMemcpy(sell64_map_.p, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice));
        sell64_entries_ = (int64_t)map.size();
        sell64_built_ = true;
        sell64_valid_ = false;
    }
    if (!sell64_valid_) {
        if (sell64_lbase_.p && !pairs_symmetric(sell64_pairs_, sell64_npairs_, (const double*)val_.p, symflag_, s)) {
            // values that differ from their mirrors: the full form from now on
            sym64_refused_ = true;
            sell64_built_ = false;
            return sell64(s);
        }
        hipLaunchKernelGGL((k_slot_values<double>), dim3((unsigned)((sell64_entries_ + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, s, sell64_entries_, (const int*)sell64_map_.p, (const double*)val_.p,
                           sell64_val_.p);
        TF_CHECK(hipGetLastError());
        sell64_valid_ = true;
    }
    return Sell<double>{n_,           sell64_off_.p,  sell64_col_.p,  sell64_val_.p, sell64_dc_.p,
                        sell64_soff_.p, sell64_mask_.p, sell64_sidx_.p, reinterpret_cast<const int2*>(sell64_lbase_.p)};
}

bool col16_from_env(const char* name) {
    const char* e = std::getenv(name);
    return !(e && e[0] == '0');
}

bool slot_offsets16(int n, const std::vector<int>& off, const std::vector<int>& col, bool pad_own,
                    std::vector<short>& dc) {
    dc.assign(col.size(), 0);
    if (col.empty()) return false;
    for (size_t sl = 0; sl + 1 < off.size(); ++sl)
        for (int q = off[sl]; q < off[sl + 1]; ++q) {
            const int row = (int)sl * kSlice + (q - off[sl]) % kSlice;
            if (row >= n) continue;                  // past the last row: never read
            if (!pad_own && col[q] < 0) {
                dc[q] = kPad16;
                continue;
            }
            const int d = col[q] - row;
            if (d < -32767 || d > 32767) return false;
            dc[q] = (short)d;
        }
    return true;
}

// Tile order of the SpMV kernels; TFLOWS_TILE_ORDER sets the default.
static int tile_order_from_env() {
    const char* e = std::getenv("TFLOWS_TILE_ORDER");
    return e ? (std::atoi(e) != 0) : 1;
}

void System::set_lanes(int lanes) {
    if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8)
        throw std::invalid_argument("lanes per row must be 1, 2, 4 or 8");
    lanes_ = lanes;
}

int max_blocks_from_env() {
    const char* e = std::getenv("TFLOWS_MAX_BLOCKS");
    int v = e ? std::atoi(e) : 2048;
    return v >= 8 ? v : 2048;
}


System::System(int n, int nonz, const int* acol, const int* arow, const int* adia, bool device_ptrs,
               hipStream_t s, bool allow_reorder)
    : n_(n), nnz_(nonz - 1), grid_(stream_grid(n, max_blocks_from_env())), lanes_(lanes_from_env()),
      strided_(tile_order_from_env()), op32_mode_(op_mode_from_env("TFLOWS_FP32_OP")),
      op64_mode_(op_mode_from_env("TFLOWS_FP64_OP")) {
    if (n <= 0) throw std::invalid_argument("N must be positive");
    if (nonz < n + 1) throw std::invalid_argument("NONZ must be at least N + 1 (diagonal entries + 1)");
    rowptr_.ensure(n_ + 1);
    col_.ensure(nnz_);
    dia_.ensure(n_);
    h_rowptr_.resize(n_ + 1);
    h_col_.resize(nnz_);
    h_dia_.resize(n_);
    if (device_ptrs) {
        TF_CHECK(hipMemcpyAsync(h_rowptr_.data(), acol, sizeof(int) * (n_ + 1), hipMemcpyDeviceToHost, s));
        TF_CHECK(hipMemcpyAsync(h_col_.data(), arow, sizeof(int) * nnz_, hipMemcpyDeviceToHost, s));
        TF_CHECK(hipMemcpyAsync(h_dia_.data(), adia, sizeof(int) * n_, hipMemcpyDeviceToHost, s));
        TF_CHECK(hipStreamSynchronize(s));
    } else {
        std::memcpy(h_rowptr_.data(), acol, sizeof(int) * (n_ + 1));
        std::memcpy(h_col_.data(), arow, sizeof(int) * nnz_);
        std::memcpy(h_dia_.data(), adia, sizeof(int) * n_);
    }
    // validate the layout on the host before any kernel sees it: every
    // index the kernels dereference must stay in bounds.
    if (h_rowptr_[0] != 1 || h_rowptr_[n_] != nnz_ + 1)
        throw std::invalid_argument("Acol must start at 1 and end at NONZ (= nnz + 1)");
    for (int i = 0; i < n_; ++i) {
        if (h_rowptr_[i + 1] < h_rowptr_[i]) throw std::invalid_argument("Acol is not non-decreasing");
        if (h_dia_[i] < h_rowptr_[i] || h_dia_[i] >= h_rowptr_[i + 1])
            throw std::invalid_argument("Adia(i) must point inside row i");
    }
    for (int i = 0; i < n_ + 1; ++i) h_rowptr_[i] -= 1;
    for (int i = 0; i < n_; ++i) h_dia_[i] -= 1;
    for (int j = 0; j < nnz_; ++j) {
        h_col_[j] -= 1;
        if (h_col_[j] < 0 || h_col_[j] >= n_) throw std::invalid_argument("Arow holds a column outside 1..N");
    }
    for (int i = 0; i < n_; ++i)
        if (h_col_[h_dia_[i]] != i) throw std::invalid_argument("Arow(Adia(i)) must equal i");
    build_storage(allow_reorder, s);
    val_.ensure(nnz_);
    TF_CHECK(hipMemsetAsync(val_.p, 0, sizeof(double) * nnz_, s));
    partials_.ensure(Sites<double>::elems(grid_) + 2 * (size_t)grid_);  // + site E
    state_.ensure(1);
    TF_CHECK(hipMemsetAsync(state_.p, 0, sizeof(SolverState), s));
    TF_CHECK(hipHostMalloc((void**)&mirror_, sizeof(HostMirror), hipHostMallocMapped | hipHostMallocCoherent));
    mirror_->done = 0;
    mirror_->iter = 0;
    TF_CHECK(hipStreamSynchronize(s));
}

System::~System() {
    if (mirror_) (void)hipHostFree(mirror_);
    if (mirror32_) (void)hipHostFree(mirror32_);
}

void System::set_values(const double* aval, bool device_ptr, hipStream_t s) {
    TF_CHECK(hipMemcpyAsync(reordered_ ? val_orig_.p : val_.p, aval, sizeof(double) * nnz_,
                            device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    if (reordered_) permute_values(s);
    levels_.vals64 = levels_.vals32 = false;
    val32_valid_ = false;
    diff32_valid_ = false;
    sell_valid_ = false;
    sell64_valid_ = false;
    transposed_valid_ = false;
    if (!device_ptr) TF_CHECK(hipStreamSynchronize(s));
}

void System::set_values_range(int64_t offset, int64_t count, const double* aval, bool device_ptr,
                              hipStream_t s) {
    if (offset < 0 || count < 0 || offset + count > nnz_) throw std::invalid_argument("value range out of bounds");
    if (reordered_) throw std::logic_error("partial value updates need the caller's storage order");
    TF_CHECK(hipMemcpyAsync(val_.p + offset, aval, sizeof(double) * count,
                            device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    levels_.vals64 = levels_.vals32 = false;
    val32_valid_ = false;
    diff32_valid_ = false;
    sell_valid_ = false;
    sell64_valid_ = false;
    transposed_valid_ = false;
    if (!device_ptr) TF_CHECK(hipStreamSynchronize(s));
}

const float* System::values32(hipStream_t s) {
    if (!val32_valid_) {
        val32_.ensure(nnz_);
        launch_f64_to_f32(nnz_, val_.p, val32_.p, s);
        TF_CHECK(hipGetLastError());
        val32_valid_ = true;
    }
    return val32_.p;
}

// Level analysis on the host (once per pattern).  Row i of the forward
// sweep depends on the rows of its lower entries (those before Adia(i)),
// row i of the backward sweep on the rows of its upper entries; the level
// of a row is the length of the longest dependency chain that ends in it,
// so the rows of one level are independent and a level is one parallel
// step of a sweep.  Levels are found by a topological walk of each
// dependency graph, which holds in any numbering of the rows (a batch
// stores its renumbered systems permuted, with each row's entries in their
// original order, so "lower" there is no longer "column < row").
static int dependency_levels(int n, const std::vector<int>& rowptr, const std::vector<int>& col,
                             const std::vector<int>& dia, bool lower, std::vector<int>& lev) {
    auto first = [&](int i) { return lower ? rowptr[i] : dia[i] + 1; };
    auto last = [&](int i) { return lower ? dia[i] : rowptr[i + 1]; };
    // dependents of each row (transpose of the dependency lists)
    std::vector<int> cnt(n + 1, 0), indeg(n);
    for (int i = 0; i < n; ++i) {
        indeg[i] = last(i) - first(i);
        for (int j = first(i); j < last(i); ++j) cnt[col[j] + 1]++;
    }
    for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<int> dep(cnt[n]), fill(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int j = first(i); j < last(i); ++j) dep[fill[col[j]]++] = i;
    std::vector<int> queue;
    queue.reserve(n);
    lev.assign(n, 0);
    for (int i = 0; i < n; ++i)
        if (indeg[i] == 0) queue.push_back(i);
    int nlev = 0;
    for (size_t h = 0; h < queue.size(); ++h) {
        const int v = queue[h];
        nlev = std::max(nlev, lev[v] + 1);
        for (int k = cnt[v]; k < cnt[v + 1]; ++k) {
            const int d = dep[k];
            lev[d] = std::max(lev[d], lev[v] + 1);
            if (--indeg[d] == 0) queue.push_back(d);
        }
    }
    if ((int)queue.size() != n)
        throw std::invalid_argument("the triangular dependencies of the pattern form a cycle");
    return nlev;
}

const Levels& System::levels() {
    if (levels_ready_) return levels_;
    std::vector<int> lev;
    auto bucket = [&](int nlev, std::vector<int>& ptr, DevArray<int>& rows) {
        ptr.assign(nlev + 1, 0);
        for (int i = 0; i < n_; ++i) ptr[lev[i] + 1]++;
        for (int l = 0; l < nlev; ++l) ptr[l + 1] += ptr[l];
        std::vector<int> out(n_), fill(ptr.begin(), ptr.end() - 1);
        // levels are found on the caller's pattern; the lists hold the rows
        // in the internal numbering the kernels use
        for (int i = 0; i < n_; ++i) out[fill[lev[i]]++] = reordered_ ? iperm_h_[i] : i;
        rows.ensure(n_);
        TF_CHECK(hipMemcpy(rows.p, out.data(), sizeof(int) * n_, hipMemcpyHostToDevice));
        return out;
    };
    // a pattern whose lower columns lie below each row and upper ones above
    // (every TopolM pattern in the caller's numbering) has its levels in
    // one pass in index order; others need the topological walk
    int nf = sorted_pattern_levels(n_, h_rowptr_.data(), h_col_.data(), h_dia_.data(), true, lev);
    const bool sorted = nf >= 0;
    if (!sorted) nf = dependency_levels(n_, h_rowptr_, h_col_, h_dia_, true, lev);
    levels_.nfwd = nf;
    levels_.fwd_rows_h = bucket(levels_.nfwd, levels_.fwd_ptr, levels_.fwd_rows);
    std::vector<int> flev;
    flev.swap(lev);
    int nb = sorted ? sorted_pattern_levels(n_, h_rowptr_.data(), h_col_.data(), h_dia_.data(), false, lev) : -1;
    if (nb < 0) nb = dependency_levels(n_, h_rowptr_, h_col_, h_dia_, false, lev);
    levels_.nbwd = nb;
    levels_.bwd_rows_h = bucket(levels_.nbwd, levels_.bwd_ptr, levels_.bwd_rows);
    build_dic_plan(flev, sorted);
    levels_ready_ = true;
    return levels_;
}

// The level-major triangles (Levels::plan) from the stored pattern.
void System::build_dic_plan(const std::vector<int>& flev, bool sorted) {
    Levels& L = levels_;
    L.mode = dic_sweep_mode();
    if (L.mode == 0) return;
    std::vector<int> rp(n_ + 1), cl(nnz_), dg(n_);
    TF_CHECK(hipMemcpy(rp.data(), rowptr_.p, sizeof(int) * (n_ + 1), hipMemcpyDeviceToHost));
    TF_CHECK(hipMemcpy(cl.data(), col_.p, sizeof(int) * nnz_, hipMemcpyDeviceToHost));
    TF_CHECK(hipMemcpy(dg.data(), dia_.p, sizeof(int) * n_, hipMemcpyDeviceToHost));
    auto up = [](DevArray<int>& d, const std::vector<int>& h) {
        d.ensure(h.size());
        if (!h.empty()) TF_CHECK(hipMemcpy(d.p, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    };
    auto arm_waits = [&] {
        L.fault.ensure(1);
        TF_CHECK(hipMemset(L.fault.p, 0, sizeof(int)));
        int dev = 0, khz = 0;
        TF_CHECK(hipGetDevice(&dev));
        TF_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        L.wait_ticks = (uint64_t)(khz > 0 ? khz : 100000) * dic_wait_ms();
    };
    if (L.mode == 3) {
        TilePlan tp;
        if (sorted && build_tile_plan(n_, h_rowptr_.data(), h_col_.data(), h_dia_.data(), flev.data(),
                                      reordered_ ? iperm_h_.data() : nullptr, rp.data(), cl.data(), tp)) {
            std::vector<int> ws, wt;
            L.twaves = std::min(tile_sweep_waves(tp.slots), tp.ncols);
            tile_schedule(tp, L.twaves, ws, wt);
            up(L.wstart, ws);
            up(L.wtiles, wt);
            up(L.dmap, tp.dmap);
            up(L.order, tp.order);
            up(L.tph, tp.tph);
            up(L.lcol, tp.lcol);
            up(L.lmap, tp.lmap);
            up(L.ucol, tp.ucol);
            up(L.umap, tp.umap);
            L.npos = n_;
            L.ntiles = tp.ntiles;
            L.tslots = tp.slots;
            L.nlow = L.nup = (size_t)tp.slots * n_;
            arm_waits();
            L.plan = true;
            L.vals64 = L.vals32 = false;
            return;
        }
        L.mode = 2;   // the per-row one-launch sweeps over the level-major copy
    }
    // positions: forward level l occupies [pos_ptr[l], pos_ptr[l+1]), its
    // start a multiple of the wavefront size; the padding positions (row
    // -1) have no entries
    L.pos_ptr.assign(L.nfwd + 1, 0);
    for (int l = 0; l < L.nfwd; ++l) {
        const int len = L.fwd_ptr[l + 1] - L.fwd_ptr[l];
        L.pos_ptr[l + 1] = L.pos_ptr[l] + (len + kWave - 1) / kWave * kWave;
    }
    const int np = L.pos_ptr[L.nfwd];
    std::vector<int> order(np, -1), pos(n_), plev(np, 0);
    for (int l = 0; l < L.nfwd; ++l)
        for (int t = L.fwd_ptr[l]; t < L.fwd_ptr[l + 1]; ++t) {
            const int k = L.pos_ptr[l] + (t - L.fwd_ptr[l]);
            order[k] = L.fwd_rows_h[t];
            pos[order[k]] = k;
        }
    for (int l = 0; l < L.nfwd; ++l)
        for (int k = L.pos_ptr[l]; k < L.pos_ptr[l + 1]; ++k) plev[k] = l;
    std::vector<int> lptr(np + 1, 0), uptr(np + 1, 0);
    for (int k = 0; k < np; ++k) {
        const int i = order[k];
        lptr[k + 1] = lptr[k] + (i < 0 ? 0 : dg[i] - rp[i]);
        uptr[k + 1] = uptr[k] + (i < 0 ? 0 : rp[i + 1] - dg[i] - 1);
    }
    std::vector<int> lcol(lptr[np]), lmap(lptr[np]), ucol(uptr[np]), umap(uptr[np]), bpos(n_);
    // one launch per triangle needs every wavefront's rows free of mutual
    // dependencies in both sweeps: rows of one forward level are (their
    // lower entries lie in earlier levels by construction); for the
    // backward sweep, run over the positions in reverse, the upper entries
    // must lie in later forward levels, which holds for a structurally
    // symmetric pattern (TopolM.f90:65-66)
    bool upper_later = true;
    for (int k = 0; k < np; ++k) {
        const int i = order[k];
        if (i < 0) continue;
        int o = lptr[k];
        for (int j = rp[i]; j < dg[i]; ++j, ++o) {
            lcol[o] = pos[cl[j]];
            lmap[o] = j;
        }
        o = uptr[k];
        for (int j = dg[i] + 1; j < rp[i + 1]; ++j, ++o) {
            ucol[o] = pos[cl[j]];
            umap[o] = j;
            upper_later = upper_later && plev[ucol[o]] > plev[k];
        }
    }
    for (int t = 0; t < n_; ++t) bpos[t] = pos[L.bwd_rows_h[t]];
    up(L.order, order);
    up(L.lptr, lptr);
    up(L.lcol, lcol);
    up(L.lmap, lmap);
    up(L.uptr, uptr);
    up(L.ucol, ucol);
    up(L.umap, umap);
    up(L.bwd_pos, bpos);
    L.npos = np;
    L.nlow = lptr[np];
    L.nup = uptr[np];
    if (L.mode == 2 && !upper_later) L.mode = 1;
    if (L.mode == 2) arm_waits();
    L.plan = true;
    L.vals64 = L.vals32 = false;
}

int dic_sweep_mode() {
    const char* e = std::getenv("TFLOWS_DIC_SWEEP");
    if (!e) return 3;
    const int m = std::atoi(e);
    return m < 0 ? 0 : (m > 3 ? 3 : m);
}

uint64_t dic_wait_ms() {
    const char* e = std::getenv("TFLOWS_DIC_WAIT_MS");
    const long v = e ? std::atol(e) : 2000;
    return v < 0 ? 0 : (uint64_t)v;
}

void System::check_dic(hipStream_t s) const {
    if (!levels_ready_ || levels_.mode < 2 || !levels_.plan) return;
    int fault = 0;
    TF_CHECK(hipMemcpyAsync(&fault, levels_.fault.p, sizeof(int), hipMemcpyDeviceToHost, s));
    TF_CHECK(hipStreamSynchronize(s));
    if (fault)
        throw DicWaitFault("DIC sweep: a row waited on its dependencies beyond TFLOWS_DIC_WAIT_MS; the "
                           "preconditioned vector is invalid");
}

void System::latch_dic_form1() {
    levels();
    if (!levels_.plan || levels_.mode < 2) return;
    levels_.mode = levels_.mode == 3 ? 0 : 1;
    std::fprintf(stderr,
                 "libtflows_krylov: a one-launch DIC sweep gave up waiting (TFLOWS_DIC_WAIT_MS); this system "
                 "now sweeps one level per launch\n");
    std::fflush(stderr);
}

// The values of the level-major (or tiled) triangles, FP64 and, when fp32,
// FP32 rounded from them, whenever the matrix has changed since the last
// gather; the work vectors of the one-launch sweeps.
void gather_dic_values(const Levels& L, const double* val, bool fp32, hipStream_t s) {
    if (!L.plan) return;
    // the slot arrays hold kTileWave spare slots past the positions (the
    // tiled sweeps' lanes without a row store there)
    const size_t slots = (size_t)L.npos + kTileWave;
    if (!L.vals64) {
        L.lval.ensure(L.nlow);
        L.uval.ensure(L.nup);
        L.up.ensure(L.npos);
        L.zp.ensure(slots);
        L.zb.ensure(slots);
        L.dp.ensure(slots);
        launch_entry_gather(L.nlow, L.lmap.p, val, L.lval.p, s);
        launch_entry_gather(L.nup, L.umap.p, val, L.uval.p, s);
        L.vals64 = true;
        L.vals32 = false;
    }
    if (fp32 && !L.vals32) {
        L.lval32.ensure(L.nlow);
        L.uval32.ensure(L.nup);
        L.up32.ensure(L.npos);
        L.zp32.ensure(slots);
        L.zb32.ensure(slots);
        L.dp32.ensure(slots);
        launch_f64_to_f32(L.nlow, L.lval.p, L.lval32.p, s);
        launch_f64_to_f32(L.nup, L.uval.p, L.uval32.p, s);
        L.vals32 = true;
    }
}

void System::prepare_dic(bool fp32, hipStream_t s) {
    const Levels& L = levels();
    if (!L.plan) return;
    gather_dic_values(L, val_.p, fp32, s);
    if (L.mode >= 2) {   // a solve starts: every forward slot empty
        launch_flow_reset(L.npos, L.zp.p, s);
        if (fp32) launch_flow_reset(L.npos, L.zp32.p, s);
    }
}

// The DIC pivots of A into D (internal numbering): the tiled one-launch
// form in mode 3 (over the zb slots, emptied first), else one launch per
// forward level.
static TileSweep tile_sweep(const Levels& L) {
    TileSweep ts;
    ts.n = L.npos;
    ts.slots = L.tslots;
    ts.waves = L.twaves;
    ts.wstart = L.wstart.p;
    ts.wtiles = L.wtiles.p;
    ts.tph = L.tph.p;
    ts.order = L.order.p;
    ts.fault = L.fault.p;
    ts.wait_ticks = L.wait_ticks;
    return ts;
}

void dic_pivots_levels(const Levels& L, const Csr<double>& A, double* D, hipStream_t s) {
    if (L.plan && L.mode == 3) {
        gather_dic_values(L, A.val, false, s);
        L.dgp.ensure(L.npos);
        launch_entry_gather(L.npos, L.dmap.p, A.val, L.dgp.p, s);
        launch_flow_reset(L.npos, L.zb.p, s);
        launch_tile_pivots(tile_sweep(L), L.dgp.p, L.lcol.p, L.lval.p, D, L.zb.p, s);
        return;
    }
    for (int l = 0; l < L.nfwd; ++l)
        launch_dic_setup_level(A, L.fwd_rows.p + L.fwd_ptr[l], L.fwd_ptr[l + 1] - L.fwd_ptr[l], D, s);
}

void System::dic_setup(double* D, hipStream_t s) { dic_pivots_levels(levels(), csr(), D, s); }

template <typename T>
void dic_apply_levels(const Levels& L, const Csr<T>& A, const T* D, const T* u, T* z, const SolverState* st,
                      hipStream_t s) {
    const bool f32 = sizeof(T) == sizeof(float);
    if (L.plan && L.mode == 3 && (f32 ? L.vals32 : L.vals64)) {
        T* zp = f32 ? (T*)L.zp32.p : (T*)L.zp.p;
        T* zb = f32 ? (T*)L.zb32.p : (T*)L.zb.p;
        T* dp = f32 ? (T*)L.dp32.p : (T*)L.dp.p;
        const T* lv = f32 ? (const T*)L.lval32.p : (const T*)L.lval.p;
        const T* uv = f32 ? (const T*)L.uval32.p : (const T*)L.uval.p;
        const TileSweep ts = tile_sweep(L);
        launch_tile_fwd(ts, u, D, L.lcol.p, lv, dp, zp, zb, st, s);
        launch_tile_bwd(ts, z, L.ucol.p, uv, (const T*)dp, zp, zb, st, s);
        return;
    }
    if (L.plan && L.mode != 0 && L.mode != 3 && (f32 ? L.vals32 : L.vals64)) {
        T* up = f32 ? (T*)L.up32.p : (T*)L.up.p;
        T* zp = f32 ? (T*)L.zp32.p : (T*)L.zp.p;
        T* zb = f32 ? (T*)L.zb32.p : (T*)L.zb.p;
        T* dp = f32 ? (T*)L.dp32.p : (T*)L.dp.p;
        const T* lv = f32 ? (const T*)L.lval32.p : (const T*)L.lval.p;
        const T* uv = f32 ? (const T*)L.uval32.p : (const T*)L.uval.p;
        const int np = L.npos;
        if (L.mode == 2) {
            launch_flow_fwd(np, L.order.p, u, D, L.lptr.p, L.lcol.p, lv, dp, zp, zb, L.fault.p, L.wait_ticks, st, s);
            launch_flow_bwd(np, L.order.p, z, L.uptr.p, L.ucol.p, uv, (const T*)dp, zp, zb, L.fault.p, L.wait_ticks,
                            st, s);
            return;
        }
        launch_plan_in(np, L.order.p, u, D, up, dp, st, s);
        for (int l = 0; l < L.nfwd; ++l)
            launch_plan_fwd(L.pos_ptr[l], L.pos_ptr[l + 1], L.lptr.p, L.lcol.p, lv, (const T*)dp, (const T*)up, zp,
                            st, s);
        for (int l = 0; l < L.nbwd; ++l)
            launch_plan_bwd(L.bwd_pos.p + L.bwd_ptr[l], L.bwd_ptr[l + 1] - L.bwd_ptr[l], L.uptr.p, L.ucol.p, uv,
                            (const T*)dp, zp, st, s);
        launch_plan_out(np, L.order.p, (const T*)zp, z, st, s);
        return;
    }
    for (int l = 0; l < L.nfwd; ++l)
        launch_dic_fwd_level(A, L.fwd_rows.p + L.fwd_ptr[l], L.fwd_ptr[l + 1] - L.fwd_ptr[l], D, u, z, st, s);
    for (int l = 0; l < L.nbwd; ++l)
        launch_dic_bwd_level(A, L.bwd_rows.p + L.bwd_ptr[l], L.bwd_ptr[l + 1] - L.bwd_ptr[l], D, z, st, s);
}

template <typename T>
void System::dic_apply(const Csr<T>& A, const T* D, const T* u, T* z, const SolverState* st, hipStream_t s) {
    dic_apply_levels(levels(), A, D, u, z, st, s);
}

void System::prec1(int prec, double* d_ext, hipStream_t s) {
    // D only; the ||b|| partial sums of k_setup run over a zero slot of the
    // workspace and are not used.
    ws_.ensure((size_t)V_COUNT * n_);
    TF_CHECK(hipMemsetAsync(ws_.p, 0, sizeof(double) * n_, s));
    double* d = reordered_ ? ext_buffer(0) : d_ext;
    if (prec == 2) {
        dic_setup(d, s);
        if (dic_form() >= 2) {
            try {
                check_dic(s);
            } catch (const DicWaitFault&) {
                latch_dic_form1();
                dic_setup(d, s);
            }
        }
    }
    launch_setup(ctx(), csr(), prec, prec != 2, d, (const double*)ws_.p, sites(), s);
    TF_CHECK(hipGetLastError());
    if (reordered_) to_external(d, d_ext, s);
}

double* System::ext_buffer(int k) {
    extbuf_.ensure(3 * (size_t)n_);
    return extbuf_.p + (size_t)k * n_;
}

Sites<double> System::sites() { return Sites<double>::carve(partials_.p, grid_); }

void System::spmv(const double* x, double* y, hipStream_t s) {
    with_op64(s, [&](const auto& A) {
        if (!reordered_) {
            launch_spmv(A, x, y, grid_, s, ctx());
            TF_CHECK(hipGetLastError());
            return;
        }
        double *xi = ext_buffer(0), *yi = ext_buffer(1);
        to_internal(x, xi, s);
        launch_spmv(A, (const double*)xi, yi, grid_, s, ctx());
        TF_CHECK(hipGetLastError());
        to_external((const double*)yi, y, s);
    });
}
void System::spmv32(const float* x, float* y, hipStream_t s) {
    with_op32(s, [&](const auto& A) {
        if (!reordered_) {
            launch_spmv(A, x, y, grid_, s, ctx());
            TF_CHECK(hipGetLastError());
            return;
        }
        extbuf32_.ensure(2 * (size_t)n_);
        float *xi = extbuf32_.p, *yi = extbuf32_.p + n_;
        to_internal(x, xi, s);
        launch_spmv(A, (const float*)xi, yi, grid_, s, ctx());
        TF_CHECK(hipGetLastError());
        to_external((const float*)yi, y, s);
    });
}
void System::resid(const double* x, double* r, hipStream_t s) {
    if (!reordered_) {
        launch_resid(csr(), x, r, grid_, s, ctx());
        TF_CHECK(hipGetLastError());
        return;
    }
    double *xi = ext_buffer(0), *ri = ext_buffer(1);
    to_internal(x, xi, s);
    to_internal((const double*)r, ri, s);
    launch_resid(csr(), (const double*)xi, ri, grid_, s, ctx());
    TF_CHECK(hipGetLastError());
    to_external((const double*)ri, r, s);
}

// One iteration k of `method`, enqueued (CG.f90:113-259, BiCG.f90:100-251,
// CGS.f90:112-334).
// Under a batch context st / hm are the per-system arrays, k is relative
// to c.kbase, and the DIC sweeps run without the early return (a finished
// system's sweeps only touch its workspace).
template <typename T, typename Op>
void enqueue_iteration(const BatchCtx& c, int method, bool jac, int k, const Op& Aop, const Csr<T>& A,
                       const KrylovVectors<T>& w, const Sites<T>& S, SolverState* st, HostMirror* hm,
                       const Levels* L, hipStream_t s) {
    const int n = A.n;
    const bool single = c.blk_sys == nullptr;
    const SolverState* dst = single ? st : nullptr;
    switch (method) {
        case 0: {  // CG
            if (single && jac && cg_form() == 2) {  // p0 = p, p1 = p2, z = u1
                launch_cgf_spmv(false, Aop, (const T*)w.u1, w.p, w.p2, w.q, S, (const double*)nullptr, st, hm, s);
                launch_cgf_update(n, w.x, w.r1, w.u1, (const T*)w.p, (const T*)w.p2, (const T*)w.q,
                                  (const T*)w.D, S, st, s);
                break;
            }
            const bool carry = carries_x(c, method, jac);
            if (single && jac && cg_form() == 3) {
                launch_cgj(false, c.strided != 0, Aop, n, w.x, w.r1, w.u1, w.p, w.q, (const T*)w.D, S, (const double*)nullptr, st,
                           hm, s, carry);
                break;
            }
            if (carry) {
                launch_cg_p_carry(c, k, n, (const T*)w.r1, (const T*)w.D, w.p, w.x, S, st, hm, s);
                launch_cg_spmv(c, Aop, (const T*)w.p, w.q, S, st, s);
                launch_cg_update_carry(c, k, n, w.r1, (const T*)w.q, (const T*)w.D, S, st, s);
                break;
            }
            if (!jac) {  // stopping test, z = M^-1 r into q, rho = (r, z)
                launch_check(c, k, S, st, hm, s);
                dic_apply_levels(*L, A, (const T*)w.D, (const T*)w.r1, w.q, dst, s);
                launch_rho(c, n, (const T*)w.r1, (const T*)w.q, S, st, s);
            }
            launch_cg_p(c, jac, k, n, (const T*)w.r1, (const T*)w.D, (const T*)w.q, w.p, S, st, hm, s);
            launch_cg_spmv(c, Aop, (const T*)w.p, w.q, S, st, s);
            launch_cg_update(c, jac, k, n, w.x, w.r1, (const T*)w.p, (const T*)w.q, (const T*)w.D, S, st, s);
            break;
        }
        case 1: {  // BiCG (reference semantics)
            if (!jac) {
                launch_check(c, k, S, st, hm, s);
                dic_apply_levels(*L, A, (const T*)w.D, (const T*)w.r1, w.q, dst, s);
                dic_apply_levels(*L, A, (const T*)w.D, (const T*)w.r2, w.q2, dst, s);
                launch_rho(c, n, (const T*)w.q, (const T*)w.r2, S, st, s);
            }
            launch_bicg_p(c, jac, k, n, (const T*)w.r1, (const T*)w.r2, (const T*)w.D, (const T*)w.q, (const T*)w.q2,
                          w.p, w.p2, S, st, hm, s);
            launch_bicg_spmv(c, Aop, (const T*)w.p, (const T*)w.p2, w.q, w.q2, S, st, s);
            launch_bicg_update(c, jac, k, n, w.x, w.r1, w.r2, (const T*)w.p, (const T*)w.q, (const T*)w.q2,
                               (const T*)w.D, S, st, s);
            break;
        }
        default: {  // CGS: u1, u2, p2 = M^-1 u2, v2 = A p2, q1 (q), p1 (p) = M^-1 (u1+q1)
            launch_cgs_u(c, jac, k, n, (const T*)w.r1, (const T*)w.q, w.u1, w.u2, (const T*)w.D, w.p2, S, st, hm, s);
            if (!jac) dic_apply_levels(*L, A, (const T*)w.D, (const T*)w.u2, w.p2, dst, s);
            launch_cgs_spmv_v(c, Aop, (const T*)w.p2, w.v2, (const T*)w.r2, S, st, s);
            launch_cgs_q(c, jac, k, n, (const T*)w.u1, (const T*)w.v2, w.q, (const T*)w.D, w.p, S, st, s);
            if (!jac) dic_apply_levels(*L, A, (const T*)w.D, (const T*)w.p, w.p, dst, s);
            launch_cgs_spmv_r(c, k, Aop, (const T*)w.p, w.x, w.r1, (const T*)w.r2, S, st, s);
            break;
        }
    }
}

// Keep the stream supplied with iterations while the device runs them.
// Iterations go out in chunks that double in size (so short solves stop
// early and long ones keep the queue deep); after queueing chunk j the host
// waits for the event of chunk j-1 and reads the `done` flag the device
// published, so one chunk is always queued behind the running one.  Kernels
// queued past convergence return at once.  After iteration `cap` a last
// stopping test (k = cap + 1) closes the loop.
Replay::~Replay() {
    for (auto& e : cache) (void)hipGraphExecDestroy(e.exec);
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

void Replay::ensure() {
    if (stream) return;
    TF_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    for (auto& e : ev) TF_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    kbase.ensure(1);
}

// TFLOWS_DIC_GRAPH=0 issues DIC iterations from the host instead.
bool dic_graph_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("TFLOWS_DIC_GRAPH");
        return !e || std::atoi(e) != 0;
    }();
    return on;
}

static const void* op_key(const Csr<double>& a) { return a.val; }
static const void* op_key(const Csr<float>& a) { return a.val; }
static const void* op_key(const DiffCsr32& a) { return a.val; }
static const void* op_key(const DiffSell32& a) { return a.val; }
static const void* op_key(const Sell<double>& a) { return a.val; }

// One DIC iteration of one system replayed from a graph until the system
// stops (or `cap` iterations were issued; then the k = cap + 1 test).
template <typename T, typename Op>
static void pace_graph(const BatchCtx& c, int method, int cap, const Op& Aop, const Csr<T>& A,
                       const KrylovVectors<T>& w, const Sites<T>& S, SolverState* st, HostMirror* hm,
                       const Levels* L, hipStream_t s, Replay& rp) {
    rp.ensure();
    hipStream_t g = rp.stream;
    TF_CHECK(hipEventRecord(rp.ev[2], s));   // the replayer's stream follows the caller's work
    TF_CHECK(hipStreamWaitEvent(g, rp.ev[2], 0));
    // the graph is bound to every buffer it was captured with
    std::string key;
    auto add = [&](const void* p) { key.append(reinterpret_cast<const char*>(&p), sizeof(p)); };
    key.append(reinterpret_cast<const char*>(&method), sizeof(method));
    key.append(reinterpret_cast<const char*>(&c.lanes), sizeof(c.lanes));
    key.append(reinterpret_cast<const char*>(&c.strided), sizeof(c.strided));
    key.append(reinterpret_cast<const char*>(&L->mode), sizeof(L->mode));   // the sweep form captured
    add(op_key(Aop));
    add(A.rowptr), add(A.col), add(A.dia), add(A.val);
    add(w.x), add(w.r1), add(w.D), add(w.p), add(w.q), add(w.r2), add(w.p2), add(w.q2), add(w.u1), add(w.u2);
    add(w.v2), add(S.A), add(S.B), add(S.C), add(S.D), add(st), add(hm);
    hipGraphExec_t exec = nullptr;
    for (auto& e : rp.cache)
        if (e.key == key) exec = e.exec;
    if (!exec) {
        if (rp.cache.size() >= 8) {   // keep a few buffer sets
            (void)hipGraphExecDestroy(rp.cache.front().exec);
            rp.cache.erase(rp.cache.begin());
        }
        BatchCtx ck = c;
        ck.kbase = rp.kbase.p;
        hipGraph_t gr = nullptr;
        TF_CHECK(hipStreamBeginCapture(g, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_iteration<T, Op>(ck, method, false, 1, Aop, A, w, S, st, hm, L, g);
            launch_advance(rp.kbase.p, 1, g);
        } catch (...) {
            (void)hipStreamEndCapture(g, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            throw;
        }
        TF_CHECK(hipStreamEndCapture(g, &gr));
        const hipError_t e = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        TF_CHECK(e);
        rp.cache.push_back({key, exec});
    }
    launch_set_int(rp.kbase.p, 0, g);
    bool pending[2] = {false, false};
    int slot = 0, issued = 0;
    bool done = false;
    while (issued < cap && !done) {
        TF_CHECK(hipGraphLaunch(exec, g));
        ++issued;
        TF_CHECK(hipEventRecord(rp.ev[slot], g));
        pending[slot] = true;
        const int other = slot ^ 1;
        if (pending[other]) {
            TF_CHECK(hipEventSynchronize(rp.ev[other]));
            pending[other] = false;
            done = hm->done != 0;
        }
        slot = other;
    }
    if (!done) launch_check(c, std::max(cap, 0) + 1, S, st, hm, g);
    TF_CHECK(hipGetLastError());
    TF_CHECK(hipStreamSynchronize(g));
}

template <typename T, typename Op>
void pace_iterations(const BatchCtx& c, int method, bool jac, int cap, const Op& Aop, const Csr<T>& A,
                     const KrylovVectors<T>& w, const Sites<T>& S, SolverState* st, HostMirror* hm,
                     const Levels* L, hipStream_t s, Replay* rp) {
    if (!jac && rp && !c.blk_sys && dic_graph_enabled()) {
        pace_graph<T, Op>(c, method, cap, Aop, A, w, S, st, hm, L, s, *rp);
        return;
    }
    hipEvent_t ev[2];
    TF_CHECK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    TF_CHECK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    bool pending[2] = {false, false};
    const int max_chunk = jac ? 64 : 2;
    int chunk = jac ? 4 : 1, slot = 0, issued = 0;
    bool done = false;
    while (issued < cap && !done) {
        const int k = std::min(chunk, cap - issued);
        for (int i = 0; i < k; ++i) {
            ++issued;
            enqueue_iteration<T, Op>(c, method, jac, issued, Aop, A, w, S, st, hm, L, s);
        }
        TF_CHECK(hipGetLastError());
        TF_CHECK(hipEventRecord(ev[slot], s));
        pending[slot] = true;
        const int other = slot ^ 1;
        if (pending[other]) {
            TF_CHECK(hipEventSynchronize(ev[other]));
            pending[other] = false;
            done = hm->done != 0;
        }
        slot = other;
        chunk = std::min(max_chunk, chunk * 2);
    }
    if (!done) {
        if (method == 0 && jac && fused_cg_enabled())
            launch_cgf_check(false, S, (const double*)nullptr, st, hm, s);
        else
            launch_check(c, std::max(cap, 0) + 1, S, st, hm, s);
    }
    // the x update of the last iteration, which no p kernel applied
    if (carries_x(c, method, jac)) launch_cg_xfinish(A.n, w.x, (const T*)w.p, st, s);
    TF_CHECK(hipGetLastError());
    TF_CHECK(hipStreamSynchronize(s));
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
}

// Set-up and initial residual of a solve in precision T (one system, or a
// batch under its context).  Aop is the operator of the SpMVs; A the CSR
// the DIC sweeps read.  With d_ready the caller has filled w.D; with
// zero_x the initial guess is 0 and the initial residual is b without a
// matrix pass.  The caller has written the initial SolverState(s) (tol,
// cap, stale error).
template <typename T, typename Op>
void start_solve(const BatchCtx& c, int method, int prec, const Op& Aop, const Csr<T>& A,
                 const KrylovVectors<T>& w, const Sites<T>& S, SolverState* st, HostMirror* hm, const Levels* L,
                 hipStream_t s, bool d_ready, bool zero_x) {
    (void)Aop;
    const bool jac = prec != 2;
    if (prec == 2 && !d_ready) {
        if constexpr (std::is_same_v<T, double>)
            dic_pivots_levels(*L, A, w.D, s);
        else
            for (int l = 0; l < L->nfwd; ++l)
                launch_dic_setup_level(A, L->fwd_rows.p + L->fwd_ptr[l], L->fwd_ptr[l + 1] - L->fwd_ptr[l], w.D,
                                       s);
    }
    launch_setup(c, A, prec, jac && !d_ready, w.D, (const T*)w.r1, S, s);
    int mode = RHO_NONE_;
    T* r2 = nullptr;
    const bool fused = c.blk_sys == nullptr && method == 0 && jac && fused_cg_enabled();
    if (method == 0) mode = (jac && !fused) ? RHO_JAC_ : RHO_NONE_;
    if (method == 1) { mode = jac ? RHO_BICG_ : RHO_NONE_; r2 = w.r2; }
    if (method == 2) { mode = RHO_RR_; r2 = w.r2; }
    launch_resid_init(mode, zero_x, c, A, (const T*)w.x, w.r1, r2, (const T*)w.D, S, st, hm, s);
    if (fused) launch_cgf_init_z(A.n, (const T*)w.r1, (const T*)w.D, w.u1, S, st, s);
    TF_CHECK(hipGetLastError());
}

// One complete solve of one system: set-up, initial residual, iterations.
template <typename T, typename Op>
void run_solve(const BatchCtx& c, int method, int prec, int cap, const Op& Aop, const Csr<T>& A,
               const KrylovVectors<T>& w, const Sites<T>& S, SolverState* st, HostMirror* hm, const Levels* L,
               hipStream_t s, bool d_ready, bool zero_x, Replay* rp) {
    hm->done = 0;
    hm->iter = 0;
    start_solve<T, Op>(c, method, prec, Aop, A, w, S, st, hm, L, s, d_ready, zero_x);
    pace_iterations<T, Op>(c, method, prec != 2, cap, Aop, A, w, S, st, hm, L, s, rp);
}

SolveResult System::solve(int method, int prec, double* x, double* r1, int cap, double tol, double stale_error,
                          hipStream_t s) {
    if (prec != 2) return solve_ext(method, prec, x, r1, cap, tol, stale_error, s);
    double* vecs[2] = {x, r1};
    return dic_retry(vecs, 2, s, [&] { return solve_ext(method, prec, x, r1, cap, tol, stale_error, s); });
}

SolveResult System::solve_ext(int method, int prec, double* x, double* r1, int cap, double tol,
                              double stale_error, hipStream_t s) {
    if (!reordered_) return solve_impl(method, prec, x, r1, cap, tol, stale_error, s);
    double *xi = ext_buffer(0), *ri = ext_buffer(1);
    to_internal((const double*)x, xi, s);
    to_internal((const double*)r1, ri, s);
    const SolveResult res = solve_impl(method, prec, xi, ri, cap, tol, stale_error, s);
    to_external((const double*)xi, x, s);
    to_external((const double*)ri, r1, s);
    TF_CHECK(hipStreamSynchronize(s));
    return res;
}

int kernel_method(int method) {
    if (method != 1) return method;
    const char* e = std::getenv("TFLOWS_BICG_SHADOW");
    return (e && std::atoi(e) != 0) ? 1 : 0;
}

SolveResult System::solve_impl(int method, int prec, double* x, double* r1, int cap, double tol,
                               double stale_error, hipStream_t s) {
    if (method < 0 || method > 2) throw std::invalid_argument("method must be 0 (cg), 1 (bicg) or 2 (cgs)");
    method = kernel_method(method);
    ws_.ensure((size_t)V_COUNT * n_);
    KrylovVectors<double> w = vectors(x, r1);
    // iteration 1 overwrites p, p2, u1, u2 before reading them; zero the
    // workspace anyway so no value of a previous solve lingers
    TF_CHECK(hipMemsetAsync(ws_.p, 0, sizeof(double) * (size_t)V_COUNT * n_, s));

    SolverState h{};
    h.tol = tol;
    h.cap = cap;
    h.error = stale_error;
    TF_CHECK(hipMemcpyAsync(state_.p, &h, sizeof(h), hipMemcpyHostToDevice, s));
    mirror_->done = 0;
    mirror_->iter = 0;

    static const Levels kNoLevels;
    const Levels* L = prec == 2 ? &levels() : &kNoLevels;
    if (prec == 2) prepare_dic(false, s);
    const Csr<double> A = csr();
    with_op64(s, [&](const auto& Aop) {
        run_solve<double, std::decay_t<decltype(Aop)>>(ctx(), method, prec, cap, Aop, A, w, sites(), state_.p,
                                                       mirror_, L, s, false, false, &replay_);
    });
    SolverState h2;
    TF_CHECK(hipMemcpy(&h2, state_.p, sizeof(h2), hipMemcpyDeviceToHost));
    if (!h2.done) throw HipError("solver finished without reaching a stopping state");
    if (prec == 2) check_dic(s);
    SolveResult res;
    res.niter = h2.niter_out;
    res.ini_written = h2.ini_written != 0;
    res.ini_res = h2.ini_res;
    res.fin_res = h2.error;
    return res;
}

KrylovVectors<double> System::vectors(double* x, double* r1) {
    KrylovVectors<double> w;
    double* b = ws_.p;
    auto vec = [&](int k) { return b + (size_t)k * n_; };
    w.x = x;
    w.r1 = r1;
    w.D = vec(V_D);
    w.p = vec(V_P);
    w.q = vec(V_Q);
    w.r2 = vec(V_R2);
    w.p2 = vec(V_P2);
    w.q2 = vec(V_Q2);
    w.u1 = vec(V_U1);
    w.u2 = vec(V_U2);
    w.v2 = vec(V_V2);
    return w;
}

#define TF_SOLVE_INST(T, Op)                                                                                  \
    template void enqueue_iteration<T, Op>(const BatchCtx&, int, bool, int, const Op&, const Csr<T>&,         \
                                           const KrylovVectors<T>&, const Sites<T>&, SolverState*,            \
                                           HostMirror*, const Levels*, hipStream_t);                          \
    template void pace_iterations<T, Op>(const BatchCtx&, int, bool, int, const Op&, const Csr<T>&,           \
                                         const KrylovVectors<T>&, const Sites<T>&, SolverState*, HostMirror*, \
                                         const Levels*, hipStream_t, Replay*);                                \
    template void start_solve<T, Op>(const BatchCtx&, int, int, const Op&, const Csr<T>&,                     \
                                     const KrylovVectors<T>&, const Sites<T>&, SolverState*, HostMirror*,     \
                                     const Levels*, hipStream_t, bool, bool);                                 \
    template void run_solve<T, Op>(const BatchCtx&, int, int, int, const Op&, const Csr<T>&,                  \
                                   const KrylovVectors<T>&, const Sites<T>&, SolverState*, HostMirror*,       \
                                   const Levels*, hipStream_t, bool, bool, Replay*);
TF_SOLVE_INST(double, Csr<double>)
TF_SOLVE_INST(float, Csr<float>)
TF_SOLVE_INST(float, DiffCsr32)
TF_SOLVE_INST(float, DiffSell32)
TF_SOLVE_INST(double, Sell<double>)
template void dic_apply_levels<double>(const Levels&, const Csr<double>&, const double*, const double*,
                                       double*, const SolverState*, hipStream_t);
template void dic_apply_levels<float>(const Levels&, const Csr<float>&, const float*, const float*, float*,
                                      const SolverState*, hipStream_t);
template void System::dic_apply<double>(const Csr<double>&, const double*, const double*, double*,
                                        const SolverState*, hipStream_t);
template void System::dic_apply<float>(const Csr<float>&, const float*, const float*, float*,
                                       const SolverState*, hipStream_t);

}  // namespace tflows
