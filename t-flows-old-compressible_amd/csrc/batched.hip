// batched.hip -- K independent systems solved by one set of launches.
//
// The K matrices are concatenated block-diagonally into one System (rows
// and entries of system k follow those of system k-1).  Every kernel of the
// CG / BiCG / CGS iteration (krylov_kernels.hip) runs under a BatchCtx that
// maps each block to one system and a range of that system's 256-row tiles;
// reductions, alfa/beta, the stopping test and the SolverState are per
// system, so every system follows exactly the reference loop
// (Process/CG.f90:113-262, BiCG.f90:100-254, CGS.f90:112-337) on its own
// matrix while they share the launches.
//
// Launch machinery:
//  * a chunk of C iterations is captured once into a HIP graph; the
//    iteration index of each kernel is a launch constant plus a device
//    counter (kbase) that the chunk's last node advances, so the same graph
//    is replayed until every system has stopped;
//  * the host keeps one replay queued behind the running one and reads the
//    per-system `done` flags the leaders publish into pinned host memory;
//  * when the finished systems leave enough idle blocks, the block maps are
//    rebuilt for the running systems only (the graph reads the maps from
//    device memory, so it stays valid); the one reduction site that crosses
//    an iteration boundary (site B) is collapsed to one value per system
//    under the old maps and re-expanded under the new ones;
//  * the grid size and the SpMV form (lanes per row) come from a small
//    autotuner that times a few iterations of every candidate on the
//    caller's own systems, once per (precision, method, preconditioner).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <functional>
#include <numeric>

#include "batched.h"

namespace tflows {

namespace {
constexpr int kMaxGrid = 4096;
// guards and inner-target back-off of the FP32 inner solves, as in the
// single-system mixed solve: diverged above 1e6 (unit start), stalled when
// the smallest residual is older than max(100, its iteration); a failed
// step is retried with a 10x looser inner target, up to 0.1
constexpr double kInnerDivLimit = 1e6;
constexpr int kInnerStallMin = 100;
constexpr double kLoosestInnerRtol = 0.1;
int round8(int v) { return (v + kNumXcd - 1) / kNumXcd * kNumXcd; }
const Levels kNoLevels{};
// bound on a barrier wait of the persistent form (TFLOWS_BARRIER_WAIT_MS,
// default 2000 ms, wall clock), read when a batch is created
uint64_t barrier_wait_ms() {
    const char* e = std::getenv("TFLOWS_BARRIER_WAIT_MS");
    const long v = e ? std::atol(e) : 2000;
    return v < 0 ? 0 : (uint64_t)v;
}
}  // namespace


// Site B of each running system summed under the current maps into
// st->aux[0..1] (by its leader).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_site_collapse(BatchCtx c, const T* __restrict__ partB,
                                                          SolverState* st_base) {
    SolverState* st = st_base + c.sys();
    if (st->done) return;
    T tot[2];
    ctx_total<T, 2>(c, partB, tot);
    if (c.leader()) {
        st->aux[0] = (double)tot[0];
        st->aux[1] = (double)tot[1];
    }
}

// Site B re-expanded under the new maps: the system's first block carries
// the collapsed totals, its other blocks zeros (x + 0 == x, so the totals
// read back unchanged).
template <typename T>
__global__ void k_site_expand(BatchCtx c, T* __restrict__ partB, const SolverState* st_base) {
    if (threadIdx.x != 0) return;
    const int b = blockIdx.x, G = gridDim.x, sys = c.sys();
    const SolverState* st = st_base + sys;
    T v0 = T(0), v1 = T(0);
    if (!st->done && b == c.sys_blk0[sys]) {
        v0 = (T)st->aux[0];
        v1 = (T)st->aux[1];
    }
    partB[b] = v0;
    partB[G + b] = v1;
}

// Per-tile sums of v^2 (a tile never straddles two systems).
__global__ __launch_bounds__(kBlock) void k_tile_norm2(int ntiles, const int* __restrict__ tile_r0,
                                                       const double* __restrict__ v, double* __restrict__ out) {
    __shared__ double sm[kWavesPerBlock];
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int r = tile_r0[t] + (int)threadIdx.x;
        double a[1] = {r < tile_r0[t + 1] ? v[r] * v[r] : 0.0};
        block_sum<double, 1>(a, sm);
        if (threadIdx.x == 0) out[t] = a[0];
    }
}

// r32 = fp32(r64 * inv[sys]); x32 = 0
__global__ __launch_bounds__(kBlock) void k_tile_scale(int ntiles, const int* __restrict__ tile_r0,
                                                       const int* __restrict__ tile_sys,
                                                       const double* __restrict__ r64,
                                                       const double* __restrict__ inv, float* __restrict__ r32,
                                                       float* __restrict__ x32) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int r = tile_r0[t] + (int)threadIdx.x;
        if (r < tile_r0[t + 1]) {
            r32[r] = (float)(r64[r] * inv[tile_sys[t]]);
            x32[r] = 0.0f;
        }
    }
}

// x64 += s[sys] * fp64(d32); r64 = b
__global__ __launch_bounds__(kBlock) void k_tile_refine(int ntiles, const int* __restrict__ tile_r0,
                                                        const int* __restrict__ tile_sys, double* __restrict__ x64,
                                                        const double* __restrict__ scale,
                                                        const float* __restrict__ d32,
                                                        const double* __restrict__ b, double* __restrict__ r64) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int r = tile_r0[t] + (int)threadIdx.x;
        if (r < tile_r0[t + 1]) {
            x64[r] = x64[r] + scale[tile_sys[t]] * (double)d32[r];
            r64[r] = b[r];
        }
    }
}

// x = xprev on the systems whose flag is set (a rejected refinement step)
__global__ __launch_bounds__(kBlock) void k_tile_restore(int ntiles, const int* __restrict__ tile_r0,
                                                         const int* __restrict__ tile_sys, double* __restrict__ x,
                                                         const double* __restrict__ xprev,
                                                         const double* __restrict__ flag) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int r = tile_r0[t] + (int)threadIdx.x;
        if (r < tile_r0[t + 1] && flag[tile_sys[t]] != 0.0) x[r] = xprev[r];
    }
}

__global__ __launch_bounds__(kBlock) void k_fill32(int n, float v, float* __restrict__ a) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------------------------------------------------------------------
Batch::Batch(int k, const int* n, const int* nonz, const int* const* acol, const int* const* arow,
             const int* const* adia)
    : k_(k) {
    if (k <= 0) throw std::invalid_argument("a batch needs at least one system");
    roff_.assign(k + 1, 0);
    eoff_.assign(k + 1, 0);
    int64_t rows = 0, ents = 0;
    for (int i = 0; i < k; ++i) {
        if (n[i] <= 0) throw std::invalid_argument("every system needs N > 0");
        if (nonz[i] < n[i] + 1) throw std::invalid_argument("NONZ must be at least N + 1 for every system");
        if (acol[i][0] != 1 || acol[i][n[i]] != nonz[i])
            throw std::invalid_argument("Acol of every system must run from 1 to its NONZ");
        rows += n[i];
        ents += nonz[i] - 1;
        if (rows >= (1LL << 31) - 1 || ents >= (1LL << 31) - 2)
            throw std::invalid_argument("the batch exceeds 2^31 rows or entries");
        roff_[i + 1] = (int)rows;
        eoff_[i + 1] = (int)ents;
    }
    n_tot_ = (int)rows;
    nnz_tot_ = ents;
    // concatenated 1-based layout; each system's columns must stay in its
    // own block (the System constructor checks the rest of the layout).  A
    // renumbered system is stored in its locality order (reorder.hip): its
    // rows permuted inside its own block, each row's entries kept in their
    // order, so every row sum is unchanged.
    std::vector<int> cacol(n_tot_ + 1), carow(nnz_tot_), cadia(n_tot_);
    std::vector<int> gperm(n_tot_), geperm(nnz_tot_);
    bool any = false;
    for (int i = 0; i < k; ++i) {
        const int r0 = roff_[i], e0 = eoff_[i], ni = n[i], nzi = nonz[i] - 1;
        std::vector<int> rp(ni + 1), cl(nzi);
        for (int r = 0; r <= ni; ++r) rp[r] = acol[i][r] - 1;
        for (int r = 0; r < ni; ++r)
            if (acol[i][r + 1] < acol[i][r] || adia[i][r] < acol[i][r] || adia[i][r] >= acol[i][r + 1])
                throw std::invalid_argument("Acol / Adia of a system are not a valid row layout");
        for (int e = 0; e < nzi; ++e) {
            const int c = arow[i][e];
            if (c < 1 || c > ni) throw std::invalid_argument("Arow of a system holds a column outside 1..N");
            cl[e] = c - 1;
        }
        const std::vector<int> perm = locality_order(ni, rp, cl, true);
        std::vector<int> inv;
        if (!perm.empty()) {
            any = true;
            inv.assign(ni, 0);
            for (int j = 0; j < ni; ++j) inv[perm[j]] = j;
        }
        int pos = e0;
        for (int j = 0; j < ni; ++j) {
            const int o = perm.empty() ? j : perm[j];
            gperm[r0 + j] = r0 + o;
            cacol[r0 + j] = pos + 1;
            cadia[r0 + j] = pos + (adia[i][o] - acol[i][o]) + 1;
            for (int e = rp[o]; e < rp[o + 1]; ++e, ++pos) {
                carow[pos] = (perm.empty() ? cl[e] : inv[cl[e]]) + r0 + 1;
                geperm[pos] = e0 + e;
            }
        }
    }
    cacol[n_tot_] = (int)nnz_tot_ + 1;
    if (any) {
        perm_.ensure(n_tot_);
        eperm_.ensure(nnz_tot_);
        val_orig_.ensure(nnz_tot_);
        val_tmp_.ensure(nnz_tot_);
        TF_CHECK(hipMemcpy(perm_.p, gperm.data(), sizeof(int) * n_tot_, hipMemcpyHostToDevice));
        TF_CHECK(hipMemcpy(eperm_.p, geperm.data(), sizeof(int) * nnz_tot_, hipMemcpyHostToDevice));
        TF_CHECK(hipMemset(val_orig_.p, 0, sizeof(double) * nnz_tot_));
        reordered_ = true;
    }
    TF_CHECK(hipStreamCreateWithFlags(&bs_, hipStreamNonBlocking));
    sys_.reset(new System(n_tot_, (int)nnz_tot_ + 1, cacol.data(), carow.data(), cadia.data(), false, bs_,
                          /*allow_reorder=*/false));

    tiles_.assign(k, 0);
    tile0_.assign(k + 1, 0);
    for (int i = 0; i < k; ++i) {
        tiles_[i] = (n[i] + kTileRows - 1) / kTileRows;
        tile0_[i + 1] = tile0_[i] + tiles_[i];
    }
    const int T = tile0_[k];
    std::vector<int> tr0(T + 1), tsys(T);
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < tiles_[i]; ++j) {
            tr0[tile0_[i] + j] = roff_[i] + j * kTileRows;
            tsys[tile0_[i] + j] = i;
        }
    tr0[T] = n_tot_;
    tile_r0_.ensure(T + 1);
    tile_sys_.ensure(T);
    TF_CHECK(hipMemcpy(tile_r0_.p, tr0.data(), sizeof(int) * (T + 1), hipMemcpyHostToDevice));
    TF_CHECK(hipMemcpy(tile_sys_.p, tsys.data(), sizeof(int) * T, hipMemcpyHostToDevice));

    gmax_ = round8(std::max(kMaxGrid, k));
    const size_t map_ints = 3 * (size_t)gmax_ + k + 1;
    maps_.ensure(map_ints);
    TF_CHECK(hipHostMalloc((void**)&h_maps_, sizeof(int) * map_ints, hipHostMallocDefault));
    TF_CHECK(hipHostMalloc((void**)&h_koff_, sizeof(int) * k, hipHostMallocDefault));
    // TFLOWS_BATCH_GROUP_MB: the working-set budget of a group (default 192)
    if (const char* e = std::getenv("TFLOWS_BATCH_GROUP_MB")) {
        const double mb = std::atof(e);
        if (mb > 0.0) mall_budget_ = mb * 1024.0 * 1024.0;
    }
    kbase_.ensure(1);
    koff_.ensure(k);
    const size_t nn = (size_t)n_tot_;
    ws_.ensure((V_COUNT + 3) * nn);
    ws32_.ensure((V_COUNT + 2) * nn);
    partials_.ensure(Sites<double>::elems(gmax_));
    partials32_.ensure(Sites<float>::elems(gmax_));
    tnorm_.ensure(T);
    scale_.ensure(k);
    st_.ensure(k);
    st32_.ensure(k);
    TF_CHECK(hipHostMalloc((void**)&hm_, sizeof(HostMirror) * k, hipHostMallocMapped | hipHostMallocCoherent));
    TF_CHECK(hipHostMalloc((void**)&hm32_, sizeof(HostMirror) * k, hipHostMallocMapped | hipHostMallocCoherent));
    TF_CHECK(hipHostMalloc((void**)&snap_, sizeof(int) * 2 * k, hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(snap_, 0, sizeof(int) * 2 * k);
    bar_.ensure(k);
    pfault_.ensure(1);
    TF_CHECK(hipMemsetAsync(pfault_.p, 0, sizeof(int), bs_));
    {
        int dev = 0, khz = 0;
        TF_CHECK(hipGetDevice(&dev));
        TF_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        wait_ticks_ = (uint64_t)(khz > 0 ? khz : 100000) * barrier_wait_ms();
    }
    for (int i = 0; i < 2; ++i) TF_CHECK(hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming));
    for (int i = 2; i < 4; ++i) TF_CHECK(hipEventCreate(&ev_[i]));
    TF_CHECK(hipEventCreateWithFlags(&ev_[4], hipEventDisableTiming));
    TF_CHECK(hipStreamSynchronize(bs_));
}

Batch::~Batch() {
    for (auto& g : graphs_) (void)hipGraphExecDestroy(g.second);
    if (bs_) (void)hipStreamSynchronize(bs_);
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (h_maps_) (void)hipHostFree(h_maps_);
    if (h_koff_) (void)hipHostFree(h_koff_);
    if (hm_) (void)hipHostFree(hm_);
    if (hm32_) (void)hipHostFree(hm32_);
    if (snap_) (void)hipHostFree(snap_);
    sys_.reset();
    if (bs_) (void)hipStreamDestroy(bs_);
}

void Batch::set_values(int sys, const double* aval, bool device_ptr, hipStream_t s) {
    if (sys < -1 || sys >= k_) throw std::invalid_argument("system index out of range");
    if (!reordered_) {
        if (sys < 0) sys_->set_values(aval, device_ptr, s);
        else sys_->set_values_range(eoff_[sys], eoff_[sys + 1] - eoff_[sys], aval, device_ptr, s);
        return;
    }
    // values arrive in the callers' entry order; the stored matrix holds
    // them in the locality order
    const int64_t e0 = sys < 0 ? 0 : eoff_[sys], cnt = sys < 0 ? nnz_tot_ : eoff_[sys + 1] - eoff_[sys];
    TF_CHECK(hipMemcpyAsync(val_orig_.p + e0, aval, sizeof(double) * cnt,
                            device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    launch_gather((int)nnz_tot_, eperm_.p, val_orig_.p, val_tmp_.p, s);
    sys_->set_values(val_tmp_.p, true, s);
    if (!device_ptr) TF_CHECK(hipStreamSynchronize(s));
}

void Batch::gather_in(const double* ext, double* in) {
    if (reordered_) launch_gather(n_tot_, perm_.p, ext, in, bs_);
    else TF_CHECK(hipMemcpyAsync(in, ext, sizeof(double) * (size_t)n_tot_, hipMemcpyDeviceToDevice, bs_));
}
void Batch::scatter_out(const double* in, double* ext) {
    if (reordered_) launch_scatter(n_tot_, perm_.p, in, ext, bs_);
    else TF_CHECK(hipMemcpyAsync(ext, in, sizeof(double) * (size_t)n_tot_, hipMemcpyDeviceToDevice, bs_));
}

void Batch::join(hipStream_t s) {
    TF_CHECK(hipEventRecord(ev_[4], s));
    TF_CHECK(hipStreamWaitEvent(bs_, ev_[4], 0));
}

// Blocks per running system in proportion to its tiles (at least one
// each); inside a system, the blocks that the hardware sends to the same
// XCD (block index mod 8) own consecutive tiles.
void Batch::build_maps(const std::vector<char>& active, int G) {
    const int K = k_;
    int* blk_sys = h_maps_;
    int* blk_tiles = h_maps_ + gmax_;
    int* sys_blk0 = h_maps_ + 3 * gmax_;
    std::vector<char> act(active);
    int64_t tact = 0;
    int ka = 0;
    for (int i = 0; i < K; ++i)
        if (act[i]) {
            tact += tiles_[i];
            ++ka;
        }
    if (ka == 0) {  // nothing left: keep every system mapped (all return at once)
        std::fill(act.begin(), act.end(), 1);
        ka = K;
        tact = tile0_[K];
    }
    if (ka > G) throw std::invalid_argument("the grid must have at least one block per system");
    std::vector<int> nb(K, 0);
    std::vector<std::pair<double, int>> frac;
    const int spare = G - ka;
    int used = 0;
    for (int i = 0; i < K; ++i) {
        if (!act[i]) continue;
        const double q = (double)spare * tiles_[i] / (double)tact;
        const int f = (int)q;
        nb[i] = 1 + f;
        used += nb[i];
        frac.push_back({q - f, i});
    }
    std::stable_sort(frac.begin(), frac.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    for (size_t j = 0; used < G; ++j, ++used) nb[frac[j % frac.size()].second]++;
    int b = 0;
    std::vector<int> order;
    for (int i = 0; i < K; ++i) {
        sys_blk0[i] = b;
        const int cnt = nb[i];
        order.resize(cnt);
        std::iota(order.begin(), order.end(), b);
        std::stable_sort(order.begin(), order.end(),
                         [](int x, int y) { return (x & (kNumXcd - 1)) < (y & (kNumXcd - 1)); });
        for (int j = 0; j < cnt; ++j) {
            const int blk = order[j];
            blk_sys[blk] = i;
            blk_tiles[2 * blk] = tile0_[i] + (int)((int64_t)tiles_[i] * j / cnt);
            blk_tiles[2 * blk + 1] = tile0_[i] + (int)((int64_t)tiles_[i] * (j + 1) / cnt);
        }
        b += cnt;
    }
    sys_blk0[K] = G;
    map_grid_ = G;
}

void Batch::upload_maps(hipStream_t s) {
    TF_CHECK(hipMemcpyAsync(maps_.p, h_maps_, sizeof(int) * (3 * (size_t)gmax_ + k_ + 1), hipMemcpyHostToDevice,
                            s));
}

BatchCtx Batch::ctx(int lanes, bool with_kbase) const {
    BatchCtx c;
    c.blk_sys = maps_.p;
    c.blk_tiles = maps_.p + gmax_;
    c.tile_r0 = tile_r0_.p;
    c.sys_blk0 = maps_.p + 3 * gmax_;
    c.kbase = with_kbase ? kbase_.p : nullptr;
    c.koff = with_kbase ? koff_.p : nullptr;
    c.lanes = lanes;
    return c;
}

BatchConfig Batch::config_for(int bits, int method, int prec) const {
    BatchConfig c;
    auto it = tuned_.find(Key{bits, method, prec});
    if (it != tuned_.end()) c = it->second;
    if (fixed_.grid) c.grid = fixed_.grid;
    if (fixed_.lanes) c.lanes = fixed_.lanes;
    if (fixed_.chunk) c.chunk = fixed_.chunk;
    if (!c.chunk) c.chunk = prec == 2 ? 1 : 8;
    if (c.grid) c.grid = std::min(gmax_, std::max(round8(k_), round8(c.grid)));
    if (form_) c.form = form_;
    if (!c.form || method != 0 || prec == 2 || !persistent_ok_) c.form = kFormKernels;
    if (groups_) c.groups = groups_;
    if (c.groups != 2 || prec == 2 || c.form != kFormKernels) c.groups = 1;
    return c;
}

int Batch::groups_in_use(int bits, int method, int prec) const {
    return config_for(bits, method, prec).groups == 2 ? (int)mall_groups(bits, method).size() : 1;
}

std::vector<std::vector<int>> Batch::mall_groups(int bits, int method) const {
    const double vb = bits == 64 ? 8.0 : 4.0;
    // vectors a solve touches per row (x, r, p, q, D and the method's extras)
    const double vecs = method == 0 ? 5.0 : (method == 1 ? 8.0 : 10.0);
    std::vector<std::vector<int>> g;
    double used = 0.0;
    for (int i = 0; i < k_; ++i) {
        const double rows = roff_[i + 1] - roff_[i], ents = eoff_[i + 1] - eoff_[i];
        const double ws = rows * vecs * vb + ents * (vb + 4.0);
        if (g.empty() || used + ws > mall_budget_) {
            g.emplace_back();
            used = 0.0;
        }
        g.back().push_back(i);
        used += ws;
    }
    return g;
}

template <typename Op>
int Batch::persistent_grid(int method, int prec, int lanes) {
    if (method != 0 || prec == 2 || !persistent_ok_) return 0;
    const int resident = batch_cg_chunk_resident<Op>(lanes);
    const int G = std::min(gmax_, resident / kNumXcd * kNumXcd);
    return G >= round8(k_) ? G : 0;
}

// One chunk of gk.chunk iterations on stream s: the kernels of each
// iteration, or one persistent launch.
template <typename T, typename Op>
void Batch::enqueue_chunk(const GraphKey& gk, const BatchCtx& c, const Op& Aop, const Csr<T>& A,
                          const KrylovVectors<T>& w, T* partials, SolverState* st, HostMirror* hm, hipStream_t s) {
    const Sites<T> S = Sites<T>::carve(partials, gk.grid);
    if (gk.form == kFormPersistent) {
        launch_batch_cg_chunk<Op, T>(c, Aop, A.n, gk.chunk, w.x, w.r1, w.p, w.q, (const T*)w.D, S, st, hm, bar_.p,
                                     k_, pfault_.p, wait_ticks_, s);
        return;
    }
    const Levels* L = gk.prec == 2 ? &sys_->levels() : &kNoLevels;
    for (int i = 1; i <= gk.chunk; ++i)
        enqueue_iteration<T, Op>(c, gk.method, gk.prec != 2, i, Aop, A, w, S, st, hm, L, s);
}

// The done flags after a chunk of `chunk` iterations into slot
// (chunk index & 1) of snap; the chunk index follows from the iteration
// base, which the chunk has just advanced.
__global__ __launch_bounds__(kBlock) void k_done_snapshot(int k, int chunk, const SolverState* __restrict__ st,
                                                          const int* __restrict__ kbase, int* snap) {
    const int m = *kbase / chunk - 1;
    for (int i = threadIdx.x; i < k; i += kBlock) snap[(m & 1) * k + i] = st[i].done;
}

template <typename T, typename Op>
hipGraphExec_t Batch::graph(const GraphKey& gk, const Op& Aop, const Csr<T>& A, const KrylovVectors<T>& w,
                            T* partials, SolverState* st, HostMirror* hm) {
    auto it = graphs_.find(gk);
    if (it != graphs_.end()) return it->second;
    const BatchCtx c = ctx(gk.lanes, true);
    hipGraph_t g = nullptr;
    TF_CHECK(hipStreamBeginCapture(bs_, hipStreamCaptureModeThreadLocal));
    try {
        enqueue_chunk<T, Op>(gk, c, Aop, A, w, partials, st, hm, bs_);
        launch_advance(kbase_.p, gk.chunk, bs_);
        hipLaunchKernelGGL(k_done_snapshot, dim3(1), dim3(kBlock), 0, bs_, k_, gk.chunk, (const SolverState*)st,
                           (const int*)kbase_.p, snap_);
    } catch (...) {
        (void)hipStreamEndCapture(bs_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    TF_CHECK(hipStreamEndCapture(bs_, &g));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    TF_CHECK(e);
    graphs_[gk] = exec;
    return exec;
}

template <typename T>
void Batch::rebalance(const std::vector<char>& active, int grid, T* partials, SolverState* st) {
    const BatchCtx c = ctx(1, false);
    const Sites<T> S = Sites<T>::carve(partials, grid);
    hipLaunchKernelGGL((k_site_collapse<T>), dim3(grid), dim3(kBlock), 0, bs_, c, (const T*)S.B, st);
    TF_CHECK(hipGetLastError());
    // the staging buffer is rewritten below: let the previous upload finish
    TF_CHECK(hipStreamSynchronize(bs_));
    build_maps(active, grid);
    upload_maps(bs_);
    hipLaunchKernelGGL((k_site_expand<T>), dim3(grid), dim3(kWave), 0, bs_, c, S.B, (const SolverState*)st);
    TF_CHECK(hipGetLastError());
    ++rebalances_;
}

// Times a few iterations of every (grid, lanes) candidate on the caller's
// systems; returns the fastest.
template <typename T, typename Op>
BatchConfig Batch::tune(int bits, int method, int prec, const Op& Aop, const Csr<T>& A,
                        const KrylovVectors<T>& w, T* partials, SolverState* st, HostMirror* hm,
                        const std::vector<SolverState>& init, bool d_ready, bool zero_x,
                        const std::function<void()>& prepare) {
    (void)bits;
    static const int kGrids[] = {1024, 2048, 4096};
    static const int kLanes[] = {1, 2, 4, 8};
    const bool jac = prec != 2;
    const int probe = jac ? 6 : 2;
    const Levels* L = prec == 2 ? &sys_->levels() : &kNoLevels;
    const std::vector<char> all(k_, 1);
    BatchConfig best;
    float best_ms = 1e30f;
    // the persistent form: one launch of `probe` iterations on the largest
    // resident grid and on half of it
    constexpr bool lane_forms_p = std::is_same<Op, Csr<T>>::value || std::is_same<Op, DiffCsr32>::value;
    const bool try_persistent = form_ != kFormKernels && persistent_grid<Op>(method, prec, 1) > 0;
    auto time_persistent = [&](int rep) {
        for (int lanes : kLanes) {
            if (!lane_forms_p && lanes != 1) continue;
            const int Gmax = persistent_grid<Op>(method, prec, lanes);
            for (int G : {Gmax, std::max(round8(k_), Gmax / 2 / kNumXcd * kNumXcd)}) {
                if (G <= 0 || !persistent_ok_) continue;
                TF_CHECK(hipStreamSynchronize(bs_));
                build_maps(all, G);
                upload_maps(bs_);
                const BatchCtx c = ctx(lanes, false);
                const Sites<T> S = Sites<T>::carve(partials, G);
                prepare();
                TF_CHECK(hipMemcpyAsync(st, init.data(), sizeof(SolverState) * k_, hipMemcpyHostToDevice, bs_));
                start_solve<T, Op>(c, method, prec, Aop, A, w, S, st, hm, L, bs_, d_ready, zero_x);
                TF_CHECK(hipEventRecord(ev_[2], bs_));
                launch_batch_cg_chunk<Op, T>(c, Aop, A.n, probe, w.x, w.r1, w.p, w.q, (const T*)w.D, S, st, hm,
                                             bar_.p, k_, pfault_.p, wait_ticks_, bs_);
                TF_CHECK(hipEventRecord(ev_[3], bs_));
                TF_CHECK(hipEventSynchronize(ev_[3]));
                int fault = 0;
                TF_CHECK(hipMemcpy(&fault, pfault_.p, sizeof(int), hipMemcpyDeviceToHost));
                if (fault) {   // not co-resident here: the batch keeps the per-kernel form
                    persistent_ok_ = false;
                    launch_set_int(pfault_.p, 0, bs_);
                    return;
                }
                float ms = 0.0f;
                TF_CHECK(hipEventElapsedTime(&ms, ev_[2], ev_[3]));
                if (rep == 1 && ms < best_ms) {
                    best_ms = ms;
                    best.grid = G;
                    best.lanes = lanes;
                    best.form = kFormPersistent;
                }
            }
        }
    };
    for (int rep = 0; rep < 2; ++rep) {  // the first pass also loads the code objects
        if (try_persistent) time_persistent(rep);
        if (form_ == kFormPersistent && try_persistent && persistent_ok_) continue;
        for (int g0 : kGrids) {
            const int G = std::min(gmax_, std::max(round8(k_), g0));
            TF_CHECK(hipStreamSynchronize(bs_));
            build_maps(all, G);
            upload_maps(bs_);
            // the sliced operators have one form; the CSR ones four
            constexpr bool lane_forms = std::is_same<Op, Csr<T>>::value || std::is_same<Op, DiffCsr32>::value;
            for (int lanes : kLanes) {
                if (!lane_forms && lanes != 1) continue;
                const BatchCtx c = ctx(lanes, false);
                const Sites<T> S = Sites<T>::carve(partials, G);
                prepare();
                TF_CHECK(hipMemcpyAsync(st, init.data(), sizeof(SolverState) * k_, hipMemcpyHostToDevice, bs_));
                start_solve<T, Op>(c, method, prec, Aop, A, w, S, st, hm, L, bs_, d_ready, zero_x);
                TF_CHECK(hipEventRecord(ev_[2], bs_));
                for (int i = 1; i <= probe; ++i) enqueue_iteration<T, Op>(c, method, jac, i, Aop, A, w, S, st, hm, L, bs_);
                TF_CHECK(hipEventRecord(ev_[3], bs_));
                TF_CHECK(hipEventSynchronize(ev_[3]));
                float ms = 0.0f;
                TF_CHECK(hipEventElapsedTime(&ms, ev_[2], ev_[3]));
                if (rep == 1 && ms < best_ms) {
                    best_ms = ms;
                    best.grid = G;
                    best.lanes = lanes;
                    best.form = kFormKernels;
                }
            }
        }
    }
    // the same configuration with the systems in MALL-sized groups, each
    // group on the whole grid in turn: kept if the groups' probes together
    // take less time than the probe of every system at once
    best.groups = 1;
    const std::vector<std::vector<int>> groups = mall_groups(bits, method);
    if (groups_ == 0 && prec != 2 && best.form == kFormKernels && best.grid && groups.size() > 1) {
        const int G = best.grid;
        float sum = 0.0f;
        for (const auto& g : groups) {
            std::vector<char> m(k_, 0);
            for (int i : g) m[i] = 1;
            TF_CHECK(hipStreamSynchronize(bs_));
            build_maps(m, G);
            upload_maps(bs_);
            const BatchCtx c = ctx(best.lanes, false);
            const Sites<T> S = Sites<T>::carve(partials, G);
            prepare();
            TF_CHECK(hipMemcpyAsync(st, init.data(), sizeof(SolverState) * k_, hipMemcpyHostToDevice, bs_));
            start_solve<T, Op>(c, method, prec, Aop, A, w, S, st, hm, L, bs_, d_ready, zero_x);
            TF_CHECK(hipEventRecord(ev_[2], bs_));
            for (int i = 1; i <= probe; ++i) enqueue_iteration<T, Op>(c, method, jac, i, Aop, A, w, S, st, hm, L, bs_);
            TF_CHECK(hipEventRecord(ev_[3], bs_));
            TF_CHECK(hipEventSynchronize(ev_[3]));
            float ms = 0.0f;
            TF_CHECK(hipEventElapsedTime(&ms, ev_[2], ev_[3]));
            sum += ms;
        }
        if (sum < best_ms) best.groups = 2;
    }
    return best;
}

template <typename T, typename Op>
void Batch::run(int bits, int method, int prec, const Op& Aop, const Csr<T>& A, const KrylovVectors<T>& w,
                T* partials, SolverState* st, HostMirror* hm, const std::vector<SolverState>& init, bool d_ready,
                bool zero_x, const std::function<void()>& prepare, std::vector<SolverState>& out) {
    const Key key{bits, method, prec};
    if (prec == 2) sys_->prepare_dic(bits == 32, bs_);
    const bool auto_cfg = !fixed_.grid || !fixed_.lanes || (form_ == 0 && method == 0 && prec != 2);
    if (auto_cfg && tuned_.find(key) == tuned_.end())
        tuned_[key] = tune<T, Op>(bits, method, prec, Aop, A, w, partials, st, hm, init, d_ready, zero_x, prepare);
    const BatchConfig cfg = config_for(bits, method, prec);
    int G = cfg.grid ? cfg.grid : std::min(gmax_, std::max(round8(k_), 2048));
    const int lanes = cfg.lanes ? cfg.lanes : 1;
    const int C = cfg.chunk;
    int form = cfg.form;
    if (form == kFormPersistent) {   // the grid must be resident all at once
        const int Gp = persistent_grid<Op>(method, prec, lanes);
        if (Gp > 0) G = std::min(G, Gp);
        else form = kFormKernels;
    }
    const bool jac = prec != 2;
    const Levels* L = prec == 2 ? &sys_->levels() : &kNoLevels;
    hipGraphExec_t exec =
        graph<T, Op>(GraphKey{bits, method, prec, G, lanes, C, form}, Aop, A, w, partials, st, hm);

    std::vector<char> act(k_, 1);
    TF_CHECK(hipStreamSynchronize(bs_));
    build_maps(act, G);
    upload_maps(bs_);
    prepare();
    TF_CHECK(hipMemcpyAsync(st, init.data(), sizeof(SolverState) * k_, hipMemcpyHostToDevice, bs_));
    for (int i = 0; i < k_; ++i) {
        hm[i].done = 0;
        hm[i].iter = 0;
    }
    const BatchCtx c0 = ctx(lanes, false);
    start_solve<T, Op>(c0, method, prec, Aop, A, w, Sites<T>::carve(partials, G), st, hm, L, bs_, d_ready, zero_x);
    launch_set_int(kbase_.p, 0, bs_);
    TF_CHECK(hipGetLastError());
    (void)jac;

    // Groups (cfg.groups == 2): the systems run group after group, each
    // group on the whole grid; a group starts with the chunk after the one
    // whose snapshot showed the previous group done, and its systems count
    // their iterations from there (koff).
    const std::vector<std::vector<int>> groups =
        cfg.groups == 2 ? mall_groups(bits, method) : std::vector<std::vector<int>>{};
    const bool grouped = groups.size() > 1;
    std::vector<int> gid(k_, 0);
    for (size_t g = 0; g < groups.size(); ++g)
        for (int i : groups[g]) gid[i] = (int)g;
    for (int i = 0; i < k_; ++i) h_koff_[i] = 0;
    TF_CHECK(hipMemcpyAsync(koff_.p, h_koff_, sizeof(int) * k_, hipMemcpyHostToDevice, bs_));
    int cur = 0;
    auto live_in = [&](int g, const int* done) {
        std::vector<char> m(k_, 0);
        for (int i : groups[g]) m[i] = done ? done[i] == 0 : 1;
        return m;
    };
    int64_t mapped = tile0_[k_];
    long long limit = 0;
    if (grouped) {
        for (const auto& g : groups) {
            int gcap = 0;
            for (int i : g) gcap = std::max(gcap, init[i].cap);
            limit += (long long)gcap + 1 + 2LL * C;
        }
        limit += 2LL * C;
        act = live_in(0, nullptr);
        rebalance<T>(act, G, partials, st);
        mapped = 0;
        for (int i : groups[0]) mapped += tiles_[i];
    } else {
        int maxcap = 0;
        for (const auto& h : init) maxcap = std::max(maxcap, h.cap);
        limit = (long long)maxcap + 1 + 2LL * C;
    }
    bool pending[2] = {false, false};
    int slot = 0;
    long long issued = 0, launched = 0;
    bool all_done = false;
    while (issued < limit) {
        TF_CHECK(hipGraphLaunch(exec, bs_));
        issued += C;
        ++launched;
        TF_CHECK(hipEventRecord(ev_[slot], bs_));
        pending[slot] = true;
        const int other = slot ^ 1;
        if (pending[other]) {
            // chunk launched - 2 has ended: its snapshot is the state at
            // that boundary, whatever the device has done since
            TF_CHECK(hipEventSynchronize(ev_[other]));
            pending[other] = false;
            const int* sn = snap_ + ((launched - 2) & 1) * k_;
            int64_t running = 0;
            all_done = true;
            for (int i = 0; i < k_; ++i) {
                if (sn[i] == 0) all_done = false;
                act[i] = sn[i] == 0 && (!grouped || gid[i] == cur);
                if (act[i]) running += tiles_[i];
            }
            if (all_done) break;
            if (grouped && running == 0) {
                // the group is done: the next group with a running system
                // takes the grid from the next chunk on
                do ++cur;
                while (cur < (int)groups.size() &&
                       std::none_of(groups[cur].begin(), groups[cur].end(), [&](int i) { return sn[i] == 0; }));
                if (cur >= (int)groups.size()) break;   // not reached: all_done covers it
                for (int i : groups[cur]) h_koff_[i] = (int)(launched * C);
                TF_CHECK(hipMemcpyAsync(koff_.p, h_koff_, sizeof(int) * k_, hipMemcpyHostToDevice, bs_));
                act = live_in(cur, sn);
                for (int i : groups[cur])
                    if (act[i]) running += tiles_[i];
                rebalance<T>(act, G, partials, st);
                mapped = running;
            } else if (rebalance_ && running * 10 <= mapped * 7) {
                rebalance<T>(act, G, partials, st);
                mapped = running;
            }
        }
        slot = other;
    }
    TF_CHECK(hipStreamSynchronize(bs_));
    if (form == kFormPersistent) {
        int fault = 0;
        TF_CHECK(hipMemcpy(&fault, pfault_.p, sizeof(int), hipMemcpyDeviceToHost));
        if (fault) {
            // a barrier wait gave up (other work held the CUs): the batch
            // keeps the per-kernel form and solves again from the inputs,
            // which prepare() restores
            persistent_ok_ = false;
            launch_set_int(pfault_.p, 0, bs_);
            std::fprintf(stderr, "libtflows_krylov: a persistent batch launch gave up waiting "
                                 "(TFLOWS_BARRIER_WAIT_MS); the batch now runs one launch per kernel\n");
            std::fflush(stderr);
            run<T, Op>(bits, method, prec, Aop, A, w, partials, st, hm, init, d_ready, zero_x, prepare, out);
            return;
        }
    }
    out.resize(k_);
    TF_CHECK(hipMemcpy(out.data(), st, sizeof(SolverState) * k_, hipMemcpyDeviceToHost));
    for (int i = 0; i < k_; ++i)
        if (!out[i].done) throw HipError("batched solve finished without every system reaching a stopping state");
    if (prec == 2) sys_->check_dic(bs_);
}

KrylovVectors<double> Batch::vectors64() {
    const size_t n = (size_t)n_tot_;
    double* b = ws_.p;
    KrylovVectors<double> w;
    w.D = b + V_D * n;
    w.p = b + V_P * n;
    w.q = b + V_Q * n;
    w.r2 = b + V_R2 * n;
    w.p2 = b + V_P2 * n;
    w.q2 = b + V_Q2 * n;
    w.u1 = b + V_U1 * n;
    w.u2 = b + V_U2 * n;
    w.v2 = b + V_V2 * n;
    w.x = b + V_COUNT * n;
    w.r1 = b + (V_COUNT + 1) * n;
    return w;
}

KrylovVectors<float> Batch::vectors32() {
    const size_t n = (size_t)n_tot_;
    float* b = ws32_.p;
    KrylovVectors<float> w;
    w.D = b + V_D * n;
    w.p = b + V_P * n;
    w.q = b + V_Q * n;
    w.r2 = b + V_R2 * n;
    w.p2 = b + V_P2 * n;
    w.q2 = b + V_Q2 * n;
    w.u1 = b + V_U1 * n;
    w.u2 = b + V_U2 * n;
    w.v2 = b + V_V2 * n;
    w.x = b + V_COUNT * n;
    w.r1 = b + (V_COUNT + 1) * n;
    return w;
}

static void check_method(int method, int prec) {
    if (method < 0 || method > 2) throw std::invalid_argument("method must be 0 (cg), 1 (bicg) or 2 (cgs)");
    if (prec < 0 || prec > 2) throw std::invalid_argument("prec must be 0, 1 or 2");
}

void Batch::solve(int method, int prec, double* x, double* r1, const int* cap, const double* tol,
                  SolveResult* res, hipStream_t s) {
    check_method(method, prec);
    method = kernel_method(method);
    join(s);
    const KrylovVectors<double> w = vectors64();
    std::vector<SolverState> init(k_);
    for (int i = 0; i < k_; ++i) {
        init[i] = SolverState{};
        init[i].tol = tol[i];
        init[i].cap = cap[i];
        init[i].error = res[i].fin_res;   // returned on the ||b|| < tol exit
    }
    auto prepare = [&] {
        gather_in(x, w.x);
        gather_in(r1, w.r1);
    };
    const Csr<double> A = sys_->csr();
    std::vector<SolverState> out;
    sys_->with_op64(bs_, [&](const auto& Aop) {
        run<double, std::decay_t<decltype(Aop)>>(64, method, prec, Aop, A, w, partials_.p, st_.p, hm_, init, false,
                                                 false, prepare, out);
    });
    scatter_out(w.x, x);
    scatter_out(w.r1, r1);
    TF_CHECK(hipStreamSynchronize(bs_));
    for (int i = 0; i < k_; ++i) {
        res[i].niter = out[i].niter_out;
        res[i].ini_written = out[i].ini_written != 0;
        if (res[i].ini_written) res[i].ini_res = out[i].ini_res;
        res[i].fin_res = out[i].error;
    }
}

// Per-system ||v||_2 from per-tile partial sums, added on the host in tile
// order.
void Batch::tile_norms(const double* v, std::vector<double>& out) {
    const int T = tile0_[k_];
    hipLaunchKernelGGL(k_tile_norm2, dim3(std::min(T, kMaxGrid)), dim3(kBlock), 0, bs_, T, (const int*)tile_r0_.p, v,
                       tnorm_.p);
    TF_CHECK(hipGetLastError());
    std::vector<double> h(T);
    TF_CHECK(hipMemcpyAsync(h.data(), tnorm_.p, sizeof(double) * T, hipMemcpyDeviceToHost, bs_));
    TF_CHECK(hipStreamSynchronize(bs_));
    out.assign(k_, 0.0);
    for (int i = 0; i < k_; ++i) {
        double a = 0.0;
        for (int t = tile0_[i]; t < tile0_[i + 1]; ++t) a += h[t];
        out[i] = std::sqrt(a);
    }
}

// Refinement per system: an inner FP32 solve on the unit-scaled residual,
// x += ||r|| d, a new FP64 residual.  A step that does not lower the FP64
// residual, or whose inner solve diverged (its residual rose above
// kInnerDivLimit), is undone and retried with a 10x
// looser inner target; once the target reaches kLoosestInnerRtol a failed
// step ends that system's refinement, which then reports non-convergence
// the reference way (niter = iterations + 1, FinRes = ||b - A x||).
void Batch::solve_mixed(int method, int prec, double* x, const double* b, double* r_out, const int* cap,
                        const double* tol, double inner_rtol, int max_outer, SolveResult* res, int* outer_steps,
                        hipStream_t s) {
    check_method(method, prec);
    method = kernel_method(method);
    if (!(inner_rtol > 0.0 && inner_rtol < 1.0)) throw std::invalid_argument("inner_rtol must be in (0, 1)");
    join(s);
    const int n = n_tot_, T = tile0_[k_];
    const size_t bytes = sizeof(double) * (size_t)n;
    double* x64 = ws_.p + (size_t)V_COUNT * n;
    double* r64 = x64 + n;
    double* b64 = r64 + n;
    double* xprev = ws_.p + (size_t)V_P * n;   // FP64 slots are free during a mixed solve
    gather_in(x, x64);
    gather_in(b, b64);

    const Csr<float> A32 = prec == 2 ? sys_->csr32(bs_) : Csr<float>{n, sys_->rowptr(), sys_->col(), sys_->dia(), nullptr};
    const KrylovVectors<float> w = vectors32();
    const int nb = (n + kBlock - 1) / kBlock;
    if (prec == 1) {
        TF_CHECK(hipMemcpyAsync(w.D, sys_->diag32(bs_), sizeof(float) * n, hipMemcpyDeviceToDevice, bs_));
    } else if (prec == 2) {
        double* d64 = ws_.p;   // FP64 scratch (the V_D slot)
        sys_->dic_setup(d64, bs_);
        launch_f64_to_f32(n, d64, w.D, bs_);
    } else {
        hipLaunchKernelGGL(k_fill32, dim3(nb), dim3(kBlock), 0, bs_, n, 1.0f, w.D);
    }
    TF_CHECK(hipGetLastError());
    const int tgrid = std::min(T, kMaxGrid);
    auto residual = [&](std::vector<double>& err) {
        TF_CHECK(hipMemcpyAsync(r64, b64, bytes, hipMemcpyDeviceToDevice, bs_));
        sys_->resid(x64, r64, bs_);
        tile_norms(r64, err);
    };
    std::vector<double> err, sc(k_);
    residual(err);
    std::vector<int> total(k_, 0);
    std::vector<char> open(k_, 0);   // still refining
    for (int i = 0; i < k_; ++i) {
        res[i].ini_res = err[i];
        res[i].ini_written = true;
        open[i] = err[i] >= tol[i] && cap[i] > 0;
    }
    int outer = 0;
    std::vector<SolverState> init(k_), out;
    std::vector<double> rtol(k_, inner_rtol);
    while (outer < max_outer && std::any_of(open.begin(), open.end(), [](char c) { return c != 0; })) {
        ++outer;
        for (int i = 0; i < k_; ++i) {
            sc[i] = open[i] ? 1.0 / err[i] : 0.0;   // closed systems: r32 = 0, an immediate exit
            init[i] = SolverState{};
            init[i].tol = rtol[i];                  // ||r32|| = 1 up to rounding
            init[i].cap = open[i] ? cap[i] - total[i] : 0;
            init[i].div_limit = kInnerDivLimit;
            init[i].stall_min = kInnerStallMin;
            init[i].err_min[0] = HUGE_VAL;
        }
        TF_CHECK(hipMemcpyAsync(scale_.p, sc.data(), sizeof(double) * k_, hipMemcpyHostToDevice, bs_));
        auto prepare = [&] {
            hipLaunchKernelGGL(k_tile_scale, dim3(tgrid), dim3(kBlock), 0, bs_, T, (const int*)tile_r0_.p,
                               (const int*)tile_sys_.p, (const double*)r64, (const double*)scale_.p, w.r1, w.x);
        };
        sys_->with_op32(bs_, [&](const auto& Aop) {
            run<float, std::decay_t<decltype(Aop)>>(32, method, prec, Aop, A32, w, partials32_.p, st32_.p, hm32_, init,
                                                    true, true, prepare, out);
        });
        std::vector<int> its(k_, 0);
        for (int i = 0; i < k_; ++i) {
            sc[i] = 0.0;
            if (!open[i]) continue;
            its[i] = out[i].niter_out > init[i].cap ? std::max(init[i].cap, 0) : out[i].niter_out;
            total[i] += its[i];
            if (!out[i].diverged) sc[i] = err[i];
        }
        TF_CHECK(hipMemcpyAsync(xprev, x64, bytes, hipMemcpyDeviceToDevice, bs_));
        TF_CHECK(hipMemcpyAsync(scale_.p, sc.data(), sizeof(double) * k_, hipMemcpyHostToDevice, bs_));
        hipLaunchKernelGGL(k_tile_refine, dim3(tgrid), dim3(kBlock), 0, bs_, T, (const int*)tile_r0_.p,
                           (const int*)tile_sys_.p, x64, (const double*)scale_.p, (const float*)w.x,
                           (const double*)b64, r64);
        TF_CHECK(hipGetLastError());
        sys_->resid(x64, r64, bs_);
        std::vector<double> err_new;
        tile_norms(r64, err_new);
        std::vector<double> reject(k_, 0.0);
        bool any_reject = false;
        for (int i = 0; i < k_; ++i)
            if (open[i] && !(err_new[i] < err[i])) {
                reject[i] = 1.0;
                any_reject = true;
            }
        if (any_reject) {   // undo those steps and recompute their residuals
            TF_CHECK(hipMemcpyAsync(scale_.p, reject.data(), sizeof(double) * k_, hipMemcpyHostToDevice, bs_));
            hipLaunchKernelGGL(k_tile_restore, dim3(tgrid), dim3(kBlock), 0, bs_, T, (const int*)tile_r0_.p,
                               (const int*)tile_sys_.p, x64, (const double*)xprev, (const double*)scale_.p);
            TF_CHECK(hipGetLastError());
            TF_CHECK(hipMemcpyAsync(r64, b64, bytes, hipMemcpyDeviceToDevice, bs_));
            sys_->resid(x64, r64, bs_);
            tile_norms(r64, err_new);
        }
        err = err_new;
        for (int i = 0; i < k_; ++i) {
            if (!open[i]) continue;
            const bool failed = reject[i] != 0.0 || out[i].diverged;
            if (failed && rtol[i] < kLoosestInnerRtol && total[i] < cap[i]) {
                rtol[i] = std::min(kLoosestInnerRtol, rtol[i] * 10.0);   // retry the step
                continue;
            }
            if (err[i] < tol[i] || its[i] == 0 || total[i] >= cap[i] || failed) open[i] = 0;
        }
    }
    scatter_out(x64, x);
    if (r_out) scatter_out(r64, r_out);
    TF_CHECK(hipStreamSynchronize(bs_));
    for (int i = 0; i < k_; ++i) {
        res[i].niter = err[i] < tol[i] ? total[i] : total[i] + 1;
        res[i].fin_res = err[i];
    }
    if (outer_steps) *outer_steps = outer;
}

}  // namespace tflows
