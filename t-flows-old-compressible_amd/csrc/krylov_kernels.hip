// krylov_kernels.hip -- the gfx950 kernels of the CG / BiCG / CGS path.
//
// One Krylov iteration is a short, fixed sequence of fused kernels.  Every
// scalar the algorithm needs between them (rho, alfa, beta, ||r||) is
// derived on the device from per-block partial sums ("reduction sites"):
// the producing kernel stores one partial per block, the consuming kernel
// sums them in a fixed order in every block.  There is no host round trip,
// no atomic and no fence inside an iteration.  Each kernel first checks
// SolverState::done and returns at once after convergence, so the host can
// run ahead by whole chunks of iterations.
//
// Every solver kernel takes a BatchCtx (kernels.h): empty for one system,
// or the block / tile / system maps of a batch of systems stored as one
// block-diagonal matrix, in which case each system has its own SolverState,
// reduction range and leader block, and systems finish independently.
//
// Reduction sites (each 2 values x G blocks):
//   A  ||b||^2                       k_setup          -> k_resid_init
//   B  ||r||^2, next rho             k_resid_init / *_update / k_cgs_spmv_r
//                                                     -> stopping test of the next iteration
//   C  rho = (r,z) for DIC forms     k_rho            -> k_cg_p / k_bicg_p
//   D  alfa denominator              *_spmv*          -> *_update / k_cgs_q / k_cgs_spmv_r
//
// Reference loops restated (single domain, Sequential build):
//   Prec1  Process/Prec1.f90:45-66          -> k_setup, k_dic_setup_level
//   Resid  Process/Resid.f90:29-34          -> k_resid_init, k_resid
//   CG     Process/CG.f90:62-262            -> k_cg_p, k_cg_spmv, k_cg_update
//   BiCG   Process/BiCG.f90:50-254          -> k_bicg_p, k_bicg_spmv, k_bicg_update
//   CGS    Process/CGS.f90:62-337           -> k_cgs_u, k_cgs_spmv_v, k_cgs_q, k_cgs_spmv_r
//   DIC    Process/CG.f90:135-156           -> k_dic_fwd_level, k_dic_bwd_level
#include <algorithm>
#include <climits>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "krylov_kernels.h"
#include "tile_plan.h"

namespace tflows {

// z_i of the Jacobi / identity preconditioner: a true division by D
// (Process/CG.f90:124-127; prec = 0 uses D = 1, and r/1 == r exactly).
template <typename T>
__device__ __forceinline__ T jac(T r, T d) { return r / d; }

// rows per thread whose operands the element-wise kernels load before they
// wait on a grid reduction
constexpr int kPrefetch = 4;

#define TF_STATE                                              \
    SolverState* st = st_base + ctx.sys();                    \
    HostMirror* hm = hm_base ? hm_base + ctx.sys() : nullptr; \
    (void)hm

// ---------------------------------------------------------------------------
// Stopping test at the start of iteration k, from residual site B written
// by the previous kernel chain: for k == 1 it is the initial-residual test
// (CG.f90:93-108), for k > 1 the test at the end of iteration k-1
// (CG.f90:248-255), and k == cap + 1 closes the DO loop with niter = cap + 1
// (CG.f90:259-262).  Every block takes the same decision; the leader
// records it.  Returns true when the solve is over.
// ---------------------------------------------------------------------------
// The state fields are read before the partial sums are waited for, so the
// two sets of loads are in flight together.  `halt`, when given, is a flag
// in shared memory that thread 0 set before the call (the system was over
// before this launch); it is read after the reduction's barriers.
template <typename T>
__device__ __forceinline__ bool check_stop(const BatchCtx& ctx, int k, const T* __restrict__ partB,
                                           SolverState* st, HostMirror* hm, T (&tot)[2],
                                           const int* halt = nullptr) {
    const double tol = st->tol;
    const int cap = st->cap;
    const double div_limit = st->div_limit;
    const double prev_min = st->err_min[(k - 1) & 1];
    const int prev_kmin = st->k_min[(k - 1) & 1];
    const int stall_min = st->stall_min;
    ctx_total<T, 2>(ctx, partB, tot);
    if (halt && *halt) return true;
    const double err = sqrt((double)tot[0]);
    const bool lead = ctx.leader();
    if (lead) {
        st->error = err;
        if (k == 1) {
            st->ini_res = err;
            st->ini_written = 1;
        }
    }
    if (err < tol) {
        if (lead) {
            st->done = 1;
            st->niter_out = k - 1;
            publish(hm, 1, k - 1);
        }
        return true;
    }
    if (div_limit > 0.0) {
        const double prev = prev_min;
        const bool lower = err < prev;
        const int kmin = lower ? k : prev_kmin;
        if (lead) {
            st->err_min[k & 1] = lower ? err : prev;
            st->k_min[k & 1] = kmin;
        }
        const bool diverged = !(err <= div_limit);
        const int window = stall_min > kmin ? stall_min : kmin;
        if (diverged || (stall_min > 0 && k - kmin > window)) {
            if (lead) {
                st->done = 1;
                st->diverged = diverged ? 1 : 0;
                st->niter_out = k - 1;
                publish(hm, 1, k - 1);
            }
            return true;
        }
    }
    if (k > cap) {
        if (lead) {
            st->done = 1;
            st->niter_out = (cap > 0 ? cap : 0) + 1;
            publish(hm, 1, k - 1);
        }
        return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// set-up: D = diag(A) (prec 1) or 1 (prec 0), and ||b||^2 -> site A
// (Prec1.f90:45-48, 63-66; CG.f90:62-67).  With prec 2 the DIC pivots are
// computed before this kernel by k_dic_setup_level and write_d is false.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_setup(BatchCtx ctx, Csr<T> A, int prec, bool write_d, T* __restrict__ D,
                                                  const T* __restrict__ b, T* __restrict__ partA) {
    int rb, re;
    ctx_rows(ctx, A.n, rb, re);
    T v[1] = {T(0)};
    for (int i = rb + threadIdx.x; i < re; i += kBlock) {
        if (write_d) D[i] = (prec == 1) ? A.val[A.dia[i]] : T(1);
        const T bi = b[i];
        v[0] += bi * bi;
    }
    site_write<T, 1>(v, partA);
}

// ---------------------------------------------------------------------------
// DIC pivots for one level of rows (Prec1.f90:51-59):
//   D(i) = 1 / (a_ii - sum_{lower j} D(k) * a_ij * a_ij)
// Rows of a level depend only on rows of earlier levels.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dic_setup_level(Csr<T> A, const int* __restrict__ rows, int count,
                                                            T* __restrict__ D) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= count) return;
    const int i = rows[t];
    const int d = A.dia[i];
    T sum1 = A.val[d];
    for (int j = A.rowptr[i]; j < d; ++j) {
        const T a = A.val[j];
        sum1 = sum1 - D[A.col[j]] * a * a;
    }
    D[i] = T(1) / sum1;
}

// Forward sweep of the DIC apply for one level (CG.f90:135-142):
//   z(i) = (u(i) - sum_{lower} a_ij z(k)) * D(i)
// `u` and `z` may alias (each row reads its own u before writing its z).
// `st` (null for a batch) lets the sweep stop after convergence.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dic_fwd_level(Csr<T> A, const int* __restrict__ rows, int count,
                                                          const T* __restrict__ D, const T* u, T* z,
                                                          const SolverState* st) {
    if (st && st->done) return;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= count) return;
    const int i = rows[t];
    const int d = A.dia[i];
    T sum1 = u[i];
    for (int j = A.rowptr[i]; j < d; ++j) sum1 = sum1 - A.val[j] * z[A.col[j]];
    z[i] = sum1 * D[i];
}

// Backward sweep for one level, fused with the scaling pass of
// CG.f90:144-146 (each row scales its own forward value just before use,
// which performs the same two roundings as the separate pass):
//   z(i) = (z(i)/(D(i)+TINY) - sum_{upper} a_ij z(k)) * D(i)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dic_bwd_level(Csr<T> A, const int* __restrict__ rows, int count,
                                                          const T* __restrict__ D, T* z,
                                                          const SolverState* st) {
    if (st && st->done) return;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= count) return;
    const int i = rows[t];
    const int d = A.dia[i];
    const T di = D[i];
    T sum1 = z[i] / (di + T(kTiny));
    const int e = A.rowptr[i + 1];
    for (int j = d + 1; j < e; ++j) sum1 = sum1 - A.val[j] * z[A.col[j]];
    z[i] = sum1 * di;
}

// The same sweeps over the level-major copy of the triangles (rows of a
// forward level at consecutive positions, padding positions with row -1
// and no entries): gather u and D into position order, forward and
// backward sweeps with the arithmetic of k_dic_fwd_level /
// k_dic_bwd_level, scatter z back.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_plan_in(int n, const int* __restrict__ order, const T* __restrict__ u,
                                                    const T* __restrict__ D, T* __restrict__ up,
                                                    T* __restrict__ dp, const SolverState* st) {
    if (st && st->done) return;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const int i = order[k];
    up[k] = i < 0 ? T(0) : u[i];
    dp[k] = i < 0 ? T(0) : D[i];
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_plan_out(int n, const int* __restrict__ order, const T* __restrict__ zp,
                                                     T* __restrict__ z, const SolverState* st) {
    if (st && st->done) return;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const int i = order[k];
    if (i >= 0) z[i] = zp[k];
}

// ---------------------------------------------------------------------------
// One launch per triangle.  A value slot holds kEmpty until its row is
// solved; a row polls the slots of its dependencies with device-coherent
// loads and publishes its own value with a device-coherent store (the
// value is its own flag, so no fence is needed).  kEmpty is a signalling
// NaN pattern: arithmetic produces only quiet NaNs, so a solved value
// never equals it.  The grid is persistent and no larger than the blocks
// that fit on the device at once; block b takes the position chunks b,
// b + G, b + 2G, ... in that order, so each wavefront visits its 64-row
// pieces in increasing position order.  Rows only wait on smaller
// positions, and the rows of one wavefront piece share one forward level
// and never wait on each other, so the unfinished piece of smallest
// position always has its inputs and its wavefront at hand: every wait
// ends.
template <typename T>
struct Slot;
template <>
struct Slot<double> {
    using U = unsigned long long;
    static constexpr U kEmpty = 0x7FF4A5A5A5A5A5A5ull;
    static constexpr U kGaveUp = 0x7FF8000000000000ull;
};
template <>
struct Slot<float> {
    using U = unsigned int;
    static constexpr U kEmpty = 0x7FA5A5A5u;
    static constexpr U kGaveUp = 0x7FC00000u;
};

// poll interval of a waiting row (s_sleep units of 64 clocks) and the
// share of the resident blocks a one-launch sweep uses (1/kFlowShare)
#ifndef TFLOWS_FLOW_SLEEP
#define TFLOWS_FLOW_SLEEP 1
#endif
#ifndef TFLOWS_FLOW_SHARE
#define TFLOWS_FLOW_SHARE 2
#endif
constexpr int kFlowShare = TFLOWS_FLOW_SHARE;

template <typename T>
__device__ __forceinline__ void publish(T* p, T v) {
    using U = typename Slot<T>::U;
    __hip_atomic_store(reinterpret_cast<U*>(p), __builtin_bit_cast(U, v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ __forceinline__ typename Slot<T>::U peek(const T* p) {
    using U = typename Slot<T>::U;
    return __hip_atomic_load(reinterpret_cast<const U*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The value of slot p once solved, given a first reading v.  Every 64
// polls the wait checks the fault word and the time spent; past wait_ticks
// (or once another row gave up) it raises the fault word and returns a
// quiet NaN.
template <typename T>
__device__ __forceinline__ T settle(typename Slot<T>::U v, const T* p, int* fault, uint64_t t0,
                                    uint64_t wait_ticks) {
    using U = typename Slot<T>::U;
    const U* q = reinterpret_cast<const U*>(p);
    unsigned polls = 0;
    while (v == Slot<T>::kEmpty) {
        __builtin_amdgcn_s_sleep(TFLOWS_FLOW_SLEEP);
        if ((++polls & 63u) == 0) {
            const bool late = (uint64_t)wall_clock64() - t0 > wait_ticks;
            if (late || __hip_atomic_load(fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                __hip_atomic_store(fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                v = Slot<T>::kGaveUp;
                break;
            }
        }
        v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return __builtin_bit_cast(T, v);
}

// Every value slot of zf back to kEmpty (before the first sweep of a
// solve; afterwards the backward sweep empties each slot it has read).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_flow_reset(int n, T* __restrict__ zf) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < n) zf[k] = __builtin_bit_cast(T, Slot<T>::kEmpty);
}

// sum1 - sum_e a_e z_{c_e} over entries [e0, e1), in entry order.  Rows
// of up to kFlowRegs entries load their columns and values, then the
// first readings of all their dependencies, before the first wait.
constexpr int kFlowRegs = 4;
template <typename T>
__device__ __forceinline__ T flow_row(T sum1, int e0, int e1, const int* __restrict__ col,
                                      const T* __restrict__ val, const T* z, int* fault, uint64_t t0,
                                      uint64_t wait_ticks) {
    using U = typename Slot<T>::U;
    if (e1 - e0 <= kFlowRegs) {
        int c[kFlowRegs];
        T a[kFlowRegs];
        U v[kFlowRegs];
#pragma unroll
        for (int j = 0; j < kFlowRegs; ++j)
            if (e0 + j < e1) {
                c[j] = col[e0 + j];
                a[j] = val[e0 + j];
            }
#pragma unroll
        for (int j = 0; j < kFlowRegs; ++j)
            if (e0 + j < e1) v[j] = peek(z + c[j]);
#pragma unroll
        for (int j = 0; j < kFlowRegs; ++j)
            if (e0 + j < e1) sum1 = sum1 - a[j] * settle(v[j], z + c[j], fault, t0, wait_ticks);
        return sum1;
    }
    for (int e = e0; e < e1; ++e) {
        const T* p = z + col[e];
        sum1 = sum1 - val[e] * settle(peek(p), p, fault, t0, wait_ticks);
    }
    return sum1;
}

// forward sweep: z_k = (u_k - sum_{lower} a_kj z_j) * D_k, in entry order,
// with u and D gathered from the caller's rows (position k holds row
// order[k], -1 for padding); D is kept in position order for the backward
// sweep, and every zb slot is emptied for it (no one reads zb now).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_flow_fwd(int n, const int* __restrict__ order, const T* __restrict__ u,
                                                     const T* __restrict__ D, const int* __restrict__ lptr,
                                                     const int* __restrict__ lcol, const T* __restrict__ lval,
                                                     T* __restrict__ dp, T* zf, T* __restrict__ zb, int* fault,
                                                     uint64_t wait_ticks, const SolverState* st) {
    if (st && st->done) return;
    const uint64_t t0 = (uint64_t)wall_clock64();
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
        const int i = order[k];
        const T uk = i < 0 ? T(0) : u[i];
        const T dk = i < 0 ? T(0) : D[i];
        dp[k] = dk;
        zb[k] = __builtin_bit_cast(T, Slot<T>::kEmpty);
        const T sum1 = flow_row(uk, lptr[k], lptr[k + 1], lcol, lval, (const T*)zf, fault, t0, wait_ticks);
        publish(zf + k, sum1 * dk);
    }
}

// backward sweep over the positions in reverse:
//   z_k = (zf_k / (D_k + TINY) - sum_{upper} a_kj z_j) * D_k
// each row empties its zf slot (read only by itself here) for the next
// forward sweep and stores its result in the caller's row as well.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_flow_bwd(int n, const int* __restrict__ order, T* __restrict__ z,
                                                     const int* __restrict__ uptr, const int* __restrict__ ucol,
                                                     const T* __restrict__ uval, const T* __restrict__ dp,
                                                     T* __restrict__ zf, T* zb, int* fault, uint64_t wait_ticks,
                                                     const SolverState* st) {
    if (st && st->done) return;
    const uint64_t t0 = (uint64_t)wall_clock64();
    for (int t = blockIdx.x * kBlock + threadIdx.x; t < n; t += gridDim.x * kBlock) {
        const int k = n - 1 - t;
        const T di = dp[k];
        const T zk = zf[k];
        zf[k] = __builtin_bit_cast(T, Slot<T>::kEmpty);
        const T sum1 = flow_row(zk / (di + T(kTiny)), uptr[k], uptr[k + 1], ucol, uval, (const T*)zb, fault, t0,
                                wait_ticks);
        const T v = sum1 * di;
        publish(zb + k, v);
        const int i = order[k];
        if (i >= 0) z[i] = v;
    }
}

// Blocks of a one-launch sweep: a share of those that fit on the device
// at once (the rest is headroom), and no more than the position chunks.
template <typename T>
int flow_grid(int n) {
    static int resident = [] {
        auto check = [](hipError_t e) {
            if (e != hipSuccess) throw std::runtime_error(std::string("sweep grid: ") + hipGetErrorString(e));
        };
        int dev = 0, cus = 0, per_cu = 0, per_cu_b = 0;
        check(hipGetDevice(&dev));
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_flow_fwd<T>, kBlock, 0));
        check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_b, k_flow_bwd<T>, kBlock, 0));
        per_cu = std::min(per_cu, per_cu_b);
        return std::max(1, cus * per_cu / kFlowShare);
    }();
    return std::max(1, std::min(resident, (n + kBlock - 1) / kBlock));
}
// ---------------------------------------------------------------------------
// Tiled one-launch sweeps (Levels mode 3; tiles as tile_plan.h builds them).
// Each block is one wavefront; all blocks are resident.  Wavefront w sweeps
// the tiles of its columns, listed in increasing index in wtiles[wstart[w]
// .. wstart[w+1]) (backward: the same lists in reverse), so it climbs each
// column level chunk by level chunk and its predecessor in the column is the
// tile it has just swept.  For a tile it issues the loads of all its phases
// at once (row, entry codes and values, u and D), reads the values it needs
// from other tiles kTileAhead phases before the phase that uses them, and
// runs the phases in order, one row per lane, with the tile's own values in
// LDS; a value not yet published when read is waited for in the phase that
// uses it, as the per-row sweeps wait.  Every wait ends: the unfinished tile
// of smallest index (largest, backward) has all its inputs published, and
// its wavefront has finished every tile of its list before it.  The
// arithmetic of each row is that of k_dic_fwd_level / k_dic_bwd_level /
// k_dic_setup_level, in entry order.
//
// The phases are straight-line code: a lane without a row in a phase loads
// from a row of the tile that exists, computes a value nobody reads, and
// stores it to a spare LDS slot and to a spare slot past the n positions of
// the slot arrays (which hold n + kTileWave entries); only its store to the
// caller's rows (z, D) is skipped.  Without branches around the loads the
// compiler counts the outstanding loads exactly and waits for each value
// alone.
constexpr int kTileAhead = 2;

// LDS stores of one phase are seen by the loads of the next: the wavefront's
// LDS operations execute in program order; the fences keep the compiler from
// moving them across the phase boundary.
__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The slow path of a tile's wait for another tile's value (out of line, so
// the phases stay straight-line): re-reads the slot until it is published,
// or, past wait_ticks or once another row gave up, raises the fault word
// and returns a quiet NaN.
template <typename T>
__device__ __noinline__ typename Slot<T>::U tile_wait(const T* p, int* fault, uint64_t t0, uint64_t wait_ticks) {
    return __builtin_bit_cast(typename Slot<T>::U, settle<T>(peek(p), p, fault, t0, wait_ticks));
}

// sum1 - sum_j a_j z_j over the E slots of one row: a code c >= 0 is the
// value v read from another tile, c < 0 the tile's own value in LDS, a
// padding slot (kTileNone, after the row's entries) leaves the sum as is
template <typename T, int E>
__device__ __forceinline__ T tile_row(T sum1, const int (&c)[E], const T (&a)[E], typename Slot<T>::U (&v)[E],
                                      const T* zl, const T* zs, int* fault, uint64_t t0, uint64_t wait_ticks) {
    bool empty = false;
#pragma unroll
    for (int j = 0; j < E; ++j) empty = empty || (c[j] >= 0 && v[j] == Slot<T>::kEmpty);
    if (__builtin_expect(empty, 0)) {
#pragma unroll
        for (int j = 0; j < E; ++j)
            if (c[j] >= 0 && v[j] == Slot<T>::kEmpty) v[j] = tile_wait<T>(zs + c[j], fault, t0, wait_ticks);
    }
#pragma unroll
    for (int j = 0; j < E; ++j) {
        const int cj = c[j];
        const T own = zl[(cj < 0 && cj != kTileNone) ? -cj - 1 : 0];
        const T zj = cj >= 0 ? __builtin_bit_cast(T, v[j]) : own;
        const T s2 = sum1 - a[j] * zj;
        sum1 = cj == kTileNone ? sum1 : s2;
    }
    return sum1;
}

// forward sweep (z_k = (u_k - sum_lower a z) * D_k) into zf, with dp = D in
// position order and every zb slot emptied for the backward sweep
template <typename T, int E>
__global__ __launch_bounds__(kTileWave) void k_tile_fwd(int n, const int* __restrict__ wstart,
                                                        const int* __restrict__ wtiles, const int* __restrict__ tph,
                                                        const int* __restrict__ order, const T* __restrict__ u,
                                                        const T* __restrict__ D, const int* __restrict__ lcol,
                                                        const T* __restrict__ lval, T* __restrict__ dp, T* zf,
                                                        T* __restrict__ zb, int* fault, uint64_t wait_ticks,
                                                        const SolverState* st) {
    if (st && st->done) return;
    using U = typename Slot<T>::U;
    __shared__ T zl[kTileMaxRows + kTileWave];
    const int lane = threadIdx.x;
    const uint64_t t0 = (uint64_t)wall_clock64();
    const int i1 = wstart[blockIdx.x + 1];
    for (int i = wstart[blockIdx.x]; i < i1; ++i) {
        const int* ph = tph + (size_t)wtiles[i] * (kTilePhases + 1);
        int bnd[kTilePhases + 1];
#pragma unroll
        for (int p = 0; p <= kTilePhases; ++p) bnd[p] = ph[p];
        const int base = bnd[0];
        bool ok[kTilePhases];
        int k[kTilePhases], c[kTilePhases][E];
        T a[kTilePhases][E], uu[kTilePhases], dd[kTilePhases];
        U v[kTilePhases][E];
#pragma unroll
        for (int p = 0; p < kTilePhases; ++p) {
            ok[p] = bnd[p] + lane < bnd[p + 1];
            k[p] = ok[p] ? bnd[p] + lane : bnd[kTilePhases] - 1;
#pragma unroll
            for (int j = 0; j < E; ++j) {
                c[p][j] = lcol[(size_t)j * n + k[p]];
                a[p][j] = lval[(size_t)j * n + k[p]];
            }
        }
#pragma unroll
        for (int p = 0; p < kTilePhases; ++p) {
            const int r = order[k[p]];
            uu[p] = u[r];
            dd[p] = D[r];
        }
#pragma unroll
        for (int p = 0; p < kTileAhead; ++p)
#pragma unroll
            for (int j = 0; j < E; ++j) v[p][j] = peek(zf + (c[p][j] >= 0 ? c[p][j] : 0));
#pragma unroll
        for (int p = 0; p < kTilePhases; ++p) {
            if (p + kTileAhead < kTilePhases) {
#pragma unroll
                for (int j = 0; j < E; ++j)
                    v[p + kTileAhead][j] = peek(zf + (c[p + kTileAhead][j] >= 0 ? c[p + kTileAhead][j] : 0));
            }
            const T zk = tile_row<T, E>(uu[p], c[p], a[p], v[p], zl, zf, fault, t0, wait_ticks) * dd[p];
            const int kp = ok[p] ? k[p] : n + lane;
            zl[ok[p] ? kp - base : kTileMaxRows + lane] = zk;
            dp[kp] = dd[p];
            zb[kp] = __builtin_bit_cast(T, Slot<T>::kEmpty);
            publish(zf + kp, zk);
            wave_lds_order();
        }
    }
}

// backward sweep (z_k = (zf_k / (D_k + TINY) - sum_upper a z) * D_k) over
// the tiles and phases in reverse, into zb and the caller's rows of z; each
// row empties its zf slot for the next forward sweep
template <typename T, int E>
__global__ __launch_bounds__(kTileWave) void k_tile_bwd(int n, const int* __restrict__ wstart,
                                                        const int* __restrict__ wtiles, const int* __restrict__ tph,
                                                        const int* __restrict__ order, T* __restrict__ z,
                                                        const int* __restrict__ ucol, const T* __restrict__ uval,
                                                        const T* __restrict__ dp, T* __restrict__ zf, T* zb,
                                                        int* fault, uint64_t wait_ticks, const SolverState* st) {
    if (st && st->done) return;
    using U = typename Slot<T>::U;
    __shared__ T zl[kTileMaxRows + kTileWave];
    const int lane = threadIdx.x;
    const uint64_t t0 = (uint64_t)wall_clock64();
    const int i0 = wstart[blockIdx.x];
    for (int i = wstart[blockIdx.x + 1] - 1; i >= i0; --i) {
        const int* ph = tph + (size_t)wtiles[i] * (kTilePhases + 1);
        int bnd[kTilePhases + 1];
#pragma unroll
        for (int p = 0; p <= kTilePhases; ++p) bnd[p] = ph[p];
        const int base = bnd[0];
        bool ok[kTilePhases];
        int k[kTilePhases], row[kTilePhases], c[kTilePhases][E];
        T a[kTilePhases][E], zk[kTilePhases], dd[kTilePhases];
        U v[kTilePhases][E];
#pragma unroll
        for (int p = 0; p < kTilePhases; ++p) {
            ok[p] = bnd[p] + lane < bnd[p + 1];
            k[p] = ok[p] ? bnd[p] + lane : base;
            row[p] = order[k[p]];
            zk[p] = zf[k[p]];
            dd[p] = dp[k[p]];
#pragma unroll
            for (int j = 0; j < E; ++j) {
                c[p][j] = ucol[(size_t)j * n + k[p]];
                a[p][j] = uval[(size_t)j * n + k[p]];
            }
        }
#pragma unroll
        for (int p = kTilePhases - 1; p >= kTilePhases - kTileAhead; --p)
#pragma unroll
            for (int j = 0; j < E; ++j) v[p][j] = peek(zb + (c[p][j] >= 0 ? c[p][j] : 0));
#pragma unroll
        for (int p = kTilePhases - 1; p >= 0; --p) {
            if (p - kTileAhead >= 0) {
#pragma unroll
                for (int j = 0; j < E; ++j)
                    v[p - kTileAhead][j] = peek(zb + (c[p - kTileAhead][j] >= 0 ? c[p - kTileAhead][j] : 0));
            }
            const int kp = ok[p] ? k[p] : n + lane;
            zf[kp] = __builtin_bit_cast(T, Slot<T>::kEmpty);
            const T di = dd[p];
            const T vk = tile_row<T, E>(zk[p] / (di + T(kTiny)), c[p], a[p], v[p], zl, zb, fault, t0, wait_ticks) * di;
            zl[ok[p] ? kp - base : kTileMaxRows + lane] = vk;
            publish(zb + kp, vk);
            if (ok[p]) z[row[p]] = vk;
            wave_lds_order();
        }
    }
}

// DIC pivots (Prec1.f90:51-59): D_i = 1 / (a_ii - sum_lower (D_k a) a), from
// the lower values and the diagonal in position order (lval, dg) into D and
// the slots zs, every one of which is empty at the start
template <int E>
__global__ __launch_bounds__(kTileWave) void k_tile_pivots(int n, const int* __restrict__ wstart,
                                                           const int* __restrict__ wtiles,
                                                           const int* __restrict__ tph,
                                                           const int* __restrict__ order,
                                                           const double* __restrict__ dg,
                                                           const int* __restrict__ lcol,
                                                           const double* __restrict__ lval, double* __restrict__ D,
                                                           double* zs, int* fault, uint64_t wait_ticks) {
    using U = typename Slot<double>::U;
    __shared__ double zl[kTileMaxRows + kTileWave];
    const int lane = threadIdx.x;
    const uint64_t t0 = (uint64_t)wall_clock64();
    const int i1 = wstart[blockIdx.x + 1];
    for (int i = wstart[blockIdx.x]; i < i1; ++i) {
        const int* ph = tph + (size_t)wtiles[i] * (kTilePhases + 1);
        int bnd[kTilePhases + 1];
#pragma unroll
        for (int p = 0; p <= kTilePhases; ++p) bnd[p] = ph[p];
        const int base = bnd[0];
        bool ok[kTilePhases];
        int k[kTilePhases], row[kTilePhases], c[kTilePhases][E];
        double a[kTilePhases][E], dd[kTilePhases];
        U v[kTilePhases][E];
#pragma unroll
        for (int p = 0; p < kTilePhases; ++p) {
            ok[p] = bnd[p] + lane < bnd[p + 1];
            k[p] = ok[p] ? bnd[p] + lane : bnd[kTilePhases] - 1;
            row[p] = order[k[p]];
            dd[p] = dg[k[p]];
#pragma unroll
            for (int j = 0; j < E; ++j) {
                c[p][j] = lcol[(size_t)j * n + k[p]];
                a[p][j] = lval[(size_t)j * n + k[p]];
            }
        }
#pragma unroll
        for (int p = 0; p < kTileAhead; ++p)
#pragma unroll
            for (int j = 0; j < E; ++j) v[p][j] = peek(zs + (c[p][j] >= 0 ? c[p][j] : 0));
#pragma unroll
        for (int p = 0; p < kTilePhases; ++p) {
            if (p + kTileAhead < kTilePhases) {
#pragma unroll
                for (int j = 0; j < E; ++j)
                    v[p + kTileAhead][j] = peek(zs + (c[p + kTileAhead][j] >= 0 ? c[p + kTileAhead][j] : 0));
            }
            // sum1 - (D_k a) a: the row form with the coefficients D_k a and a
            bool empty = false;
#pragma unroll
            for (int j = 0; j < E; ++j) empty = empty || (c[p][j] >= 0 && v[p][j] == Slot<double>::kEmpty);
            if (__builtin_expect(empty, 0)) {
#pragma unroll
                for (int j = 0; j < E; ++j)
                    if (c[p][j] >= 0 && v[p][j] == Slot<double>::kEmpty)
                        v[p][j] = tile_wait<double>(zs + c[p][j], fault, t0, wait_ticks);
            }
            double sum1 = dd[p];
#pragma unroll
            for (int j = 0; j < E; ++j) {
                const int cj = c[p][j];
                const double own = zl[(cj < 0 && cj != kTileNone) ? -cj - 1 : 0];
                const double dj = cj >= 0 ? __builtin_bit_cast(double, v[p][j]) : own;
                const double s2 = sum1 - dj * a[p][j] * a[p][j];
                sum1 = cj == kTileNone ? sum1 : s2;
            }
            const double dk = 1.0 / sum1;
            const int kp = ok[p] ? k[p] : n + lane;
            zl[ok[p] ? kp - base : kTileMaxRows + lane] = dk;
            if (ok[p]) D[row[p]] = dk;
            publish(zs + kp, dk);
            wave_lds_order();
        }
    }
}

// Wavefronts of the tiled launches: a share of those that fit on the device
// at once, for the kernel of the lowest occupancy among the tiled sweeps
// and pivots (every tiled launch runs the one schedule built for it).
static int tile_resident(const void* k) {
    auto check = [](hipError_t e) {
        if (e != hipSuccess) throw std::runtime_error(std::string("tiled sweep grid: ") + hipGetErrorString(e));
    };
    int dev = 0, cus = 0, per_cu = 0;
    check(hipGetDevice(&dev));
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kTileWave, 0));
    return std::max(1, cus * per_cu / kFlowShare);
}
template <int E>
static int tile_waves_for() {
    static const int r = [] {
        const void* ks[] = {(const void*)k_tile_fwd<double, E>, (const void*)k_tile_bwd<double, E>,
                            (const void*)k_tile_fwd<float, E>, (const void*)k_tile_bwd<float, E>,
                            (const void*)k_tile_pivots<E>};
        int m = INT_MAX;
        for (const void* k : ks) m = std::min(m, tile_resident(k));
        return m;
    }();
    return r;
}
int tile_sweep_waves(int slots) { return slots == 3 ? tile_waves_for<3>() : tile_waves_for<4>(); }

template <typename T>
__global__ __launch_bounds__(kBlock) void k_plan_fwd(int a, int b, const int* __restrict__ lptr,
                                                     const int* __restrict__ lcol, const T* __restrict__ lval,
                                                     const T* __restrict__ dp, const T* __restrict__ up, T* zp,
                                                     const SolverState* st) {
    if (st && st->done) return;
    const int k = a + blockIdx.x * kBlock + threadIdx.x;
    if (k >= b) return;
    T sum1 = up[k];
    for (int e = lptr[k]; e < lptr[k + 1]; ++e) sum1 = sum1 - lval[e] * zp[lcol[e]];
    zp[k] = sum1 * dp[k];
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_plan_bwd(const int* __restrict__ pos, int count,
                                                     const int* __restrict__ uptr, const int* __restrict__ ucol,
                                                     const T* __restrict__ uval, const T* __restrict__ dp, T* zp,
                                                     const SolverState* st) {
    if (st && st->done) return;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= count) return;
    const int k = pos[t];
    const T di = dp[k];
    T sum1 = zp[k] / (di + T(kTiny));
    for (int e = uptr[k]; e < uptr[k + 1]; ++e) sum1 = sum1 - uval[e] * zp[ucol[e]];
    zp[k] = sum1 * di;
}

__global__ __launch_bounds__(kBlock) void k_entry_gather(int count, const int* __restrict__ map,
                                                         const double* __restrict__ src, double* __restrict__ dst) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e < count) {
        const int m = map[e];
        dst[e] = m < 0 ? 0.0 : src[m];
    }
}

// ---------------------------------------------------------------------------
// Early exit on ||b|| < tol (CG.f90:69-72) from site A, then the initial
// residual r = b - A x in place (Resid.f90:29-34: subtraction in column
// order starting from b) with ||r||^2 and the first rho -> site B:
//   RHO_JAC : rho = sum r*(r/D)          (CG with prec 0/1)
//   RHO_BICG: rho = sum (r/D)*r2, r2 = r (BiCG with prec 0/1)
//   RHO_RR  : rho = sum r*r2,     r2 = r (CGS, any prec)
//   RHO_NONE: rho comes later from the DIC z (CG/BiCG with prec 2)
// r2 (if not null) receives a copy of r (BiCG.f90:93-95, CGS.f90:84-86).
// ---------------------------------------------------------------------------
enum RhoMode { RHO_NONE = 0, RHO_JAC = 1, RHO_BICG = 2, RHO_RR = 3 };

template <typename T, int MODE>
struct ResidEpi {
    T* __restrict__ r;
    T* __restrict__ r2;
    const T* __restrict__ D;
    T acc[2];
    __device__ void operator()(int row, T (&y)[1]) {
        const T ri = y[0];
        r[row] = ri;
        if (r2) r2[row] = ri;
        acc[0] += ri * ri;
        if (MODE == RHO_JAC) acc[1] += ri * jac(ri, D[row]);
        if (MODE == RHO_BICG) acc[1] += jac(ri, D[row]) * ri;
        if (MODE == RHO_RR) acc[1] += ri * ri;
    }
};

template <typename T, int MODE, bool ZERO_X, int L>
__global__ __launch_bounds__(kBlock) void k_resid_init(BatchCtx ctx, Csr<T> A, const T* __restrict__ x,
                                                       T* __restrict__ r, T* __restrict__ r2,
                                                       const T* __restrict__ D, const T* __restrict__ partA,
                                                       T* __restrict__ partB, SolverState* st_base,
                                                       HostMirror* hm_base) {
    TF_STATE;
    T bb[1];
    ctx_total<T, 1>(ctx, partA, bb);
    const double bn = sqrt((double)bb[0]);
    ResidEpi<T, MODE> epi{r, r2, D, {T(0), T(0)}};
    if (bn < st->tol) {
        if (ctx.leader()) {
            st->bnrm2 = bn;
            st->done = 1;
            st->niter_out = 0;
            publish(hm, 1, 0);
        }
        site_write<T, 2>(epi.acc, partB);   // zeros: the system is over
        return;
    }
    if (ctx.leader()) st->bnrm2 = bn;
    if (ZERO_X) {  // x = 0: r = b - A*0 = b, no matrix pass
        int rb, re;
        ctx_rows(ctx, A.n, rb, re);
        for (int i = rb + threadIdx.x; i < re; i += kBlock) {
            T y[1] = {r[i]};
            epi(i, y);
        }
    } else {
        const TileSpan sp = ctx_span(ctx, A.n);
        Gather1<T> g{x};
        InitVec<T> init{r};
        for (int t = sp.first; t < sp.end; t += sp.step) {
            int r0, r1;
            ctx_tile_rows(ctx, A.n, t, r0, r1);
            resid_tile<L>(A, r0, r1, g, init, epi);
        }
    }
    site_write<T, 2>(epi.acc, partB);
}

// Stand-alone stopping test (DIC forms start their iteration with it, and
// every solve ends with it at k = cap + 1).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_check(BatchCtx ctx, int k_arg, const T* __restrict__ partB,
                                                  SolverState* st_base, HostMirror* hm_base) {
    TF_STATE;
    if (st->done) return;
    T tot[2];
    check_stop<T>(ctx, ctx.k(k_arg), partB, st, hm, tot);
}

// rho = sum a_i*b_i -> site C, for the DIC forms (CG.f90:172-176: (r,z);
// BiCG.f90:169-173: (z1,r2)); the product is a*b in the reference's order.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_rho(BatchCtx ctx, int n, const T* __restrict__ a,
                                                const T* __restrict__ b, T* __restrict__ partC,
                                                const SolverState* st_base) {
    const SolverState* st = st_base + ctx.sys();
    T v[2] = {T(0), T(0)};
    if (!st->done) {
        int rb, re;
        ctx_rows(ctx, n, rb, re);
        for (int i = rb + threadIdx.x; i < re; i += kBlock) v[0] += a[i] * b[i];
    }
    site_write<T, 2>(v, partC);
}

// rho_k of iteration k: from site B (Jacobi forms, after the stopping test)
// or site C (DIC forms); recorded by the leader in st->rho[k & 1].
template <typename T, bool JAC>
__device__ __forceinline__ bool iteration_rho(const BatchCtx& ctx, int k, const T* partB, const T* partC,
                                              SolverState* st, HostMirror* hm, T& rho, const int* halt = nullptr) {
    T tot[2];
    if (JAC) {
        if (check_stop<T>(ctx, k, partB, st, hm, tot, halt)) return false;
        rho = tot[1];
    } else {
        ctx_total<T, 2>(ctx, partC, tot);
        if (halt && *halt) return false;
        rho = tot[0];
    }
    if (ctx.leader()) st->rho[k & 1] = (double)rho;
    return true;
}

// In a batch, blocks of a finished system still write their (zero) partial
// sums so that every partial a consumer reads is defined; for one system a
// finished solve simply returns.
#define TF_DONE_RETURN(site, NV)                 \
    if (st->done) {                              \
        if (ctx.blk_sys) {                       \
            T z_[NV] = {};                       \
            site_write<T, NV>(z_, site);         \
        }                                        \
        return;                                  \
    }

// ===========================================================================
// CG  (Process/CG.f90:113-259)
// ===========================================================================
// [stopping test,] rho, p = z + beta*p  (z = r/D, or the DIC output z)
// CG.f90:172-187.  DEFER (one system): before p is overwritten, the
// x += alfa p of iteration k - 1 (CG.f90:226-227), whose update kernel left
// alfa in the state; the pass then moves p once instead of twice.
template <typename T, bool JAC, bool DEFER>
__global__ __launch_bounds__(kBlock) void k_cg_p(BatchCtx ctx, int k_arg, int n, const T* __restrict__ r,
                                                 const T* __restrict__ D, const T* __restrict__ z,
                                                 T* __restrict__ p, T* __restrict__ x,
                                                 const T* __restrict__ partB, const T* __restrict__ partC,
                                                 SolverState* st_base, HostMirror* hm_base) {
    TF_STATE;
    // done is read once per block (its leader may set it in this launch)
    // and tested after the reduction, so that it, rho_{k-1}, the first
    // kPrefetch rows of this thread and the partial sums are loaded together
    __shared__ int halt;
    if (threadIdx.x == 0) halt = st->done;
    const int k = ctx.k(k_arg);
    const double rho_prev = st->rho[(k - 1) & 1];
    const bool carry = DEFER && k > 1;
    const T alfa_prev = DEFER ? (T)st->alfa : T(0);
    int rb, re;
    ctx_rows(ctx, n, rb, re);
    T av[kPrefetch], dv[kPrefetch], pv[kPrefetch], xv[kPrefetch];
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
        const int i = rb + (int)threadIdx.x + u * kBlock;
        const bool ok = i < re;
        av[u] = ok ? (JAC ? r[i] : z[i]) : T(0);
        dv[u] = (JAC && ok) ? D[i] : T(1);
        pv[u] = ok ? p[i] : T(0);
        xv[u] = (carry && ok) ? x[i] : T(0);
    }
    T rho;
    if (!iteration_rho<T, JAC>(ctx, k, partB, partC, st, hm, rho, &halt)) return;
    const bool first = k == 1;
    const T beta = first ? T(0) : rho / (T)rho_prev;
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
        const int i = rb + (int)threadIdx.x + u * kBlock;
        if (i < re) {
            if (carry) x[i] = xv[u] + alfa_prev * pv[u];
            const T zi = JAC ? jac(av[u], dv[u]) : av[u];
            p[i] = first ? zi : zi + beta * pv[u];
        }
    }
    for (int i = rb + (int)threadIdx.x + kPrefetch * kBlock; i < re; i += kBlock) {
        const T zi = JAC ? jac(r[i], D[i]) : z[i];
        const T pi = p[i];
        if (carry) x[i] = x[i] + alfa_prev * pi;
        p[i] = first ? zi : zi + beta * pi;
    }
}

// q = A p (CG.f90:193-199) with (p,q) -> site D (:214-218)
template <typename T>
struct DotEpi {
    const T* __restrict__ w;   // dot partner
    T* __restrict__ q;         // output (may be null)
    T acc[1];
    __device__ void operator()(int row, T (&y)[1]) {
        if (q) q[row] = y[0];
        acc[0] += w[row] * y[0];
    }
};

template <typename Op, int L, typename T = typename OpValue<Op>::type>
__global__ __launch_bounds__(kBlock) void k_cg_spmv(BatchCtx ctx, Op A, const T* __restrict__ p, T* __restrict__ q,
                                                    T* __restrict__ partD, const SolverState* st_base) {
    const SolverState* st = st_base + ctx.sys();
    TF_DONE_RETURN(partD, 1);
    const TileSpan sp = ctx_span(ctx, A.n);
    Gather1<T> g{p};
    DotEpi<T> epi{p, q, {T(0)}};
    for (int t = sp.first; t < sp.end; t += sp.step) {
        int r0, r1;
        ctx_tile_rows(ctx, A.n, t, r0, r1);
        op_tile<1, L>(A, r0, r1, g, epi);
    }
    site_write<T, 1>(epi.acc, partD);
}

// alfa = rho/(p,q) (:219); x += alfa p; r -= alfa q (CG.f90:226-229), with
// ||r||^2 and, for the Jacobi forms, the next rho = (r, r/D) -> site B.
// DEFER (one system): x is left to the next p kernel; alfa and k go to the
// state for it.
template <typename T, bool JAC, bool DEFER>
__global__ __launch_bounds__(kBlock) void k_cg_update(BatchCtx ctx, int k_arg, int n, T* __restrict__ x,
                                                      T* __restrict__ r, const T* __restrict__ p,
                                                      const T* __restrict__ q, const T* __restrict__ D,
                                                      const T* __restrict__ partD, T* __restrict__ partB,
                                                      SolverState* st_base) {
    SolverState* st = st_base + ctx.sys();
    // done is set by earlier kernels only, so this return is the same in
    // every thread; rho_k is loaded with the first rows and the partial sums
    // of (p,q) (testing done after them as well costs this kernel its
    // eighth wave per SIMD)
    TF_DONE_RETURN(partB, 2);
    const int k = ctx.k(k_arg);
    const double rho_k = st->rho[k & 1];
    int rb, re;
    ctx_rows(ctx, n, rb, re);
    T xv[kPrefetch], pv[kPrefetch], rv[kPrefetch], qv[kPrefetch], dv[kPrefetch];
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
        const int i = rb + (int)threadIdx.x + u * kBlock;
        const bool ok = i < re;
        xv[u] = (!DEFER && ok) ? x[i] : T(0);
        pv[u] = (!DEFER && ok) ? p[i] : T(0);
        rv[u] = ok ? r[i] : T(0);
        qv[u] = ok ? q[i] : T(0);
        dv[u] = (JAC && ok) ? D[i] : T(1);
    }
    T pq[1];
    ctx_total<T, 1>(ctx, partD, pq);
    const T alfa = (T)rho_k / pq[0];
    if (DEFER && ctx.leader()) {
        st->alfa = (double)alfa;
        st->it = k;
    }
    T v[2] = {T(0), T(0)};
#pragma unroll
    for (int u = 0; u < kPrefetch; ++u) {
        const int i = rb + (int)threadIdx.x + u * kBlock;
        if (i < re) {
            if (!DEFER) x[i] = xv[u] + alfa * pv[u];
            const T ri = rv[u] - alfa * qv[u];
            r[i] = ri;
            v[0] += ri * ri;
            if (JAC) v[1] += ri * jac(ri, dv[u]);
        }
    }
    for (int i = rb + (int)threadIdx.x + kPrefetch * kBlock; i < re; i += kBlock) {
        if (!DEFER) x[i] = x[i] + alfa * p[i];
        const T ri = r[i] - alfa * q[i];
        r[i] = ri;
        v[0] += ri * ri;
        if (JAC) v[1] += ri * jac(ri, D[i]);
    }
    site_write<T, 2>(v, partB);
}

// The update x += alfa p (CG.f90:226-227) that the last iteration left to a
// p kernel which did not run its rows (the solve stopped first); nothing
// when a merge has applied it already or no iteration ran.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_xfinish(int n, T* __restrict__ x, const T* __restrict__ p,
                                                       const SolverState* st) {
    if (st->it <= st->xupd) return;
    const T alfa = (T)st->alfa;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] = x[i] + alfa * p[i];
}

// ===========================================================================
// BiCG  (Process/BiCG.f90:100-251).  The shadow system is multiplied by A
// (BiCG.f90:196-200) and starts from r2 = r1, exactly as in the reference.
// ===========================================================================
template <typename T, bool JAC>
__global__ __launch_bounds__(kBlock) void k_bicg_p(BatchCtx ctx, int k_arg, int n, const T* __restrict__ r1,
                                                   const T* __restrict__ r2, const T* __restrict__ D,
                                                   const T* __restrict__ z1, const T* __restrict__ z2,
                                                   T* __restrict__ p1, T* __restrict__ p2,
                                                   const T* __restrict__ partB, const T* __restrict__ partC,
                                                   SolverState* st_base, HostMirror* hm_base) {
    TF_STATE;
    if (st->done) return;
    const int k = ctx.k(k_arg);
    T rho;
    if (!iteration_rho<T, JAC>(ctx, k, partB, partC, st, hm, rho)) return;
    const bool first = k == 1;
    const T beta = first ? T(0) : rho / (T)st->rho[(k - 1) & 1];
    int rb, re;
    ctx_rows(ctx, n, rb, re);
    for (int i = rb + threadIdx.x; i < re; i += kBlock) {
        const T a = JAC ? jac(r1[i], D[i]) : z1[i];
        const T b = JAC ? jac(r2[i], D[i]) : z2[i];
        p1[i] = first ? a : a + beta * p1[i];
        p2[i] = first ? b : b + beta * p2[i];
    }
}

template <typename T>
struct BiDotEpi {
    const T* __restrict__ p2;
    T* __restrict__ q1;
    T* __restrict__ q2;
    T acc[1];
    __device__ void operator()(int row, T (&y)[2]) {
        q1[row] = y[0];
        q2[row] = y[1];
        acc[0] += p2[row] * y[0];          // (p2, q1)  BiCG.f90:218-222
    }
};

template <typename Op, int L, typename T = typename OpValue<Op>::type>
__global__ __launch_bounds__(kBlock) void k_bicg_spmv(BatchCtx ctx, Op A, const T* __restrict__ p1,
                                                      const T* __restrict__ p2, T* __restrict__ q1,
                                                      T* __restrict__ q2, T* __restrict__ partD,
                                                      const SolverState* st_base) {
    const SolverState* st = st_base + ctx.sys();
    TF_DONE_RETURN(partD, 1);
    const TileSpan sp = ctx_span(ctx, A.n);
    Gather2<T> g{p1, p2};
    BiDotEpi<T> epi{p2, q1, q2, {T(0)}};
    for (int t = sp.first; t < sp.end; t += sp.step) {
        int r0, r1;
        ctx_tile_rows(ctx, A.n, t, r0, r1);
        op_tile<2, L>(A, r0, r1, g, epi);
    }
    site_write<T, 1>(epi.acc, partD);
}

template <typename T, bool JAC>
__global__ __launch_bounds__(kBlock) void k_bicg_update(BatchCtx ctx, int k_arg, int n, T* __restrict__ x,
                                                        T* __restrict__ r1, T* __restrict__ r2,
                                                        const T* __restrict__ p1, const T* __restrict__ q1,
                                                        const T* __restrict__ q2, const T* __restrict__ D,
                                                        const T* __restrict__ partD, T* __restrict__ partB,
                                                        const SolverState* st_base) {
    const SolverState* st = st_base + ctx.sys();
    TF_DONE_RETURN(partB, 2);
    const int k = ctx.k(k_arg);
    T pq[1];
    ctx_total<T, 1>(ctx, partD, pq);
    const T alfa = (T)st->rho[k & 1] / pq[0];        // BiCG.f90:223
    int rb, re;
    ctx_rows(ctx, n, rb, re);
    T v[2] = {T(0), T(0)};
    for (int i = rb + threadIdx.x; i < re; i += kBlock) {
        x[i] = x[i] + alfa * p1[i];                   // BiCG.f90:230-234
        const T a = r1[i] - alfa * q1[i];
        const T b = r2[i] - alfa * q2[i];
        r1[i] = a;
        r2[i] = b;
        v[0] += a * a;
        if (JAC) v[1] += jac(a, D[i]) * b;            // next rho = (z1, r2)
    }
    site_write<T, 2>(v, partB);
}

// ===========================================================================
// CGS  (Process/CGS.f90:112-334), fixed shadow r2 = r0.
// ===========================================================================
// stopping test; rho = (r1, r2) from site B; u1 = r1 + beta q1;
// u2 = u1 + beta (q1 + beta u2); p2 = u2/D  (:118-148)
template <typename T, bool JAC>
__global__ __launch_bounds__(kBlock) void k_cgs_u(BatchCtx ctx, int k_arg, int n, const T* __restrict__ r1,
                                                  const T* __restrict__ q1, T* __restrict__ u1,
                                                  T* __restrict__ u2, const T* __restrict__ D,
                                                  T* __restrict__ p2, const T* __restrict__ partB,
                                                  SolverState* st_base, HostMirror* hm_base) {
    TF_STATE;
    if (st->done) return;
    const int k = ctx.k(k_arg);
    T tot[2];
    if (check_stop<T>(ctx, k, partB, st, hm, tot)) return;
    const T rho = tot[1];
    if (ctx.leader()) st->rho[k & 1] = (double)rho;
    const bool first = k == 1;
    const T beta = first ? T(0) : rho / (T)st->rho[(k - 1) & 1];
    int rb, re;
    ctx_rows(ctx, n, rb, re);
    for (int i = rb + threadIdx.x; i < re; i += kBlock) {
        T a, b;
        if (first) {
            a = r1[i];
            b = a;
        } else {
            const T qi = q1[i];
            a = r1[i] + beta * qi;
            b = a + beta * (qi + beta * u2[i]);
        }
        u1[i] = a;
        u2[i] = b;
        if (JAC) p2[i] = jac(b, D[i]);
    }
}

// v2 = A p2 (:193-198) with (r2, v2) -> site D (:215-219)
template <typename Op, int L, typename T = typename OpValue<Op>::type>
__global__ __launch_bounds__(kBlock) void k_cgs_spmv_v(BatchCtx ctx, Op A, const T* __restrict__ p2,
                                                       T* __restrict__ v2, const T* __restrict__ r2,
                                                       T* __restrict__ partD, const SolverState* st_base) {
    const SolverState* st = st_base + ctx.sys();
    TF_DONE_RETURN(partD, 1);
    const TileSpan sp = ctx_span(ctx, A.n);
    Gather1<T> g{p2};
    DotEpi<T> epi{r2, v2, {T(0)}};
    for (int t = sp.first; t < sp.end; t += sp.step) {
        int r0, r1;
        ctx_tile_rows(ctx, A.n, t, r0, r1);
        op_tile<1, L>(A, r0, r1, g, epi);
    }
    site_write<T, 1>(epi.acc, partD);
}

// alfa = rho/(r2,v2) (:220); q1 = u1 - alfa v2 (:226-228);
// p1 = (u1+q1)/D (:238-241), or w = u1+q1 into p1 for the DIC apply.
template <typename T, bool JAC>
__global__ __launch_bounds__(kBlock) void k_cgs_q(BatchCtx ctx, int k_arg, int n, const T* __restrict__ u1,
                                                  const T* __restrict__ v2, T* __restrict__ q1,
                                                  const T* __restrict__ D, T* __restrict__ p1,
                                                  const T* __restrict__ partD, const SolverState* st_base) {
    const SolverState* st = st_base + ctx.sys();
    if (st->done) return;
    const int k = ctx.k(k_arg);
    T rv[1];
    ctx_total<T, 1>(ctx, partD, rv);
    const T alfa = (T)st->rho[k & 1] / rv[0];
    int rb, re;
    ctx_rows(ctx, n, rb, re);
    for (int i = rb + threadIdx.x; i < re; i += kBlock) {
        const T ui = u1[i];
        const T qi = ui - alfa * v2[i];
        q1[i] = qi;
        const T w = ui + qi;
        p1[i] = JAC ? jac(w, D[i]) : w;
    }
}

// x += alfa p1 (:286-288); q2 = A p1 (:294-300); r1 -= alfa q2 (:315-317),
// with ||r1||^2 and the next rho = (r1, r2) (:118-122) -> site B.
template <typename T>
struct CgsREpi {
    T* __restrict__ x;
    T* __restrict__ r1;
    const T* __restrict__ r2;
    const T* __restrict__ p1;
    T alfa;
    T acc[2];
    __device__ void operator()(int row, T (&y)[1]) {
        x[row] = x[row] + alfa * p1[row];
        const T ri = r1[row] - alfa * y[0];
        r1[row] = ri;
        acc[0] += ri * ri;
        acc[1] += ri * r2[row];
    }
};

template <typename Op, int L, typename T = typename OpValue<Op>::type>
__global__ __launch_bounds__(kBlock) void k_cgs_spmv_r(BatchCtx ctx, int k_arg, Op A, const T* __restrict__ p1,
                                                       T* __restrict__ x, T* __restrict__ r1,
                                                       const T* __restrict__ r2, const T* __restrict__ partD,
                                                       T* __restrict__ partB, const SolverState* st_base) {
    const SolverState* st = st_base + ctx.sys();
    TF_DONE_RETURN(partB, 2);
    const int k = ctx.k(k_arg);
    T rv[1];
    ctx_total<T, 1>(ctx, partD, rv);
    const T alfa = (T)st->rho[k & 1] / rv[0];
    const TileSpan sp = ctx_span(ctx, A.n);
    Gather1<T> g{p1};
    CgsREpi<T> epi{x, r1, r2, p1, alfa, {T(0), T(0)}};
    for (int t = sp.first; t < sp.end; t += sp.step) {
        int r0, r1_;
        ctx_tile_rows(ctx, A.n, t, r0, r1_);
        op_tile<1, L>(A, r0, r1_, g, epi);
    }
    site_write<T, 2>(epi.acc, partB);
}

// ===========================================================================
// Stand-alone SpMV (q = A p) and residual (r -= A x), for the prec1_/resid_
// entry points, tests and the bandwidth benchmark.
// ===========================================================================
template <typename T>
struct StoreEpi {
    T* __restrict__ y;
    __device__ void operator()(int row, T (&v)[1]) { y[row] = v[0]; }
};

template <typename Op, int L, typename T = typename OpValue<Op>::type>
__global__ __launch_bounds__(kBlock) void k_spmv(BatchCtx ctx, Op A, const T* __restrict__ x, T* __restrict__ y) {
    const TileSpan sp = ctx_span(ctx, A.n);
    Gather1<T> g{x};
    StoreEpi<T> epi{y};
    for (int t = sp.first; t < sp.end; t += sp.step) {
        if constexpr (std::is_same_v<Op, DiffSell32>)
            sell_tile<1, true>(A, t * kTileRows, min((t + 1) * kTileRows, A.n), g, epi);
        else
            op_tile<1, L>(A, t * kTileRows, min((t + 1) * kTileRows, A.n), g, epi);
    }
}

template <typename T, int L>
__global__ __launch_bounds__(kBlock) void k_resid(BatchCtx ctx, Csr<T> A, const T* __restrict__ x,
                                                  T* __restrict__ r) {
    const TileSpan sp = ctx_span(ctx, A.n);
    Gather1<T> g{x};
    InitVec<T> init{r};
    StoreEpi<T> epi{r};
    for (int t = sp.first; t < sp.end; t += sp.step)
        resid_tile<L>(A, t * kTileRows, min((t + 1) * kTileRows, A.n), g, init, epi);
}

// Device iteration base of replayed graphs: set, and advanced by a chunk
// (the last node of each captured chunk).  Kernels, not memsets, so that
// they can sit inside a graph.
__global__ void k_set_int(int* p, int v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}
__global__ void k_set_double(double* p, double v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}
__global__ void k_advance(int* kbase, int c) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *kbase += c;
}

template <typename TO, typename TI>
__global__ void k_convert(int64_t count, const TI* __restrict__ in, TO* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = (TO)in[i];
}

// ---------------------------------------------------------------------------
// host-side launchers (explicitly instantiated below)
// ---------------------------------------------------------------------------
// ===========================================================================
// Persistent batched CG (Jacobi or no preconditioner): one launch runs
// `chunk` iterations of every system of a batch.  The three kernels of an
// iteration (k_cg_p, k_cg_spmv, k_cg_update) become three phases of one
// block, and the kernel boundaries between them become barriers over the
// blocks of the block's own system, so each system advances at its own
// pace.  Every phase does exactly what its kernel does, on the same tiles
// and the same reduction sites, so the results are the per-kernel form's
// bit for bit.
//
// Barrier: every thread releases its stores at agent scope (they reach the
// L2 and the L2 is written back, so blocks on other XCDs see them); thread
// 0 adds one arrival to the system's counter and waits until all of the
// system's blocks have arrived; every thread then acquires at agent scope.
// Co-residency of the grid is what makes the wait end: the host sizes the
// grid to the blocks the device can hold at once (occupancy query).  The
// wait is bounded: every 64 polls it checks the fault word and the wall
// clock, and past wait_ticks (other work held the CUs) it raises the fault
// word and the block stops; every other waiting block sees the fault and
// stops too, so every wavefront ends, and the host reruns the solve in the
// per-kernel form.
// ===========================================================================
__device__ __forceinline__ bool system_barrier(int* ctr, int target, int* fault, uint64_t t0, uint64_t wait_ticks) {
    __shared__ int s_ok;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        unsigned polls = 0;
        while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if ((++polls & 63u) == 0) {
                const bool late = (uint64_t)wall_clock64() - t0 > wait_ticks;
                if (late || __hip_atomic_load(fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    __hip_atomic_store(fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
            }
        }
        s_ok = ok;
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return s_ok != 0;
}

template <typename Op, int L, typename T = typename OpValue<Op>::type>
__global__ __launch_bounds__(kBlock) void k_batch_cg_chunk(BatchCtx ctx, Op A, int n, int chunk, T* __restrict__ x,
                                                           T* __restrict__ r, T* __restrict__ p, T* __restrict__ q,
                                                           const T* __restrict__ D, T* __restrict__ partB,
                                                           T* __restrict__ partD, SolverState* st_base,
                                                           HostMirror* hm_base, int* __restrict__ bar,
                                                           int* __restrict__ fault, uint64_t wait_ticks) {
    TF_STATE;
    const int sys = ctx.sys();
    const int nblk = ctx.sys_blk0[sys + 1] - ctx.sys_blk0[sys];
    // A system that has stopped returns at once.  Its partial sums are not
    // rewritten: nothing reads a finished system's sites again (the
    // rebalancing skips it), and a block that stops inside a phase must
    // not write a site that the other blocks of its system may still be
    // reading in that phase.
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = st->done;
    __syncthreads();
    if (s_done) return;
    int rb, re;
    ctx_rows(ctx, n, rb, re);
    const TileSpan sp = ctx_span(ctx, A.n);
    const uint64_t t0 = (uint64_t)wall_clock64();
    int arrived = 0;
    for (int j = 1; j <= chunk; ++j) {
        const int k = ctx.k(j);
        // k_cg_p: stopping test (site B), rho, p = z + beta p  (CG.f90:172-187)
        const double rho_prev = st->rho[(k - 1) & 1];
        T tot[2];
        if (check_stop<T>(ctx, k, partB, st, hm, tot)) return;   // every block of the system takes this exit
        const T rho = tot[1];
        if (ctx.leader()) st->rho[k & 1] = (double)rho;
        const bool first = k == 1;
        const T beta = first ? T(0) : rho / (T)rho_prev;
        for (int i = rb + (int)threadIdx.x; i < re; i += kBlock) {
            const T zi = jac(r[i], D[i]);
            p[i] = first ? zi : zi + beta * p[i];
        }
        if (!system_barrier(bar + sys, ++arrived * nblk, fault, t0, wait_ticks)) return;
        // k_cg_spmv: q = A p, (p, q) -> site D  (CG.f90:193-218)
        {
            Gather1<T> g{p};
            DotEpi<T> epi{p, q, {T(0)}};
            for (int t = sp.first; t < sp.end; t += sp.step) {
                int r0, r1;
                ctx_tile_rows(ctx, A.n, t, r0, r1);
                op_tile<1, L>(A, r0, r1, g, epi);
            }
            site_write<T, 1>(epi.acc, partD);
        }
        if (!system_barrier(bar + sys, ++arrived * nblk, fault, t0, wait_ticks)) return;
        // k_cg_update: alfa, x += alfa p, r -= alfa q, ||r||^2 and the next
        // rho -> site B  (CG.f90:219-253)
        {
            const double rho_k = st->rho[k & 1];
            T pq[1];
            ctx_total<T, 1>(ctx, partD, pq);
            const T alfa = (T)rho_k / pq[0];
            T v[2] = {T(0), T(0)};
            for (int i = rb + (int)threadIdx.x; i < re; i += kBlock) {
                x[i] = x[i] + alfa * p[i];
                const T ri = r[i] - alfa * q[i];
                r[i] = ri;
                v[0] += ri * ri;
                v[1] += ri * jac(ri, D[i]);
            }
            site_write<T, 2>(v, partB);
        }
        if (j < chunk && !system_barrier(bar + sys, ++arrived * nblk, fault, t0, wait_ticks)) return;
    }
}

__global__ void k_zero_ints(int n, int* __restrict__ a) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = 0;
}

#define TF_LAUNCH(k, grid, stream) hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, stream

// Calls f(std::integral_constant<int, L>) for the SpMV form `lanes`.
template <class F>
static void with_lanes(int lanes, F&& f) {
    switch (lanes) {
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        default: f(std::integral_constant<int, 1>{}); break;
    }
}
#define TF_L constexpr int L = decltype(lc_)::value

template <typename T>
void launch_setup(const BatchCtx& c, const Csr<T>& A, int prec, bool write_d, T* D, const T* b, const Sites<T>& S,
                  hipStream_t s) {
    TF_LAUNCH((k_setup<T>), S.grid, s), c, A, prec, write_d, D, b, S.A);
}
template <typename T>
void launch_dic_setup_level(const Csr<T>& A, const int* rows, int count, T* D, hipStream_t s) {
    if (count > 0) TF_LAUNCH((k_dic_setup_level<T>), (count + kBlock - 1) / kBlock, s), A, rows, count, D);
}
template <typename T>
void launch_dic_fwd_level(const Csr<T>& A, const int* rows, int count, const T* D, const T* u, T* z,
                          const SolverState* st, hipStream_t s) {
    if (count > 0) TF_LAUNCH((k_dic_fwd_level<T>), (count + kBlock - 1) / kBlock, s), A, rows, count, D, u, z, st);
}
template <typename T>
void launch_dic_bwd_level(const Csr<T>& A, const int* rows, int count, const T* D, T* z,
                          const SolverState* st, hipStream_t s) {
    if (count > 0) TF_LAUNCH((k_dic_bwd_level<T>), (count + kBlock - 1) / kBlock, s), A, rows, count, D, z, st);
}
template <typename T>
void launch_plan_in(int n, const int* order, const T* u, const T* D, T* up, T* dp, const SolverState* st,
                    hipStream_t s) {
    if (n > 0) TF_LAUNCH((k_plan_in<T>), (n + kBlock - 1) / kBlock, s), n, order, u, D, up, dp, st);
}
template <typename T>
void launch_flow_reset(int n, T* zf, hipStream_t s) {
    if (n > 0) TF_LAUNCH((k_flow_reset<T>), (n + kBlock - 1) / kBlock, s), n, zf);
}
template <typename T>
void launch_flow_fwd(int n, const int* order, const T* u, const T* D, const int* lptr, const int* lcol,
                     const T* lval, T* dp, T* zf, T* zb, int* fault, uint64_t wait_ticks, const SolverState* st,
                     hipStream_t s) {
    if (n > 0)
        TF_LAUNCH((k_flow_fwd<T>), flow_grid<T>(n), s), n, order, u, D, lptr, lcol, lval, dp, zf, zb, fault,
            wait_ticks, st);
}
template <typename T>
void launch_flow_bwd(int n, const int* order, T* z, const int* uptr, const int* ucol, const T* uval, const T* dp,
                     T* zf, T* zb, int* fault, uint64_t wait_ticks, const SolverState* st, hipStream_t s) {
    if (n > 0)
        TF_LAUNCH((k_flow_bwd<T>), flow_grid<T>(n), s), n, order, z, uptr, ucol, uval, dp, zf, zb, fault,
            wait_ticks, st);
}
template <typename T>
void launch_tile_fwd(const TileSweep& ts, const T* u, const T* D, const int* lcol, const T* lval, T* dp, T* zf, T* zb,
                     const SolverState* st, hipStream_t s) {
    if (ts.waves <= 0) return;
    if (ts.slots == 3)
        hipLaunchKernelGGL((k_tile_fwd<T, 3>), dim3(ts.waves), dim3(kTileWave), 0, s, ts.n, ts.wstart, ts.wtiles,
                           ts.tph, ts.order, u, D, lcol, lval, dp, zf, zb, ts.fault, ts.wait_ticks, st);
    else
        hipLaunchKernelGGL((k_tile_fwd<T, 4>), dim3(ts.waves), dim3(kTileWave), 0, s, ts.n, ts.wstart, ts.wtiles,
                           ts.tph, ts.order, u, D, lcol, lval, dp, zf, zb, ts.fault, ts.wait_ticks, st);
}
template <typename T>
void launch_tile_bwd(const TileSweep& ts, T* z, const int* ucol, const T* uval, const T* dp, T* zf, T* zb,
                     const SolverState* st, hipStream_t s) {
    if (ts.waves <= 0) return;
    if (ts.slots == 3)
        hipLaunchKernelGGL((k_tile_bwd<T, 3>), dim3(ts.waves), dim3(kTileWave), 0, s, ts.n, ts.wstart, ts.wtiles,
                           ts.tph, ts.order, z, ucol, uval, dp, zf, zb, ts.fault, ts.wait_ticks, st);
    else
        hipLaunchKernelGGL((k_tile_bwd<T, 4>), dim3(ts.waves), dim3(kTileWave), 0, s, ts.n, ts.wstart, ts.wtiles,
                           ts.tph, ts.order, z, ucol, uval, dp, zf, zb, ts.fault, ts.wait_ticks, st);
}
void launch_tile_pivots(const TileSweep& ts, const double* dg, const int* lcol, const double* lval, double* D,
                        double* zs, hipStream_t s) {
    if (ts.waves <= 0) return;
    if (ts.slots == 3)
        hipLaunchKernelGGL((k_tile_pivots<3>), dim3(ts.waves), dim3(kTileWave), 0, s, ts.n, ts.wstart, ts.wtiles,
                           ts.tph, ts.order, dg, lcol, lval, D, zs, ts.fault, ts.wait_ticks);
    else
        hipLaunchKernelGGL((k_tile_pivots<4>), dim3(ts.waves), dim3(kTileWave), 0, s, ts.n, ts.wstart, ts.wtiles,
                           ts.tph, ts.order, dg, lcol, lval, D, zs, ts.fault, ts.wait_ticks);
}
template <typename T>
void launch_plan_out(int n, const int* order, const T* zp, T* z, const SolverState* st, hipStream_t s) {
    if (n > 0) TF_LAUNCH((k_plan_out<T>), (n + kBlock - 1) / kBlock, s), n, order, zp, z, st);
}
template <typename T>
void launch_plan_fwd(int a, int b, const int* lptr, const int* lcol, const T* lval, const T* dp, const T* up, T* zp,
                     const SolverState* st, hipStream_t s) {
    if (b > a) TF_LAUNCH((k_plan_fwd<T>), (b - a + kBlock - 1) / kBlock, s), a, b, lptr, lcol, lval, dp, up, zp, st);
}
template <typename T>
void launch_plan_bwd(const int* pos, int count, const int* uptr, const int* ucol, const T* uval, const T* dp, T* zp,
                     const SolverState* st, hipStream_t s) {
    if (count > 0)
        TF_LAUNCH((k_plan_bwd<T>), (count + kBlock - 1) / kBlock, s), pos, count, uptr, ucol, uval, dp, zp, st);
}
void launch_entry_gather(size_t count, const int* map, const double* src, double* dst, hipStream_t s) {
    if (count > 0)
        hipLaunchKernelGGL(k_entry_gather, dim3((unsigned)((count + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                           (int)count, map, src, dst);
}

template <typename T, bool Z, int L>
static void resid_init_mode(int mode, const BatchCtx& c, const Csr<T>& A, const T* x, T* r, T* r2, const T* D,
                            const Sites<T>& S, SolverState* st, HostMirror* hm, hipStream_t s) {
    switch (mode) {
        case RHO_JAC: TF_LAUNCH((k_resid_init<T, RHO_JAC, Z, L>), S.grid, s), c, A, x, r, r2, D, S.A, S.B, st, hm); break;
        case RHO_BICG: TF_LAUNCH((k_resid_init<T, RHO_BICG, Z, L>), S.grid, s), c, A, x, r, r2, D, S.A, S.B, st, hm); break;
        case RHO_RR: TF_LAUNCH((k_resid_init<T, RHO_RR, Z, L>), S.grid, s), c, A, x, r, r2, D, S.A, S.B, st, hm); break;
        default: TF_LAUNCH((k_resid_init<T, RHO_NONE, Z, L>), S.grid, s), c, A, x, r, r2, D, S.A, S.B, st, hm); break;
    }
}
template <typename T>
void launch_resid_init(int mode, bool zero_x, const BatchCtx& c, const Csr<T>& A, const T* x, T* r, T* r2,
                       const T* D, const Sites<T>& S, SolverState* st, HostMirror* hm, hipStream_t s) {
    if (zero_x) {  // no matrix pass: one form suffices
        resid_init_mode<T, true, 1>(mode, c, A, x, r, r2, D, S, st, hm, s);
        return;
    }
    with_lanes(c.lanes, [&](auto lc_) {
        TF_L;
        resid_init_mode<T, false, L>(mode, c, A, x, r, r2, D, S, st, hm, s);
    });
}
template <typename T>
void launch_check(const BatchCtx& c, int k, const Sites<T>& S, SolverState* st, HostMirror* hm, hipStream_t s) {
    TF_LAUNCH((k_check<T>), S.grid, s), c, k, S.B, st, hm);
}
template <typename T>
void launch_rho(const BatchCtx& c, int n, const T* a, const T* b, const Sites<T>& S, const SolverState* st,
                hipStream_t s) {
    TF_LAUNCH((k_rho<T>), S.grid, s), c, n, a, b, S.C, st);
}
template <typename T>
void launch_cg_p(const BatchCtx& c, bool jac, int k, int n, const T* r, const T* D, const T* z, T* p,
                 const Sites<T>& S, SolverState* st, HostMirror* hm, hipStream_t s) {
    if (jac) TF_LAUNCH((k_cg_p<T, true, false>), S.grid, s), c, k, n, r, D, z, p, (T*)nullptr, S.B, S.C, st, hm);
    else TF_LAUNCH((k_cg_p<T, false, false>), S.grid, s), c, k, n, r, D, z, p, (T*)nullptr, S.B, S.C, st, hm);
}
template <typename T>
void launch_cg_p_carry(const BatchCtx& c, int k, int n, const T* r, const T* D, T* p, T* x, const Sites<T>& S,
                       SolverState* st, HostMirror* hm, hipStream_t s) {
    if (c.blk_sys) throw std::logic_error("the carried x update is for one system");
    TF_LAUNCH((k_cg_p<T, true, true>), S.grid, s), c, k, n, r, D, (const T*)nullptr, p, x, S.B, S.C, st, hm);
}
template <typename Op, typename T>
void launch_cg_spmv(const BatchCtx& c, const Op& A, const T* p, T* q, const Sites<T>& S, const SolverState* st,
                    hipStream_t s) {
    with_lanes(c.lanes, [&](auto lc_) {
        TF_L;
        TF_LAUNCH((k_cg_spmv<Op, L>), S.grid, s), c, A, p, q, S.D, st);
    });
}
template <typename T>
void launch_cg_update(const BatchCtx& c, bool jac, int k, int n, T* x, T* r, const T* p, const T* q, const T* D,
                      const Sites<T>& S, const SolverState* st, hipStream_t s) {
    if (jac) TF_LAUNCH((k_cg_update<T, true, false>), S.grid, s), c, k, n, x, r, p, q, D, S.D, S.B, (SolverState*)st);
    else TF_LAUNCH((k_cg_update<T, false, false>), S.grid, s), c, k, n, x, r, p, q, D, S.D, S.B, (SolverState*)st);
}
template <typename T>
void launch_cg_update_carry(const BatchCtx& c, int k, int n, T* r, const T* q, const T* D, const Sites<T>& S,
                            SolverState* st, hipStream_t s) {
    if (c.blk_sys) throw std::logic_error("the carried x update is for one system");
    TF_LAUNCH((k_cg_update<T, true, true>), S.grid, s), c, k, n, (T*)nullptr, r, (const T*)nullptr, q, D, S.D,
        S.B, st);
}
template <typename T>
void launch_cg_xfinish(int n, T* x, const T* p, const SolverState* st, hipStream_t s) {
    TF_LAUNCH((k_cg_xfinish<T>), (n + kBlock - 1) / kBlock, s), n, x, p, st);
}
template <typename T>
void launch_bicg_p(const BatchCtx& c, bool jac, int k, int n, const T* r1, const T* r2, const T* D, const T* z1,
                   const T* z2, T* p1, T* p2, const Sites<T>& S, SolverState* st, HostMirror* hm, hipStream_t s) {
    if (jac) TF_LAUNCH((k_bicg_p<T, true>), S.grid, s), c, k, n, r1, r2, D, z1, z2, p1, p2, S.B, S.C, st, hm);
    else TF_LAUNCH((k_bicg_p<T, false>), S.grid, s), c, k, n, r1, r2, D, z1, z2, p1, p2, S.B, S.C, st, hm);
}
template <typename Op, typename T>
void launch_bicg_spmv(const BatchCtx& c, const Op& A, const T* p1, const T* p2, T* q1, T* q2, const Sites<T>& S,
                      const SolverState* st, hipStream_t s) {
    with_lanes(c.lanes, [&](auto lc_) {
        TF_L;
        TF_LAUNCH((k_bicg_spmv<Op, L>), S.grid, s), c, A, p1, p2, q1, q2, S.D, st);
    });
}
template <typename T>
void launch_bicg_update(const BatchCtx& c, bool jac, int k, int n, T* x, T* r1, T* r2, const T* p1, const T* q1,
                        const T* q2, const T* D, const Sites<T>& S, const SolverState* st, hipStream_t s) {
    if (jac) TF_LAUNCH((k_bicg_update<T, true>), S.grid, s), c, k, n, x, r1, r2, p1, q1, q2, D, S.D, S.B, st);
    else TF_LAUNCH((k_bicg_update<T, false>), S.grid, s), c, k, n, x, r1, r2, p1, q1, q2, D, S.D, S.B, st);
}
template <typename T>
void launch_cgs_u(const BatchCtx& c, bool jac, int k, int n, const T* r1, const T* q1, T* u1, T* u2, const T* D,
                  T* p2, const Sites<T>& S, SolverState* st, HostMirror* hm, hipStream_t s) {
    if (jac) TF_LAUNCH((k_cgs_u<T, true>), S.grid, s), c, k, n, r1, q1, u1, u2, D, p2, S.B, st, hm);
    else TF_LAUNCH((k_cgs_u<T, false>), S.grid, s), c, k, n, r1, q1, u1, u2, D, p2, S.B, st, hm);
}
template <typename Op, typename T>
void launch_cgs_spmv_v(const BatchCtx& c, const Op& A, const T* p2, T* v2, const T* r2, const Sites<T>& S,
                       const SolverState* st, hipStream_t s) {
    with_lanes(c.lanes, [&](auto lc_) {
        TF_L;
        TF_LAUNCH((k_cgs_spmv_v<Op, L>), S.grid, s), c, A, p2, v2, r2, S.D, st);
    });
}
template <typename T>
void launch_cgs_q(const BatchCtx& c, bool jac, int k, int n, const T* u1, const T* v2, T* q1, const T* D, T* p1,
                  const Sites<T>& S, const SolverState* st, hipStream_t s) {
    if (jac) TF_LAUNCH((k_cgs_q<T, true>), S.grid, s), c, k, n, u1, v2, q1, D, p1, S.D, st);
    else TF_LAUNCH((k_cgs_q<T, false>), S.grid, s), c, k, n, u1, v2, q1, D, p1, S.D, st);
}
template <typename Op, typename T>
void launch_cgs_spmv_r(const BatchCtx& c, int k, const Op& A, const T* p1, T* x, T* r1, const T* r2,
                       const Sites<T>& S, const SolverState* st, hipStream_t s) {
    with_lanes(c.lanes, [&](auto lc_) {
        TF_L;
        TF_LAUNCH((k_cgs_spmv_r<Op, L>), S.grid, s), c, k, A, p1, x, r1, r2, S.D, S.B, st);
    });
}
template <typename Op, typename T>
void launch_spmv(const Op& A, const T* x, T* y, int grid, hipStream_t s, const BatchCtx& c) {
    if (c.blk_sys) throw std::invalid_argument("stand-alone SpMV takes a one-system context");
    with_lanes(c.lanes, [&](auto lc_) {
        TF_L;
        TF_LAUNCH((k_spmv<Op, L>), grid, s), c, A, x, y);
    });
}
template <typename T>
void launch_resid(const Csr<T>& A, const T* x, T* r, int grid, hipStream_t s, const BatchCtx& c) {
    if (c.blk_sys) throw std::invalid_argument("stand-alone residual takes a one-system context");
    with_lanes(c.lanes, [&](auto lc_) {
        TF_L;
        TF_LAUNCH((k_resid<T, L>), grid, s), c, A, x, r);
    });
}
template <typename Op, typename T>
void launch_batch_cg_chunk(const BatchCtx& c, const Op& A, int n, int chunk, T* x, T* r, T* p, T* q, const T* D,
                           const Sites<T>& S, SolverState* st, HostMirror* hm, int* bar, int nsys, int* fault,
                           uint64_t wait_ticks, hipStream_t s) {
    if (!c.blk_sys) throw std::invalid_argument("the persistent chunk runs under a batch context");
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(kBlock), 0, s, nsys, bar);
    with_lanes(c.lanes, [&](auto lc_) {
        TF_L;
        TF_LAUNCH((k_batch_cg_chunk<Op, L>), S.grid, s), c, A, n, chunk, x, r, p, q, D, S.B, S.D, st, hm, bar, fault,
            wait_ticks);
    });
}

template <typename Op>
int batch_cg_chunk_resident(int lanes) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("hipGetDevice failed");
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        throw std::runtime_error("hipDeviceGetAttribute failed");
    with_lanes(lanes, [&](auto lc_) {
        TF_L;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_batch_cg_chunk<Op, L>, kBlock, 0) !=
            hipSuccess)
            throw std::runtime_error("occupancy query of the persistent batch kernel failed");
    });
    return per_cu * cus;
}

void launch_set_int(int* p, int v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(kWave), 0, s, p, v);
}
void launch_set_double(double* p, double v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_double, dim3(1), dim3(kWave), 0, s, p, v);
}
void launch_advance(int* kbase, int c, hipStream_t s) {
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(kWave), 0, s, kbase, c);
}

void launch_f64_to_f32(int64_t count, const double* in, float* out, hipStream_t s) {
    if (count > 0)
        hipLaunchKernelGGL((k_convert<float, double>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s,
                           count, in, out);
}

#define TF_INSTANTIATE_PLAN(T)                                                                                   \
    template void launch_plan_in<T>(int, const int*, const T*, const T*, T*, T*, const SolverState*, hipStream_t); \
    template void launch_plan_out<T>(int, const int*, const T*, T*, const SolverState*, hipStream_t);            \
    template void launch_plan_fwd<T>(int, int, const int*, const int*, const T*, const T*, const T*, T*,         \
                                     const SolverState*, hipStream_t);                                          \
    template void launch_plan_bwd<T>(const int*, int, const int*, const int*, const T*, const T*, T*,           \
                                     const SolverState*, hipStream_t);                                          \
    template void launch_flow_reset<T>(int, T*, hipStream_t);                                                   \
    template void launch_flow_fwd<T>(int, const int*, const T*, const T*, const int*, const int*, const T*, T*,  \
                                     T*, T*, int*, uint64_t, const SolverState*, hipStream_t);                  \
    template void launch_flow_bwd<T>(int, const int*, T*, const int*, const int*, const T*, const T*, T*, T*,    \
                                     int*, uint64_t, const SolverState*, hipStream_t);                         \
    template void launch_tile_fwd<T>(const TileSweep&, const T*, const T*, const int*, const T*, T*, T*, T*,     \
                                     const SolverState*, hipStream_t);                                         \
    template void launch_tile_bwd<T>(const TileSweep&, T*, const int*, const T*, const T*, T*, T*,               \
                                     const SolverState*, hipStream_t);
TF_INSTANTIATE_PLAN(double)
TF_INSTANTIATE_PLAN(float)

#define TF_INSTANTIATE(T)                                                                                        \
    template void launch_setup<T>(const BatchCtx&, const Csr<T>&, int, bool, T*, const T*, const Sites<T>&,     \
                                  hipStream_t);                                                                  \
    template void launch_dic_setup_level<T>(const Csr<T>&, const int*, int, T*, hipStream_t);                   \
    template void launch_dic_fwd_level<T>(const Csr<T>&, const int*, int, const T*, const T*, T*,               \
                                          const SolverState*, hipStream_t);                                      \
    template void launch_dic_bwd_level<T>(const Csr<T>&, const int*, int, const T*, T*, const SolverState*,      \
                                          hipStream_t);                                                          \
    template void launch_resid_init<T>(int, bool, const BatchCtx&, const Csr<T>&, const T*, T*, T*, const T*,   \
                                       const Sites<T>&, SolverState*, HostMirror*, hipStream_t);                 \
    template void launch_check<T>(const BatchCtx&, int, const Sites<T>&, SolverState*, HostMirror*,             \
                                  hipStream_t);                                                                  \
    template void launch_rho<T>(const BatchCtx&, int, const T*, const T*, const Sites<T>&, const SolverState*,  \
                                hipStream_t);                                                                    \
    template void launch_cg_p<T>(const BatchCtx&, bool, int, int, const T*, const T*, const T*, T*,             \
                                 const Sites<T>&, SolverState*, HostMirror*, hipStream_t);                       \
    template void launch_cg_update<T>(const BatchCtx&, bool, int, int, T*, T*, const T*, const T*, const T*,    \
                                      const Sites<T>&, const SolverState*, hipStream_t);                         \
    template void launch_cg_p_carry<T>(const BatchCtx&, int, int, const T*, const T*, T*, T*, const Sites<T>&,  \
                                       SolverState*, HostMirror*, hipStream_t);                                  \
    template void launch_cg_update_carry<T>(const BatchCtx&, int, int, T*, const T*, const T*, const Sites<T>&, \
                                            SolverState*, hipStream_t);                                          \
    template void launch_cg_xfinish<T>(int, T*, const T*, const SolverState*, hipStream_t);                      \
    template void launch_bicg_p<T>(const BatchCtx&, bool, int, int, const T*, const T*, const T*, const T*,     \
                                   const T*, T*, T*, const Sites<T>&, SolverState*, HostMirror*, hipStream_t);   \
    template void launch_bicg_update<T>(const BatchCtx&, bool, int, int, T*, T*, T*, const T*, const T*,        \
                                        const T*, const T*, const Sites<T>&, const SolverState*, hipStream_t);   \
    template void launch_cgs_u<T>(const BatchCtx&, bool, int, int, const T*, const T*, T*, T*, const T*, T*,    \
                                  const Sites<T>&, SolverState*, HostMirror*, hipStream_t);                      \
    template void launch_cgs_q<T>(const BatchCtx&, bool, int, int, const T*, const T*, T*, const T*, T*,        \
                                  const Sites<T>&, const SolverState*, hipStream_t);                             \
    template void launch_resid<T>(const Csr<T>&, const T*, T*, int, hipStream_t, const BatchCtx&);

#define TF_INSTANTIATE_OP(Op, T)                                                                                 \
    template void launch_cg_spmv<Op, T>(const BatchCtx&, const Op&, const T*, T*, const Sites<T>&,              \
                                        const SolverState*, hipStream_t);                                        \
    template void launch_bicg_spmv<Op, T>(const BatchCtx&, const Op&, const T*, const T*, T*, T*,               \
                                          const Sites<T>&, const SolverState*, hipStream_t);                     \
    template void launch_cgs_spmv_v<Op, T>(const BatchCtx&, const Op&, const T*, T*, const T*, const Sites<T>&, \
                                           const SolverState*, hipStream_t);                                     \
    template void launch_cgs_spmv_r<Op, T>(const BatchCtx&, int, const Op&, const T*, T*, T*, const T*,         \
                                           const Sites<T>&, const SolverState*, hipStream_t);                    \
    template void launch_spmv<Op, T>(const Op&, const T*, T*, int, hipStream_t, const BatchCtx&);          \
    template void launch_batch_cg_chunk<Op, T>(const BatchCtx&, const Op&, int, int, T*, T*, T*, T*, const T*,  \
                                               const Sites<T>&, SolverState*, HostMirror*, int*, int, int*,     \
                                               uint64_t, hipStream_t);                                          \
    template int batch_cg_chunk_resident<Op>(int);

TF_INSTANTIATE(double)
TF_INSTANTIATE(float)
TF_INSTANTIATE_OP(Csr<double>, double)
TF_INSTANTIATE_OP(Csr<float>, float)
TF_INSTANTIATE_OP(DiffCsr32, float)
TF_INSTANTIATE_OP(DiffSell32, float)
TF_INSTANTIATE_OP(Sell<double>, double)

}  // namespace tflows
