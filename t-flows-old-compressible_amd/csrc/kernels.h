// kernels.h -- device state and the fused CSR-stream machinery shared by the
// Krylov kernels.  Everything here runs on gfx950 (MI355X).
#pragma once
#include "device_common.h"

namespace tflows {

constexpr double kTiny = 1.0e-64;  // Library/allp_mod.f90:31 (TINY)

// Per-solve scalars in device memory.  Kernels never read a field that a
// block of the same kernel writes: rho is kept in two slots by iteration
// parity, the iteration number is a launch argument, and only block 0
// (the "leader") writes.
struct SolverState {
    double rho[2];                  // rho of iteration k in rho[k & 1]
    double error;                   // last ||r||_2 (FinRes; stale on the early exit)
    double ini_res;                 // IniRes
    double tol;                     // absolute tolerance
    double bnrm2;                   // ||b||_2
    double aux[4];                  // scratch for composite solvers
    int done;                       // 1 once converged, capped or exited early
    int niter_out;                  // the reference's `niter` on exit
    int cap;                        // iteration cap
    int ini_written;                // 1 when IniRes was set
    // device-side iteration bookkeeping of the fused CG kernels: `it` is the
    // number of completed iterations (written by the second kernel of an
    // iteration), `kcur` the one in progress (written by the first); each is
    // read only by the other kernel, so no kernel reads what it writes.
    int it;
    int kcur;
    // mixed precision with residual replacement: `suspend` is set when the
    // FP32 recurrence residual reaches tol (iterations pause until the next
    // FP64 replacement verifies); `replaced_before` = the iteration that
    // follows the last replacement, whose stopping test uses the FP64 norm.
    int suspend;
    int replaced_before;
    double err_rep;                 // FP64 residual norm at the last replacement
    // guards of the FP32 inner solves of the mixed-precision path (0 = off,
    // the reference behaviour): the solve stops, marked `diverged`, once
    // ||r|| exceeds div_limit or is not finite, and stops (stalled) once
    // the smallest ||r|| so far is more than max(stall_min, its iteration)
    // iterations old.  err_min / k_min (the smallest ||r|| and where) are
    // kept by iteration parity (read [(k-1)&1], written [k&1]); the host
    // sets err_min[0] = +inf.
    double div_limit;
    double err_min[2];
    int k_min[2];
    int stall_min;
    int diverged;
    // x += alfa p carried into the next iteration's p kernel (one system,
    // Jacobi CG): the update kernel of iteration `it` leaves its alfa here
    // and does not touch x; the p kernel of iteration it + 1 applies it, and
    // when the solve stops, the finishing pass (or the FP64 merge of the
    // mixed-precision path) does.  `xupd` = the last iteration whose update
    // a merge has applied, so the p kernel that resumes after a residual
    // replacement does not apply it again.
    double alfa;
    int xupd;
};

// Host-visible mirror of the end of a solve, written by the device with
// system-scope stores into coherent pinned memory so the host can pace its
// launches without events.  Only the stopping test that ends a solve writes
// it (done = 1 and the last iteration): a system-scope release makes the
// writing wave write back its XCD's L2 and wait for it, which is not worth
// paying every iteration for a flag that stays 0.
struct HostMirror {
    volatile int done;
    volatile int iter;
};

__device__ __forceinline__ void publish(HostMirror* hm, int done, int iter) {
    if (hm) {
        __hip_atomic_store(&hm->iter, iter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&hm->done, done, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Launch context of the solver kernels.  For one system every member is
// null: block b owns an even share of the 256-row tiles (XCD-aware), the
// leader is block 0, reductions span the whole grid and the iteration index
// is the launch argument.  For a batch of K systems stored as one
// block-diagonal matrix, tiles never straddle two systems and:
//   blk_sys[b]      system of block b
//   blk_tiles[2b], blk_tiles[2b+1]  tiles [t0, t1) of block b
//   tile_r0[t..t+1] rows of tile t (tiles are contiguous in row order)
//   sys_blk0[k..k+1] blocks of system k (its reduction range; the first is
//                    its leader)
// kbase (optional) offsets the launch's iteration index by a device value,
// so one captured graph of a chunk of iterations can be replayed; koff
// (optional, per system) is subtracted from it, so a system that starts
// iterating after others (a later group of a batch) counts from 1.
// `lanes` selects the SpMV form of the launch: 1 = one row per thread with
// the tile's products staged in LDS (64 rows per wavefront), 2/4/8 = that
// many lanes per row streaming col/val directly (32/16/8 rows per
// wavefront).  Every form sums a row in column order, so all give the same
// bits.
struct BatchCtx {
    const int* blk_sys = nullptr;
    const int* blk_tiles = nullptr;
    const int* tile_r0 = nullptr;
    const int* sys_blk0 = nullptr;
    const int* kbase = nullptr;
    const int* koff = nullptr;
    int lanes = 1;
    // tile order of the SpMV kernels of one system: 0 = each block a
    // contiguous range, 1 = strided inside each XCD's chunk
    int strided = 0;

    __device__ __forceinline__ int sys() const { return blk_sys ? blk_sys[blockIdx.x] : 0; }
    __device__ __forceinline__ bool leader() const {
        return threadIdx.x == 0 && (blk_sys ? (int)blockIdx.x == sys_blk0[sys()] : blockIdx.x == 0);
    }
    __device__ __forceinline__ int k(int k_arg) const {
        const int k = kbase ? *kbase + k_arg : k_arg;
        return koff ? k - koff[sys()] : k;
    }
};

// Matrix view with 0-based indices.
template <typename T>
struct Csr {
    int n;
    const int* __restrict__ rowptr;  // n+1
    const int* __restrict__ col;     // nnz
    const int* __restrict__ dia;     // n   (position of the diagonal)
    const T* __restrict__ val;       // nnz
};

// FP32 operator in difference ("flux") form, used by the inner solves of
// the mixed-precision path:
//   (A p)_i = sum_{j != i} a_ij (p_j - p_i) + rs_i p_i,   rs_i = sum_j a_ij
// with the row sums rs_i formed in FP64 and rounded once.  It is the same
// operator as A, but its FP32 evaluation keeps relative accuracy on smooth
// vectors: rows of a pressure-correction matrix sum to zero, and the
// plain form a_ii p_i + sum a_ij p_j cancels terms of size |a_ii p_i| that
// FP32 cannot resolve against the tiny eigenvalues of the near-null mode.
struct DiffCsr32 {
    int n;
    const int* __restrict__ rowptr;  // n+1, off-diagonal entries only
    const int* __restrict__ col;     // nnz - n
    const float* __restrict__ val;   // off-diagonal a_ij
    const float* __restrict__ rs;    // row sums (from FP64)
};

// The same FP32 difference-form operator in sliced ELL layout: rows in
// slices of 64 (one wavefront), the off-diagonal entries of a slice stored
// slot-major -- slot j of row 64 s + l at off[s] + 64 j + l -- and padded
// to the longest row of the slice with (column = own row, a = 0), which
// adds exactly 0.  One thread per row: every column and value load of a
// wavefront is one contiguous segment, with no LDS staging.
struct DiffSell32 {
    int n;
    const int* __restrict__ off;     // ceil(n/64)+1 slice offsets
    const int* __restrict__ col;
    const float* __restrict__ val;   // off-diagonal a_ij, 0 in padding
    const float* __restrict__ rs;    // row sums (from FP64)
    // when not null: col - row of every slot as 16 bits, read instead of
    // col (made where every column lies within 32767 rows of its row)
    const short* __restrict__ dc = nullptr;
    // when not null: shared slot offsets (see Sell), read instead of col
    const int* __restrict__ soff = nullptr;
    const unsigned char* __restrict__ mask = nullptr;
    // when not null: the symmetric shared-offset form (see SymSlots)
    const int* __restrict__ sidx = nullptr;
    const int2* __restrict__ lbase = nullptr;
    // 1: the shared-offset loop runs over all kMaxSharedSlots slots without
    // branches (every slot's loads in flight at once; a slot past the
    // slice's width reads slot 0 and, like an absent one, adds a (p_i - p_i)
    // with a = 0), in the kernels compiled with it (sell_tile's WIDE); 0: two
    // slots at a time.  The same sums either way.
    int unroll8 = 1;
};
constexpr int kSlice = 64;   // kSliceRows of slot_layout.h
// most slots of a slice in the shared-offset form (bits of the row mask)
constexpr int kMaxSharedSlots = 8;

// A matrix in sliced ELL layout: rows in slices of 64 (one wavefront), the
// entries of a slice stored slot-major -- slot j of row 64 s + l at
// off[s] + 64 j + l -- each row's entries in their CSR order (diagonal
// included where it stands), padded to the slice's longest row with column
// -1, which is skipped.  One thread per row sums its products in that
// order, as the reference loop does, so the result is the CSR stream
// form's bit for bit; the loads of a wavefront are contiguous segments.
template <typename T>
struct Sell {
    int n;
    const int* __restrict__ off;     // ceil(n/64)+1 slice offsets
    const int* __restrict__ col;     // -1: padding
    const T* __restrict__ val;
    // when not null: col - row of every slot as 16 bits (kPad16: padding),
    // read instead of col (made where every column lies within 32767 rows)
    const short* __restrict__ dc = nullptr;
    // when not null: shared slot offsets.  Every row of slice s uses the
    // slice's ascending list of column offsets, slot j holding offset
    // soff[off[s] / 64 + j] (column = row + offset) for every row that has
    // it, which bit j of mask[row] marks; no column index is stored per
    // entry.  Made where each row's columns ascend without repeats, so the
    // slots of a row run in its CSR order.
    const int* __restrict__ soff = nullptr;
    const unsigned char* __restrict__ mask = nullptr;
    // When not null: the symmetric form of the shared offsets, for a matrix
    // whose values equal their mirrors (a_ij = a_ji bit for bit).  Only the
    // slots of non-negative offsets are stored (slice s: off[s] + 64 (j -
    // nlow) + l, nlow = its number of negative offsets); the offset list of
    // slice s starts at soff[sidx[s]] and has sidx[s + 1] - sidx[s] slots.
    // Lower slot j of slice s (offset d < 0) reads a_{i,i+d} as the stored
    // a_{i+d,i} of row i' = i + d: at lbase[sidx[s] + j].x + i' when i' lies
    // in slice (64 s + d) >> 6, else at .y + i' (the slice after).  The
    // slots still run in ascending offset order, so every row's sum is the
    // same as in the full form.
    const int* __restrict__ sidx = nullptr;
    const int2* __restrict__ lbase = nullptr;
};
constexpr short kPad16 = -32768;

// Where lower slot j (offset d < 0, entry present) of row `row` in slice sl
// finds its mirror value in a symmetric shared-offset layout.
__device__ __forceinline__ int sym_mirror_pos(const int2* __restrict__ lbase, int sb, int j, int sl, int row, int d) {
    const int2 lb = lbase[sb + j];
    const int ip = row + d;
    const int sa = (sl * 64 + d) >> 6;          // arithmetic shift: floor division by 64
    return ((ip >> 6) == sa ? lb.x : lb.y) + ip;
}

template <typename Op> struct OpValue;
template <typename T> struct OpValue<Csr<T>> { using type = T; };
template <typename T> struct OpValue<Sell<T>> { using type = T; };
template <> struct OpValue<DiffCsr32> { using type = float; };
template <> struct OpValue<DiffSell32> { using type = float; };

// Rows per tile of the stream kernels: one row per thread in the summation
// phase.  Tiles are the unit of work split across blocks.
constexpr int kTileRows = kBlock;
// LDS entries per right-hand side (per chunk).  A 7-point tile needs
// 256*7 = 1792; longer rows are handled by chunking.
constexpr int kLdsCap = 2048;

// ---------------------------------------------------------------------------
// Stream one tile of rows [r0, r1) of a CSR matrix:
//   phase 1: every thread loads val/col entries of the tile coalesced and
//            stores val[e]*g_v(col[e]) for each right-hand side v into LDS;
//   phase 2: thread t sums the products of local row t sequentially, in
//            column order, starting from init_v(row) (0 for q = A p, or the
//            row's rhs for the residual form, which subtracts).
// The products are rounded before they are summed, and the sum runs in the
// reference's order, so the FP64 result equals the reference loop
// (Process/CG.f90:193-199, Resid.f90:29-34) bit for bit.
// epi(row, y[NRHS]) is called once per row by its owning thread.
// ---------------------------------------------------------------------------
template <typename T, int NRHS, bool SUBTRACT, class Gather, class Init, class Epi>
__device__ __forceinline__ void stream_tile(const Csr<T>& A, int r0, int r1, Gather& g, Init& init, Epi& epi) {
    __shared__ int s_rp[kTileRows + 1];
    __shared__ T s_prod[NRHS * kLdsCap];
    const int tid = threadIdx.x;
    const int nr = r1 - r0;
    __syncthreads();  // the previous tile's readers of s_rp / s_prod are done
    for (int i = tid; i <= nr; i += kBlock) s_rp[i] = A.rowptr[r0 + i];
    __syncthreads();
    const int e0 = s_rp[0], e1 = s_rp[nr];
    T acc[NRHS];
    int my_b = 0, my_e = 0;
    const bool has_row = tid < nr;
    if (has_row) {
        my_b = s_rp[tid];
        my_e = s_rp[tid + 1];
#pragma unroll
        for (int v = 0; v < NRHS; ++v) acc[v] = SUBTRACT ? init(r0 + tid, v) : T(0);
    }
    for (int cs = e0; cs < e1; cs += kLdsCap) {
        const int ce = min(cs + kLdsCap, e1);
        for (int e = cs + tid; e < ce; e += kBlock) {
            const int c = A.col[e];
            const T a = A.val[e];
            T gv[NRHS];
            g(c, gv);
#pragma unroll
            for (int v = 0; v < NRHS; ++v) s_prod[v * kLdsCap + (e - cs)] = a * gv[v];
        }
        __syncthreads();
        if (has_row) {
            const int b = max(my_b, cs), e = min(my_e, ce);
            for (int k = b; k < e; ++k) {
#pragma unroll
                for (int v = 0; v < NRHS; ++v) {
                    if (SUBTRACT) acc[v] = acc[v] - s_prod[v * kLdsCap + (k - cs)];
                    else acc[v] = acc[v] + s_prod[v * kLdsCap + (k - cs)];
                }
            }
        }
        __syncthreads();
    }
    if (has_row) epi(r0 + tid, acc);
}

// The same streaming scheme for the difference form: phase 1 stores a_ij
// and the gathered p_j; phase 2 forms sum a_ij (p_j - p_i) in column order
// and adds rs_i p_i.  With one right-hand side the pair (a_ij, p_j) is
// stored as one 8-byte LDS word, so each entry costs one LDS write and one
// read.
template <class Gather, class Epi>
__device__ __forceinline__ void stream_tile_diff1(const DiffCsr32& A, int r0, int r1, Gather& g, Epi& epi) {
    __shared__ int s_rp[kTileRows + 1];
    __shared__ float2 s_ag[kLdsCap];
    const int tid = threadIdx.x;
    const int nr = r1 - r0;
    __syncthreads();
    for (int i = tid; i <= nr; i += kBlock) s_rp[i] = A.rowptr[r0 + i];
    __syncthreads();
    const int e0 = s_rp[0], e1 = s_rp[nr];
    float acc[1] = {0.0f}, own[1] = {0.0f};
    int my_b = 0, my_e = 0;
    const bool has_row = tid < nr;
    if (has_row) {
        my_b = s_rp[tid];
        my_e = s_rp[tid + 1];
        g(r0 + tid, own);
    }
    for (int cs = e0; cs < e1; cs += kLdsCap) {
        const int ce = min(cs + kLdsCap, e1);
        for (int e = cs + tid; e < ce; e += kBlock) {
            const int c = A.col[e];
            const float a = A.val[e];
            float gv[1];
            g(c, gv);
            s_ag[e - cs] = make_float2(a, gv[0]);
        }
        __syncthreads();
        if (has_row) {
            const int b = max(my_b, cs), e = min(my_e, ce);
            for (int k = b; k < e; ++k) {
                const float2 ag = s_ag[k - cs];
                acc[0] = acc[0] + ag.x * (ag.y - own[0]);
            }
        }
        __syncthreads();
    }
    if (has_row) {
        const float rsi = A.rs[r0 + tid];
        acc[0] = acc[0] + rsi * own[0];
        epi(r0 + tid, acc);
    }
}

template <int NRHS, class Gather, class Epi>
__device__ __forceinline__ void stream_tile_diff(const DiffCsr32& A, int r0, int r1, Gather& g, Epi& epi) {
    if constexpr (NRHS == 1) {
        stream_tile_diff1(A, r0, r1, g, epi);
        return;
    }
    __shared__ int s_rp[kTileRows + 1];
    __shared__ float s_a[kLdsCap];
    __shared__ float s_g[NRHS * kLdsCap];
    const int tid = threadIdx.x;
    const int nr = r1 - r0;
    __syncthreads();
    for (int i = tid; i <= nr; i += kBlock) s_rp[i] = A.rowptr[r0 + i];
    __syncthreads();
    const int e0 = s_rp[0], e1 = s_rp[nr];
    float acc[NRHS], own[NRHS];
    int my_b = 0, my_e = 0;
    const bool has_row = tid < nr;
    if (has_row) {
        my_b = s_rp[tid];
        my_e = s_rp[tid + 1];
        g(r0 + tid, own);
#pragma unroll
        for (int v = 0; v < NRHS; ++v) acc[v] = 0.0f;
    }
    for (int cs = e0; cs < e1; cs += kLdsCap) {
        const int ce = min(cs + kLdsCap, e1);
        for (int e = cs + tid; e < ce; e += kBlock) {
            const int c = A.col[e];
            s_a[e - cs] = A.val[e];
            float gv[NRHS];
            g(c, gv);
#pragma unroll
            for (int v = 0; v < NRHS; ++v) s_g[v * kLdsCap + (e - cs)] = gv[v];
        }
        __syncthreads();
        if (has_row) {
            const int b = max(my_b, cs), e = min(my_e, ce);
            for (int k = b; k < e; ++k) {
                const float a = s_a[k - cs];
#pragma unroll
                for (int v = 0; v < NRHS; ++v) acc[v] = acc[v] + a * (s_g[v * kLdsCap + (k - cs)] - own[v]);
            }
        }
        __syncthreads();
    }
    if (has_row) {
        const float rsi = A.rs[r0 + tid];
#pragma unroll
        for (int v = 0; v < NRHS; ++v) acc[v] = acc[v] + rsi * own[v];
        epi(r0 + tid, acc);
    }
}

template <typename T>
struct InitZero {
    __device__ T operator()(int, int) const { return T(0); }
};
template <typename T>
struct InitVec {
    const T* __restrict__ b;
    __device__ T operator()(int r, int) const { return b[r]; }
};
template <typename T>
struct Gather1 {
    const T* __restrict__ x;
    __device__ void operator()(int c, T (&g)[1]) const { g[0] = x[c]; }
};
template <typename T>
struct Gather2 {
    const T* __restrict__ x;
    const T* __restrict__ y;
    __device__ void operator()(int c, T (&g)[2]) const { g[0] = x[c]; g[1] = y[c]; }
};

// ---------------------------------------------------------------------------
// Work split under a BatchCtx.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void ctx_tiles(const BatchCtx& c, int n, int& t0, int& t1) {
    if (!c.blk_sys) {
        block_tiles((n + kTileRows - 1) / kTileRows, t0, t1);
    } else {
        t0 = c.blk_tiles[2 * blockIdx.x];
        t1 = c.blk_tiles[2 * blockIdx.x + 1];
    }
}
// Tiles of this block for the SpMV kernels.
__device__ __forceinline__ TileSpan ctx_span(const BatchCtx& c, int n) {
    if (!c.blk_sys) {
        const int ntiles = (n + kTileRows - 1) / kTileRows;
        if (c.strided) return block_span_strided(ntiles);
        int t0, t1;
        block_tiles(ntiles, t0, t1);
        return TileSpan{t0, t1, 1};
    }
    return TileSpan{c.blk_tiles[2 * blockIdx.x], c.blk_tiles[2 * blockIdx.x + 1], 1};
}
__device__ __forceinline__ void ctx_tile_rows(const BatchCtx& c, int n, int t, int& r0, int& r1) {
    if (!c.tile_r0) {
        r0 = t * kTileRows;
        r1 = min(r0 + kTileRows, n);
    } else {
        r0 = c.tile_r0[t];
        r1 = c.tile_r0[t + 1];
    }
}
__device__ __forceinline__ void ctx_rows(const BatchCtx& c, int n, int& rb, int& re) {
    int t0, t1;
    ctx_tiles(c, n, t0, t1);
    if (!c.tile_r0) {
        rb = t0 * kTileRows;
        re = min(t1 * kTileRows, n);
    } else {
        rb = c.tile_r0[t0];
        re = c.tile_r0[t1];
    }
}
template <typename T, int NV>
__device__ __forceinline__ void ctx_total(const BatchCtx& c, const T* __restrict__ partials, T (&tot)[NV]) {
    if (!c.blk_sys) {
        site_total<T, NV>(partials, tot);
    } else {
        const int s = c.sys();
        site_total_range<T, NV>(partials, c.sys_blk0[s], c.sys_blk0[s + 1], tot);
    }
}

// ---------------------------------------------------------------------------
// Lane-group forms: L lanes (L = 2, 4, 8; aligned groups inside a 16-lane
// DPP row) share one row.  Lane j of the group loads entries e0+j, e0+j+L,
// ... of the row, so a wavefront's col/val loads cover 64/L consecutive rows
// contiguously with no LDS staging and no barrier.  The products of each
// chunk of L entries are folded into the row sum left to right: after the
// DPP row_shr:1 step i, lane i holds acc + p_0 + ... + p_i, so the row is
// summed in column order exactly as the thread-per-row form does.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T shr1(T v);
template <>
__device__ __forceinline__ double shr1<double>(double v) { return dpp_d<0x111>(v); }
template <>
__device__ __forceinline__ float shr1<float>(float v) { return dpp_f<0x111>(v); }

template <typename T>
__device__ __forceinline__ T group_last(T v, int L) {   // value of the group's last lane
    const int src = (int)(threadIdx.x & (kWave - 1) & ~(L - 1)) + L - 1;
    return __shfl(v, src, kWave);
}

template <int L, typename T, int NRHS, bool SUBTRACT, class Gather, class Init, class Epi>
__device__ __forceinline__ void group_tile(const Csr<T>& A, int r0, int r1, Gather& g, Init& init, Epi& epi) {
    constexpr int kRowsPerPass = kBlock / L;
    const int j = threadIdx.x & (L - 1);
    for (int rb = r0; rb < r1; rb += kRowsPerPass) {
        const int row = rb + (int)threadIdx.x / L;
        if (row >= r1) continue;                  // the whole group leaves together
        const int e0 = A.rowptr[row], e1 = A.rowptr[row + 1];
        T acc[NRHS];
#pragma unroll
        for (int v = 0; v < NRHS; ++v) acc[v] = SUBTRACT ? init(row, v) : T(0);
        for (int cs = e0; cs < e1; cs += L) {
            const int e = cs + j;
            const bool ok = e < e1;
            T prod[NRHS];
            if (ok) {
                T gv[NRHS];
                const T a = A.val[e];
                g(A.col[e], gv);
#pragma unroll
                for (int v = 0; v < NRHS; ++v) prod[v] = a * gv[v];
            } else {
#pragma unroll
                for (int v = 0; v < NRHS; ++v) prod[v] = T(0);
            }
            T s[NRHS];
#pragma unroll
            for (int v = 0; v < NRHS; ++v) s[v] = ok ? (SUBTRACT ? acc[v] - prod[v] : acc[v] + prod[v]) : acc[v];
#pragma unroll
            for (int i = 1; i < L; ++i) {
#pragma unroll
                for (int v = 0; v < NRHS; ++v) {
                    const T t = shr1(s[v]);
                    if (j == i) s[v] = ok ? (SUBTRACT ? t - prod[v] : t + prod[v]) : t;
                }
            }
#pragma unroll
            for (int v = 0; v < NRHS; ++v) acc[v] = group_last(s[v], L);
        }
        if (j == 0) epi(row, acc);
    }
}

template <int L, int NRHS, class Gather, class Epi>
__device__ __forceinline__ void group_tile_diff(const DiffCsr32& A, int r0, int r1, Gather& g, Epi& epi) {
    constexpr int kRowsPerPass = kBlock / L;
    const int j = threadIdx.x & (L - 1);
    for (int rb = r0; rb < r1; rb += kRowsPerPass) {
        const int row = rb + (int)threadIdx.x / L;
        if (row >= r1) continue;
        const int e0 = A.rowptr[row], e1 = A.rowptr[row + 1];
        float own[NRHS], acc[NRHS];
        g(row, own);
#pragma unroll
        for (int v = 0; v < NRHS; ++v) acc[v] = 0.0f;
        for (int cs = e0; cs < e1; cs += L) {
            const int e = cs + j;
            const bool ok = e < e1;
            float term[NRHS];
            if (ok) {
                float gv[NRHS];
                const float a = A.val[e];
                g(A.col[e], gv);
#pragma unroll
                for (int v = 0; v < NRHS; ++v) term[v] = a * (gv[v] - own[v]);
            } else {
#pragma unroll
                for (int v = 0; v < NRHS; ++v) term[v] = 0.0f;
            }
            float s[NRHS];
#pragma unroll
            for (int v = 0; v < NRHS; ++v) s[v] = ok ? acc[v] + term[v] : acc[v];
#pragma unroll
            for (int i = 1; i < L; ++i) {
#pragma unroll
                for (int v = 0; v < NRHS; ++v) {
                    const float t = shr1(s[v]);
                    if (j == i) s[v] = ok ? t + term[v] : t;
                }
            }
#pragma unroll
            for (int v = 0; v < NRHS; ++v) acc[v] = group_last(s[v], L);
        }
        if (j == 0) {
            const float rsi = A.rs[row];
#pragma unroll
            for (int v = 0; v < NRHS; ++v) acc[v] = acc[v] + rsi * own[v];
            epi(row, acc);
        }
    }
}

// q = A p over one tile for either operator; L = lanes per row (1 = the
// LDS-staged thread-per-row form).
template <int NRHS, int L = 1, typename T, class Gather, class Epi>
__device__ __forceinline__ void op_tile(const Csr<T>& A, int r0, int r1, Gather& g, Epi& epi) {
    InitZero<T> z;
    if constexpr (L == 1) stream_tile<T, NRHS, false>(A, r0, r1, g, z, epi);
    else group_tile<L, T, NRHS, false>(A, r0, r1, g, z, epi);
}
template <int NRHS, int L = 1, class Gather, class Epi>
__device__ __forceinline__ void op_tile(const DiffCsr32& A, int r0, int r1, Gather& g, Epi& epi) {
    if constexpr (L == 1) stream_tile_diff<NRHS>(A, r0, r1, g, epi);
    else group_tile_diff<L, NRHS>(A, r0, r1, g, epi);
}
// Rows [r0, r1) of the sliced operator, row r0 + t on thread t (no early
// return: the caller's block reductions follow).
// The slice of a row.  When the tile starts on a slice boundary (every
// tile of one system does) a wavefront's rows are one slice: the index is
// then made wavefront-uniform, so the slice's offsets and width come from
// scalar loads and the slot loop is uniform.
__device__ __forceinline__ int row_slice(int r0, int row) {
    const int sl = row / kSlice;
    return (r0 & (kSlice - 1)) == 0 ? __builtin_amdgcn_readfirstlane(sl) : sl;
}

// WIDE: the kernel also holds the all-slot loop (chosen at run time by
// A.unroll8).  Only the single-system SpMV kernels hold it: in the batch
// kernel the extra registers cost more than the loop gains.
template <int NRHS, bool WIDE = false, class Gather, class Epi>
__device__ __forceinline__ void sell_tile(const DiffSell32& A, int r0, int r1, Gather& g, Epi& epi) {
    const int row = r0 + (int)threadIdx.x;
    if (row < r1) {
        const int sl = row_slice(r0, row);
        const int b = A.off[sl];
        const int w = (A.off[sl + 1] - b) / kSlice;
        const float* __restrict__ vp = A.val + b + (row - sl * kSlice);
        float own[NRHS], acc[NRHS];
        g(row, own);
#pragma unroll
        for (int v = 0; v < NRHS; ++v) acc[v] = 0.0f;
        auto slot = [&](int c, float a) {
            float gv[NRHS];
            g(c, gv);
#pragma unroll
            for (int v = 0; v < NRHS; ++v) acc[v] = acc[v] + a * (gv[v] - own[v]);
        };
        if (A.lbase) {                               // symmetric: upper slots stored, lower ones mirrored
            const int sb = A.sidx[sl];
            const int ws = A.sidx[sl + 1] - sb;
            const int nl = ws - w;
            const int* __restrict__ so = A.soff + sb;
            const unsigned m = A.mask[row];
            // as in the full form, a missing slot adds a (p_i - p_i) with a = 0;
            // a lower slot's load stays in bounds (index 0) when the row lacks
            // it.  The lower slots come first in the ascending offsets: two
            // straight-line loops keep each slot's loads independent
#pragma unroll 2
            for (int j = 0; j < nl; ++j) {
                const bool has = ((m >> j) & 1u) != 0;
                const int d = so[j];
                const int pos = sym_mirror_pos(A.lbase, sb, j, sl, row, d);
                const float av = A.val[has ? pos : 0];
                slot(has ? row + d : row, has ? av : 0.0f);
            }
#pragma unroll 2
            for (int j = nl; j < ws; ++j) {
                const bool has = ((m >> j) & 1u) != 0;
                slot(has ? row + so[j] : row, vp[(j - nl) * kSlice]);
            }
        } else if (A.soff) {                         // offsets shared by the slice
            // a slot the row lacks takes the row's own column with a = 0 and
            // adds a (p_i - p_i) = +-0, which leaves every sum as it is (one
            // that starts at +0 never becomes -0); no branch, so the loads of
            // all slots are in flight together
            const int* __restrict__ so = A.soff + b / kSlice;
            const unsigned m = A.mask[row];
            if (WIDE && A.unroll8) {
                if (w > 0) {
#pragma unroll
                    for (int j = 0; j < kMaxSharedSlots; ++j) {
                        const bool in = j < w;
                        const int jj = in ? j : 0;
                        const float a = vp[jj * kSlice];
                        const bool has = in && ((m >> j) & 1u);
                        slot(has ? row + so[jj] : row, in ? a : 0.0f);
                    }
                }
            } else {
#pragma unroll 2
                for (int j = 0; j < w; ++j) slot(((m >> j) & 1u) ? row + so[j] : row, vp[j * kSlice]);
            }
        } else if (A.dc) {                           // the same columns, 16-bit
            const short* __restrict__ dp = A.dc + b + (row - sl * kSlice);
#pragma unroll 2
            for (int j = 0; j < w; ++j) slot(row + (int)dp[j * kSlice], vp[j * kSlice]);
        } else {
            const int* __restrict__ cp = A.col + b + (row - sl * kSlice);
#pragma unroll 2
            for (int j = 0; j < w; ++j) slot(cp[j * kSlice], vp[j * kSlice]);
        }
        const float rsi = A.rs[row];
#pragma unroll
        for (int v = 0; v < NRHS; ++v) acc[v] = acc[v] + rsi * own[v];
        epi(row, acc);
    }
}
template <int NRHS, int L = 1, class Gather, class Epi>
__device__ __forceinline__ void op_tile(const DiffSell32& A, int r0, int r1, Gather& g, Epi& epi) {
    sell_tile<NRHS>(A, r0, r1, g, epi);
}

// Rows [r0, r1) of a sliced matrix, row r0 + t on thread t: products
// rounded, then added to 0 (or subtracted from init) in CSR order.
template <typename T, int NRHS, bool SUBTRACT, class Gather, class Init, class Epi>
__device__ __forceinline__ void sell_rows(const Sell<T>& A, int r0, int r1, Gather& g, Init& init, Epi& epi) {
    const int row = r0 + (int)threadIdx.x;
    if (row < r1) {
        const int sl = row_slice(r0, row);
        const int b = A.off[sl];
        const int w = (A.off[sl + 1] - b) / kSlice;
        const int* __restrict__ cp = A.col + b + (row - sl * kSlice);
        const T* __restrict__ vp = A.val + b + (row - sl * kSlice);
        T acc[NRHS];
#pragma unroll
        for (int v = 0; v < NRHS; ++v) acc[v] = SUBTRACT ? init(row, v) : T(0);
        auto slot = [&](int c, T a) {
            T gv[NRHS];
            g(c, gv);
#pragma unroll
            for (int v = 0; v < NRHS; ++v) {
                const T prod = a * gv[v];
                acc[v] = SUBTRACT ? acc[v] - prod : acc[v] + prod;
            }
        };
        if (A.lbase) {                               // symmetric: upper slots stored, lower ones mirrored
            const int sb = A.sidx[sl];
            const int ws = A.sidx[sl + 1] - sb;
            const int nl = ws - w;
            const int* __restrict__ so = A.soff + sb;
            const unsigned m = A.mask[row];
            // the lower slots (mirrored) come first in the ascending offsets
            // (two slots at a time: a branch-free loop over all slots, the
            // FP32 operator's form, held more registers and ran slower here)
#pragma unroll 2
            for (int j = 0; j < nl; ++j)
                if ((m >> j) & 1u) slot(row + so[j], A.val[sym_mirror_pos(A.lbase, sb, j, sl, row, so[j])]);
#pragma unroll 2
            for (int j = nl; j < ws; ++j)
                if ((m >> j) & 1u) slot(row + so[j], vp[(j - nl) * kSlice]);
        } else if (A.soff) {                         // offsets shared by the slice
            const int* __restrict__ so = A.soff + b / kSlice;
            const unsigned m = A.mask[row];
#pragma unroll 2
            for (int j = 0; j < w; ++j)
                if ((m >> j) & 1u) slot(row + so[j], vp[j * kSlice]);
        } else if (A.dc) {                           // the same columns, 16-bit
            const short* __restrict__ dp = A.dc + b + (row - sl * kSlice);
#pragma unroll 2
            for (int j = 0; j < w; ++j) {
                const short d = dp[j * kSlice];
                if (d != kPad16) slot(row + (int)d, vp[j * kSlice]);
            }
        } else {
#pragma unroll 2
            for (int j = 0; j < w; ++j) {
                const int c = cp[j * kSlice];
                if (c >= 0) slot(c, vp[j * kSlice]);
            }
        }
        epi(row, acc);
    }
}
template <int NRHS, int L = 1, typename T, class Gather, class Epi>
__device__ __forceinline__ void op_tile(const Sell<T>& A, int r0, int r1, Gather& g, Epi& epi) {
    InitZero<T> z;
    sell_rows<T, NRHS, false>(A, r0, r1, g, z, epi);
}

// r = b - A x over one tile (the residual form: subtraction from b).
template <int L, typename T, class Gather, class Init, class Epi>
__device__ __forceinline__ void resid_tile(const Csr<T>& A, int r0, int r1, Gather& g, Init& init, Epi& epi) {
    if constexpr (L == 1) stream_tile<T, 1, true>(A, r0, r1, g, init, epi);
    else group_tile<L, T, 1, true>(A, r0, r1, g, init, epi);
}
// The same on the sliced copy: the products subtracted from b in CSR
// order, the same bits.
template <int L, typename T, class Gather, class Init, class Epi>
__device__ __forceinline__ void resid_tile(const Sell<T>& A, int r0, int r1, Gather& g, Init& init, Epi& epi) {
    sell_rows<T, 1, true>(A, r0, r1, g, init, epi);
}

}  // namespace tflows
