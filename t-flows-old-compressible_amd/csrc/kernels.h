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
};

// Host-visible mirror of the progress, written by the device with
// system-scope stores into coherent pinned memory so the host can pace its
// launches without events.
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

// Matrix view with 0-based indices.
template <typename T>
struct Csr {
    int n;
    const int* __restrict__ rowptr;  // n+1
    const int* __restrict__ col;     // nnz
    const int* __restrict__ dia;     // n   (position of the diagonal)
    const T* __restrict__ val;       // nnz
};

// Rows per tile of the stream kernels: one row per thread in the summation
// phase.  Tiles are the unit of work split across blocks.
constexpr int kTileRows = kBlock;
// LDS products per right-hand side (per chunk).  A 7-point tile needs
// 256*7 = 1792; longer rows are handled by chunking.
template <typename T> struct LdsCap { static constexpr int value = 2048; };

// ---------------------------------------------------------------------------
// Stream one tile of rows [r0, r1):
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
__device__ __forceinline__ void stream_tile(const Csr<T>& A, int r0, int r1, Gather& g, Init& init,
                                            Epi& epi, int* s_rp, T* s_prod) {
    constexpr int CAP = LdsCap<T>::value;
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
    for (int cs = e0; cs < e1; cs += CAP) {
        const int ce = min(cs + CAP, e1);
        for (int e = cs + tid; e < ce; e += kBlock) {
            const int c = A.col[e];
            const T a = A.val[e];
            T gv[NRHS];
            g(c, gv);
#pragma unroll
            for (int v = 0; v < NRHS; ++v) s_prod[v * CAP + (e - cs)] = a * gv[v];
        }
        __syncthreads();
        if (has_row) {
            const int b = max(my_b, cs), e = min(my_e, ce);
            for (int k = b; k < e; ++k) {
#pragma unroll
                for (int v = 0; v < NRHS; ++v) {
                    if (SUBTRACT) acc[v] = acc[v] - s_prod[v * CAP + (k - cs)];
                    else acc[v] = acc[v] + s_prod[v * CAP + (k - cs)];
                }
            }
        }
        __syncthreads();
    }
    if (has_row) epi(r0 + tid, acc);
}

}  // namespace tflows
