// krylov_kernels.h -- host-side declarations of the kernel launchers
// (definitions and explicit instantiations for double and float live in
// krylov_kernels.hip).
#pragma once
#include "kernels.h"

namespace tflows {

enum { RHO_NONE_ = 0, RHO_JAC_ = 1, RHO_BICG_ = 2, RHO_RR_ = 3 };

// Partial-sum buffers of the four reduction sites of a solve; each holds
// 2 values x `grid` blocks.  Producer and consumer kernels of a site are
// launched with the same grid.
template <typename T>
struct Sites {
    T* A;
    T* B;
    T* C;
    T* D;
    int grid;
    static Sites carve(T* base, int grid) { return Sites{base, base + 2 * grid, base + 4 * grid, base + 6 * grid, grid}; }
    static size_t elems(int grid) { return (size_t)8 * grid; }
};

// Solver-kernel launchers.  `c` is the launch context: BatchCtx{} for one
// system, or the maps of a block-diagonal batch (st and hm then point at
// the first of the per-system arrays).
template <typename T>
void launch_setup(const BatchCtx& c, const Csr<T>& A, int prec, bool write_d, T* D, const T* b, const Sites<T>& S,
                  hipStream_t s);
template <typename T>
void launch_dic_setup_level(const Csr<T>& A, const int* rows, int count, T* D, hipStream_t s);
template <typename T>
void launch_dic_fwd_level(const Csr<T>& A, const int* rows, int count, const T* D, const T* u, T* z,
                          const SolverState* st, hipStream_t s);
template <typename T>
void launch_dic_bwd_level(const Csr<T>& A, const int* rows, int count, const T* D, T* z,
                          const SolverState* st, hipStream_t s);
template <typename T>
void launch_resid_init(int mode, bool zero_x, const BatchCtx& c, const Csr<T>& A, const T* x, T* r, T* r2,
                       const T* D, const Sites<T>& S, SolverState* st, HostMirror* hm, hipStream_t s);
template <typename T>
void launch_check(const BatchCtx& c, int k, const Sites<T>& S, SolverState* st, HostMirror* hm, hipStream_t s);
template <typename T>
void launch_rho(const BatchCtx& c, int n, const T* a, const T* b, const Sites<T>& S, const SolverState* st,
                hipStream_t s);
template <typename T>
void launch_cg_p(const BatchCtx& c, bool jac, int k, int n, const T* r, const T* D, const T* z, T* p,
                 const Sites<T>& S, SolverState* st, HostMirror* hm, hipStream_t s);
template <typename Op, typename T>
void launch_cg_spmv(const BatchCtx& c, const Op& A, const T* p, T* q, const Sites<T>& S, const SolverState* st,
                    hipStream_t s);
template <typename T>
void launch_cg_update(const BatchCtx& c, bool jac, int k, int n, T* x, T* r, const T* p, const T* q, const T* D,
                      const Sites<T>& S, const SolverState* st, hipStream_t s);
// Jacobi CG of one system with x += alfa p carried from iteration k's update
// kernel into iteration k + 1's p kernel (which reads p anyway); after the
// solve stops, launch_cg_xfinish applies the update left pending.
template <typename T>
void launch_cg_p_carry(const BatchCtx& c, int k, int n, const T* r, const T* D, T* p, T* x, const Sites<T>& S,
                       SolverState* st, HostMirror* hm, hipStream_t s);
template <typename T>
void launch_cg_update_carry(const BatchCtx& c, int k, int n, T* r, const T* q, const T* D, const Sites<T>& S,
                            SolverState* st, hipStream_t s);
template <typename T>
void launch_cg_xfinish(int n, T* x, const T* p, const SolverState* st, hipStream_t s);
template <typename T>
void launch_bicg_p(const BatchCtx& c, bool jac, int k, int n, const T* r1, const T* r2, const T* D, const T* z1,
                   const T* z2, T* p1, T* p2, const Sites<T>& S, SolverState* st, HostMirror* hm, hipStream_t s);
template <typename Op, typename T>
void launch_bicg_spmv(const BatchCtx& c, const Op& A, const T* p1, const T* p2, T* q1, T* q2, const Sites<T>& S,
                      const SolverState* st, hipStream_t s);
template <typename T>
void launch_bicg_update(const BatchCtx& c, bool jac, int k, int n, T* x, T* r1, T* r2, const T* p1, const T* q1,
                        const T* q2, const T* D, const Sites<T>& S, const SolverState* st, hipStream_t s);
template <typename T>
void launch_cgs_u(const BatchCtx& c, bool jac, int k, int n, const T* r1, const T* q1, T* u1, T* u2, const T* D,
                  T* p2, const Sites<T>& S, SolverState* st, HostMirror* hm, hipStream_t s);
template <typename Op, typename T>
void launch_cgs_spmv_v(const BatchCtx& c, const Op& A, const T* p2, T* v2, const T* r2, const Sites<T>& S,
                       const SolverState* st, hipStream_t s);
template <typename T>
void launch_cgs_q(const BatchCtx& c, bool jac, int k, int n, const T* u1, const T* v2, T* q1, const T* D, T* p1,
                  const Sites<T>& S, const SolverState* st, hipStream_t s);
template <typename Op, typename T>
void launch_cgs_spmv_r(const BatchCtx& c, int k, const Op& A, const T* p1, T* x, T* r1, const T* r2,
                       const Sites<T>& S, const SolverState* st, hipStream_t s);
// stand-alone q = A p and r -= A x of one system; the context selects the
// SpMV form (lanes) and tile order
template <typename Op, typename T>
void launch_spmv(const Op& A, const T* x, T* y, int grid, hipStream_t s, const BatchCtx& c = BatchCtx{});
template <typename T>
void launch_resid(const Csr<T>& A, const T* x, T* r, int grid, hipStream_t s, const BatchCtx& c = BatchCtx{});
// fused two-kernel CG (cg_fused.hip)
template <typename Op, typename T>
void launch_cgf_spmv(bool mixed, const Op& A, const T* z, T* p0, T* p1, T* q, const Sites<T>& S,
                     const double* partE, SolverState* st, HostMirror* hm, hipStream_t s);
template <typename T>
void launch_cgf_update(int n, T* x, T* r, T* z, const T* p0, const T* p1, const T* q, const T* D,
                       const Sites<T>& S, SolverState* st, hipStream_t s);
template <typename T>
void launch_cgf_check(bool mixed, const Sites<T>& S, const double* partE, SolverState* st, HostMirror* hm,
                      hipStream_t s);
// three-kernel Jacobi CG with device-side iteration counters (cg_fused.hip);
// with `carry` each x += alfa p goes to the next iteration's p kernel, and
// the solve's end applies the one left pending (launch_cg_xfinish, or the
// FP64 merge of the mixed-precision path)
template <typename Op, typename T>
void launch_cgj(bool mixed, bool strided, const Op& A, int n, T* x, T* r, T* z, T* p, T* q, const T* D, const Sites<T>& S,
                const double* partE, SolverState* st, HostMirror* hm, hipStream_t s, bool carry);
template <typename T>
void launch_cgf_init_z(int n, const T* r, const T* D, T* z, const Sites<T>& S, const SolverState* st,
                       hipStream_t s);

// Persistent batched CG with Jacobi (or no) preconditioning: one launch
// runs `chunk` iterations (kbase + 1 .. kbase + chunk) of every system of a
// batch, the three kernels of an iteration as three phases separated by
// barriers over each system's blocks (bar: one counter per system, zeroed
// by the launcher).  The grid (S.grid) must be resident all at once:
// batch_cg_chunk_resident gives how many blocks the device holds.  A
// barrier wait beyond wait_ticks raises *fault and every block stops.
template <typename Op, typename T>
void launch_batch_cg_chunk(const BatchCtx& c, const Op& A, int n, int chunk, T* x, T* r, T* p, T* q, const T* D,
                           const Sites<T>& S, SolverState* st, HostMirror* hm, int* bar, int nsys, int* fault,
                           uint64_t wait_ticks, hipStream_t s);
template <typename Op>
int batch_cg_chunk_resident(int lanes);

void launch_f64_to_f32(int64_t count, const double* in, float* out, hipStream_t s);
void launch_set_int(int* p, int v, hipStream_t s);
// *p = v, ordered on stream s (a kernel, so it may sit inside a graph)
void launch_set_double(double* p, double v, hipStream_t s);
// level-major DIC sweeps (see Levels in system.h)
template <typename T>
void launch_plan_in(int n, const int* order, const T* u, const T* D, T* up, T* dp, const SolverState* st,
                    hipStream_t s);
template <typename T>
void launch_plan_out(int n, const int* order, const T* zp, T* z, const SolverState* st, hipStream_t s);
template <typename T>
void launch_flow_reset(int n, T* zf, hipStream_t s);
template <typename T>
void launch_flow_fwd(int n, const int* order, const T* u, const T* D, const int* lptr, const int* lcol,
                     const T* lval, T* dp, T* zf, T* zb, int* fault, uint64_t wait_ticks, const SolverState* st,
                     hipStream_t s);
template <typename T>
void launch_flow_bwd(int n, const int* order, T* z, const int* uptr, const int* ucol, const T* uval, const T* dp,
                     T* zf, T* zb, int* fault, uint64_t wait_ticks, const SolverState* st, hipStream_t s);
template <typename T>
void launch_plan_fwd(int a, int b, const int* lptr, const int* lcol, const T* lval, const T* dp, const T* up, T* zp,
                     const SolverState* st, hipStream_t s);
template <typename T>
void launch_plan_bwd(const int* pos, int count, const int* uptr, const int* ucol, const T* uval, const T* dp, T* zp,
                     const SolverState* st, hipStream_t s);
// tiled one-launch DIC sweeps and pivots (Levels mode 3, tile_plan.h)
struct TileSweep {
    int n = 0, slots = 0, waves = 0;
    const int* wstart = nullptr;  // waves + 1: each wavefront's tiles in wtiles
    const int* wtiles = nullptr;
    const int* tph = nullptr;     // phase starts, kTilePhases + 1 per tile
    const int* order = nullptr;   // position -> row
    int* fault = nullptr;
    uint64_t wait_ticks = 0;
};
// wavefronts a tiled launch with `slots` entry slots may use (all resident at once)
int tile_sweep_waves(int slots);
template <typename T>
void launch_tile_fwd(const TileSweep& ts, const T* u, const T* D, const int* lcol, const T* lval, T* dp, T* zf, T* zb,
                     const SolverState* st, hipStream_t s);
template <typename T>
void launch_tile_bwd(const TileSweep& ts, T* z, const int* ucol, const T* uval, const T* dp, T* zf, T* zb,
                     const SolverState* st, hipStream_t s);
void launch_tile_pivots(const TileSweep& ts, const double* dg, const int* lcol, const double* lval, double* D,
                        double* zs, hipStream_t s);
// dst[e] = src[map[e]], 0 where map[e] < 0
void launch_entry_gather(size_t count, const int* map, const double* src, double* dst, hipStream_t s);
void launch_advance(int* kbase, int c, hipStream_t s);

// Number of blocks of the row-streaming kernels for n rows: one block per
// 256-row tile up to 8 blocks per CU (2048 on MI355X), rounded to a
// multiple of 8 so that the XCD-aware block mapping is a bijection.
inline int stream_grid(int n, int max_blocks = 2048) {
    int tiles = (n + kTileRows - 1) / kTileRows;
    int g = tiles < max_blocks ? tiles : max_blocks;
    if (g < 1) g = 1;
    if (g >= kNumXcd) g = (g + kNumXcd - 1) / kNumXcd * kNumXcd;
    return g;
}

}  // namespace tflows
