// cg_fused.hip -- CG with the diagonal (or no) preconditioner in two fused
// kernels per iteration, in FP64 (the drop-in path) and FP32 (the inner
// solve of the mixed-precision path).
//
// Process/CG.f90:113-259 per iteration reads r, forms z = r/D, rho = (r,z),
// p = z + beta p, q = A p, alfa = rho/(p,q), x += alfa p, r -= alfa q,
// ||r||.  Here:
//
//   k_cgf_spmv    stopping test of the previous iteration; rho; beta;
//                 q = A p_new with p_new = z + beta p_old formed inside the
//                 SpMV gathers (so p needs no separate pass); writes p_new
//                 and q; (p_new, q) -> site D
//   k_cgf_update  alfa = rho/(p,q); x += alfa p; r -= alfa q; z = r/D;
//                 ||r||^2 and the next rho = (r,z) -> site B
//
// p alternates between two buffers by iteration parity (the SpMV reads the
// old p at neighbouring rows while it writes the new one).  Every value is
// formed with the reference's operations in the reference's order; only
// the dot-product summation order differs (block tree instead of a serial
// loop).
//
// The iteration index lives on the device: k_cgf_update stores the number
// of completed iterations in `it`, k_cgf_spmv stores the one in progress in
// `kcur`, and each kernel only reads the field the other writes.
#include <cstdint>
#include <cstdlib>
#include <initializer_list>
#include <type_traits>

#include "krylov_kernels.h"

namespace tflows {

namespace {
__device__ __forceinline__ bool leader() { return blockIdx.x == 0 && threadIdx.x == 0; }

// True when the solve is over or paused.  The leader of the same launch may
// set these flags (in its stopping test) while the block's waves start, so
// thread 0 reads them once and the whole block takes its answer: a block
// whose waves disagreed would sum the block totals of only some waves.
// Blocks that continue reach the leader's decision themselves, from the
// same totals.
__device__ __forceinline__ bool block_halted(const SolverState* st) {
    __shared__ int h;
    if (threadIdx.x == 0) h = st->done | st->suspend;
    __syncthreads();
    return h != 0;
}
}

// Stopping test before iteration kd (CG.f90:93-108 for kd == 1, :248-257
// otherwise; kd == cap + 1 closes the loop with niter = cap + 1).  In MIXED
// mode the decision is taken on the FP64 true residual (site E) when the
// last FP64 residual replacement came right before iteration kd; otherwise
// the FP32 recurrence residual only pauses the iterations until the next
// replacement once it reaches tol, or once it has dropped below
// kReplaceDrop times the FP64 norm of the last replacement (so the FP32
// recurrence never runs far past the accuracy FP32 can attain).  Returns
// true when iteration kd must not run.
constexpr double kReplaceDrop = 1e-2;

// The state fields a kernel's stopping test and scalars read, loaded
// together at its start: the test then waits for one memory round trip
// (these loads and the partial sums in flight together) instead of a chain
// of dependent ones.  A kernel never reads here a field it writes itself
// (err_rep is written only when the test does not read it).
struct StopIn {
    double tol, err_rep, rho0, rho1, alfa;
    int cap, replaced_before, it, kcur, halt, xupd;
    // rho of iteration k; a select, not an indexed array, so that the
    // struct stays in registers
    __device__ __forceinline__ double rho(int k) const { return (k & 1) ? rho1 : rho0; }
};
__device__ __forceinline__ StopIn load_stop_in(const SolverState* st) {
    StopIn in;
    in.tol = st->tol;
    in.err_rep = st->err_rep;
    in.rho0 = st->rho[0];
    in.rho1 = st->rho[1];
    in.cap = st->cap;
    in.replaced_before = st->replaced_before;
    in.it = st->it;
    in.kcur = st->kcur;
    in.halt = st->done | st->suspend;
    in.alfa = st->alfa;
    in.xupd = st->xupd;
    return in;
}

template <typename T, bool MIXED>
__device__ __forceinline__ bool stop_decision(int kd, T rr, const StopIn& in, const double* __restrict__ partE,
                                              SolverState* st, HostMirror* hm) {
    const double tol = in.tol;
    const int cap = in.cap;
    bool authoritative = true;
    double err;
    if (MIXED && in.replaced_before == kd) {
        double e64[1];
        site_total<double, 1>(partE, e64);
        err = sqrt(e64[0]);
    } else {
        err = sqrt((double)rr);
        authoritative = !MIXED;
    }
    if (leader() && authoritative) {
        st->error = err;
        if (MIXED) st->err_rep = err;
        if (kd == 1) {
            st->ini_res = err;
            st->ini_written = 1;
        }
    }
    if (MIXED && !authoritative && err < kReplaceDrop * in.err_rep) {
        if (leader()) st->suspend = 1;
        return true;
    }
    if (err < tol) {
        if (leader()) {
            if (authoritative) {
                st->done = 1;
                st->niter_out = kd - 1;
                publish(hm, 1, kd - 1);
            } else {
                st->suspend = 1;
            }
        }
        return true;
    }
    if (kd > cap) {
        if (leader()) {
            st->done = 1;
            st->niter_out = (cap > 0 ? cap : 0) + 1;
            publish(hm, 1, kd - 1);
        }
        return true;
    }
    return false;
}

// The stopping test on site B, whose two values are ||r||^2 and the next
// rho = (r, z).
template <typename T, bool MIXED>
__device__ __forceinline__ bool fused_check(int kd, const T* __restrict__ partB, const double* __restrict__ partE,
                                            SolverState* st, HostMirror* hm, T& rho_next) {
    const StopIn in = load_stop_in(st);
    T tot[2];
    site_total<T, 2>(partB, tot);
    rho_next = tot[1];
    return stop_decision<T, MIXED>(kd, tot[0], in, partE, st, hm);
}

template <typename T>
struct GatherPNew {
    const T* __restrict__ z;
    const T* __restrict__ p_old;
    T beta;
    bool first;
    __device__ __forceinline__ T at(int c) const {
        const T zc = z[c];
        return first ? zc : zc + beta * p_old[c];   // CG.f90:180, :185
    }
    __device__ void operator()(int c, T (&g)[1]) const { g[0] = at(c); }
};

template <typename T>
struct PNewEpi {
    GatherPNew<T> g;
    T* __restrict__ p_new;
    T* __restrict__ q;
    T acc[1];
    __device__ void operator()(int row, T (&y)[1]) {
        const T pr = g.at(row);
        p_new[row] = pr;
        q[row] = y[0];
        acc[0] += pr * y[0];                         // (p, q)  CG.f90:216
    }
};

template <typename Op, bool MIXED, typename T = typename OpValue<Op>::type>
__global__ __launch_bounds__(kBlock) void k_cgf_spmv(Op A, const T* __restrict__ z, T* __restrict__ p0,
                                                     T* __restrict__ p1, T* __restrict__ q,
                                                     const T* __restrict__ partB, const double* __restrict__ partE,
                                                     T* __restrict__ partD, SolverState* st, HostMirror* hm) {
    if (block_halted(st)) return;
    const int kd = st->it + 1;
    T rho;
    if (fused_check<T, MIXED>(kd, partB, partE, st, hm, rho)) return;
    if (leader()) {
        st->rho[kd & 1] = (double)rho;
        st->kcur = kd;
    }
    const bool first = kd == 1;
    const T beta = first ? T(0) : rho / (T)st->rho[(kd - 1) & 1];   // CG.f90:183
    T* p_new = (kd & 1) ? p1 : p0;
    const T* p_old = (kd & 1) ? p0 : p1;
    GatherPNew<T> g{z, p_old, beta, first};
    PNewEpi<T> epi{g, p_new, q, {T(0)}};
    const int ntiles = (A.n + kTileRows - 1) / kTileRows;
    int t0, t1;
    block_tiles(ntiles, t0, t1);
    for (int t = t0; t < t1; ++t) op_tile<1>(A, t * kTileRows, min((t + 1) * kTileRows, A.n), g, epi);
    site_write<T, 1>(epi.acc, partD);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cgf_update(int n, T* __restrict__ x, T* __restrict__ r,
                                                       T* __restrict__ z, const T* __restrict__ p0,
                                                       const T* __restrict__ p1, const T* __restrict__ q,
                                                       const T* __restrict__ D, const T* __restrict__ partD,
                                                       T* __restrict__ partB, SolverState* st) {
    if (st->done || st->suspend) return;
    const int kd = st->kcur;
    T pq[1];
    site_total<T, 1>(partD, pq);
    const T alfa = (T)st->rho[kd & 1] / pq[0];          // CG.f90:219
    const T* p = (kd & 1) ? p1 : p0;
    const int ntiles = (n + kTileRows - 1) / kTileRows;
    int t0, t1;
    block_tiles(ntiles, t0, t1);
    const int rb = t0 * kTileRows, re = min(t1 * kTileRows, n);
    T v[2] = {T(0), T(0)};
    for (int i = rb + threadIdx.x; i < re; i += kBlock) {
        x[i] = x[i] + alfa * p[i];                       // CG.f90:227
        const T ri = r[i] - alfa * q[i];                 // CG.f90:228
        r[i] = ri;
        const T zi = ri / D[i];                          // next z, CG.f90:126
        z[i] = zi;
        v[0] += ri * ri;                                 // ||r||^2, CG.f90:250
        v[1] += ri * zi;                                 // next rho, CG.f90:174
    }
    site_write<T, 2>(v, partB);
    if (leader()) st->it = kd;
}

template <typename T, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_cgf_check(const T* __restrict__ partB, const double* __restrict__ partE,
                                                      SolverState* st, HostMirror* hm) {
    if (block_halted(st)) return;
    T rho;
    fused_check<T, MIXED>(st->it + 1, partB, partE, st, hm, rho);
}

// z = r / D and the sums (r,r), (r,z) of the initial residual: site B of the
// first fused iteration (the residual itself is made by k_resid_init or by
// the FP64 replacement of the mixed path).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cgf_init_z(int n, const T* __restrict__ r, const T* __restrict__ D,
                                                       T* __restrict__ z, T* __restrict__ partB,
                                                       const SolverState* st) {
    const int ntiles = (n + kTileRows - 1) / kTileRows;
    int t0, t1;
    block_tiles(ntiles, t0, t1);
    const int rb = t0 * kTileRows, re = min(t1 * kTileRows, n);
    T v[2] = {T(0), T(0)};
    if (!st->done) {
        for (int i = rb + threadIdx.x; i < re; i += kBlock) {
            const T ri = r[i];
            const T zi = ri / D[i];
            z[i] = zi;
            v[0] += ri * ri;
            v[1] += ri * zi;
        }
    }
    site_write<T, 2>(v, partB);
}

// ---------------------------------------------------------------------------
// The same iteration in three kernels: the p update is its own streaming
// pass (z and p only), and the SpMV gathers p alone.  Each element-wise loop
// keeps kUnroll rows per thread in flight, issuing their loads before it
// uses any of them.  A thread takes its rows in packs of V consecutive rows
// (V = 1, or 16 bytes of T where every vector is 16-byte aligned: one
// dwordx4 access per pack); the U = kUnroll / V packs of a group lie
// kBlock * V rows apart, so V = 1 is the plain row-per-lane layout.
// ---------------------------------------------------------------------------
constexpr int kUnroll = 4;

template <typename T, int V>
struct alignas(sizeof(T) * V) Pack {
    T v[V];
};

template <typename T, int V>
__device__ __forceinline__ Pack<T, V> load_pack(const T* __restrict__ a, int i, int re, T fill) {
    Pack<T, V> r;
    if (i + V <= re) {
        r = *reinterpret_cast<const Pack<T, V>*>(a + i);
    } else {
#pragma unroll
        for (int e = 0; e < V; ++e) r.v[e] = i + e < re ? a[i + e] : fill;
    }
    return r;
}

template <typename T, int V>
__device__ __forceinline__ void store_pack(T* __restrict__ a, int i, int re, const Pack<T, V>& v) {
    if (i + V <= re) {
        *reinterpret_cast<Pack<T, V>*>(a + i) = v;
    } else {
#pragma unroll
        for (int e = 0; e < V; ++e)
            if (i + e < re) a[i + e] = v.v[e];
    }
}

// DEFER: the p kernel also applies x += alfa p of the iteration before
// (CG.f90:226-227), which its update kernel left with alfa in the state,
// unless a merge of the mixed-precision path has applied it already; the
// update kernel then reads neither x nor p.
template <typename T, bool MIXED, bool DEFER, int V>
__global__ __launch_bounds__(kBlock) void k_cgj_p(const T* __restrict__ z, T* __restrict__ p, T* __restrict__ x,
                                                  const T* __restrict__ partB, const double* __restrict__ partE,
                                                  int n, SolverState* st, HostMirror* hm) {
    constexpr int U = kUnroll / V;
    using P = Pack<T, V>;
    // issued together: the state, this thread's first group of rows and
    // (in site_total) the partial sums of the stopping test
    const StopIn in = load_stop_in(st);
    __shared__ int halt;                 // thread 0's flags, for the whole block (block_halted)
    if (threadIdx.x == 0) halt = in.halt;
    const int kd = in.it + 1;
    const int ntiles = (n + kTileRows - 1) / kTileRows;
    int t0, t1;
    block_tiles(ntiles, t0, t1);
    const int rb = t0 * kTileRows, re = min(t1 * kTileRows, n);
    int i0 = rb + (int)threadIdx.x * V;
    P zv[U], pv[U], xv[U];
    // x is loaded before the state says whether its update is pending (it
    // is, except right after a merge)
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = i0 + u * kBlock * V;
        zv[u] = load_pack<T, V>(z, i, re, T(0));
        pv[u] = load_pack<T, V>(p, i, re, T(0));
        if (DEFER) xv[u] = load_pack<T, V>(x, i, re, T(0));
    }
    T tot[2];
    site_total<T, 2>(partB, tot);        // its barriers publish `halt`
    if (halt) return;
    if (stop_decision<T, MIXED>(kd, tot[0], in, partE, st, hm)) return;
    const T rho = tot[1];
    if (leader()) {
        st->rho[kd & 1] = (double)rho;
        st->kcur = kd;
    }
    const bool first = kd == 1;
    const bool carry = DEFER && in.it > in.xupd;
    const T alfa_prev = (T)in.alfa;
    const T beta = first ? T(0) : rho / (T)in.rho(kd - 1);        // CG.f90:183
    for (;;) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * kBlock * V;
            P xo, po;
#pragma unroll
            for (int e = 0; e < V; ++e) {
                if (carry) xo.v[e] = xv[u].v[e] + alfa_prev * pv[u].v[e];            // CG.f90:227
                po.v[e] = first ? zv[u].v[e] : zv[u].v[e] + beta * pv[u].v[e];       // CG.f90:180, :185
            }
            if (carry) store_pack<T, V>(x, i, re, xo);
            store_pack<T, V>(p, i, re, po);
        }
        i0 += kUnroll * kBlock;
        if (i0 >= re) break;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * kBlock * V;
            zv[u] = load_pack<T, V>(z, i, re, T(0));
            pv[u] = load_pack<T, V>(p, i, re, T(0));
            if (carry) xv[u] = load_pack<T, V>(x, i, re, T(0));
        }
    }
}

template <typename T, bool DEFER, int V>
__global__ __launch_bounds__(kBlock) void k_cgj_update(int n, T* __restrict__ x, T* __restrict__ r,
                                                       T* __restrict__ z, const T* __restrict__ p,
                                                       const T* __restrict__ q, const T* __restrict__ D,
                                                       const T* __restrict__ partD, T* __restrict__ partB,
                                                       SolverState* st) {
    constexpr int U = kUnroll / V;
    using P = Pack<T, V>;
    // issued together: the state, this thread's first group of rows and
    // (in site_total) the partial sums of (p,q); the flags are not written
    // by this kernel, so every thread reads the same ones
    const StopIn in = load_stop_in(st);
    const int kd = in.kcur;
    const int ntiles = (n + kTileRows - 1) / kTileRows;
    int t0, t1;
    block_tiles(ntiles, t0, t1);
    const int rb = t0 * kTileRows, re = min(t1 * kTileRows, n);
    T v[2] = {T(0), T(0)};
    P xv[U], pv[U], rv[U], qv[U], dv[U];
    // loads of the next group of rows
    auto load = [&](int i0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * kBlock * V;
            if (!DEFER) {
                xv[u] = load_pack<T, V>(x, i, re, T(0));
                pv[u] = load_pack<T, V>(p, i, re, T(0));
            }
            rv[u] = load_pack<T, V>(r, i, re, T(0));
            qv[u] = load_pack<T, V>(q, i, re, T(0));
            dv[u] = load_pack<T, V>(D, i, re, T(1));
        }
    };
    const int ib = rb + (int)threadIdx.x * V;
    load(ib);
    T pq[1];
    site_total<T, 1>(partD, pq);
    if (in.halt) return;
    const T alfa = (T)in.rho(kd) / pq[0];               // CG.f90:219
    if (leader()) {
        if (DEFER) st->alfa = (double)alfa;
        st->it = kd;
    }
    for (int i0 = ib; i0 < re; i0 += kUnroll * kBlock) {
        if (i0 != ib) load(i0);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * kBlock * V;
            P xo, ro, zo;
#pragma unroll
            for (int e = 0; e < V; ++e) {
                if (!DEFER) xo.v[e] = xv[u].v[e] + alfa * pv[u].v[e];     // CG.f90:227
                const T ri = rv[u].v[e] - alfa * qv[u].v[e];              // CG.f90:228
                ro.v[e] = ri;
                const T zi = ri / dv[u].v[e];                             // next z, CG.f90:126
                zo.v[e] = zi;
                if (i + e < re) {
                    v[0] += ri * ri;                                      // ||r||^2, CG.f90:250
                    v[1] += ri * zi;                                      // next rho, CG.f90:174
                }
            }
            if (!DEFER) store_pack<T, V>(x, i, re, xo);
            store_pack<T, V>(r, i, re, ro);
            store_pack<T, V>(z, i, re, zo);
        }
    }
    site_write<T, 2>(v, partB);
}

template <typename Op, typename T = typename OpValue<Op>::type>
__global__ __launch_bounds__(kBlock) void k_cgj_spmv(Op A, const T* __restrict__ p, T* __restrict__ q,
                                                     T* __restrict__ partD, const SolverState* st, int strided) {
    if (st->done || st->suspend) return;
    BatchCtx one;
    one.strided = strided;
    const TileSpan sp = ctx_span(one, A.n);
    Gather1<T> g{p};
    struct Epi {
        const T* __restrict__ p;
        T* __restrict__ q;
        T acc[1];
        __device__ void operator()(int row, T (&y)[1]) {
            q[row] = y[0];
            acc[0] += p[row] * y[0];                     // (p, q)  CG.f90:216
        }
    } epi{p, q, {T(0)}};
    for (int t = sp.first; t < sp.end; t += sp.step) {
        if constexpr (std::is_same_v<Op, DiffSell32>)
            sell_tile<1, true>(A, t * kTileRows, min((t + 1) * kTileRows, A.n), g, epi);
        else
            op_tile<1>(A, t * kTileRows, min((t + 1) * kTileRows, A.n), g, epi);
    }
    site_write<T, 1>(epi.acc, partD);
}

#define TF_FL(k, grid, s) hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, s

template <typename Op, typename T>
void launch_cgf_spmv(bool mixed, const Op& A, const T* z, T* p0, T* p1, T* q, const Sites<T>& S,
                     const double* partE, SolverState* st, HostMirror* hm, hipStream_t s) {
    if (mixed) TF_FL((k_cgf_spmv<Op, true>), S.grid, s), A, z, p0, p1, q, S.B, partE, S.D, st, hm);
    else TF_FL((k_cgf_spmv<Op, false>), S.grid, s), A, z, p0, p1, q, S.B, partE, S.D, st, hm);
}
template <typename T>
void launch_cgf_update(int n, T* x, T* r, T* z, const T* p0, const T* p1, const T* q, const T* D,
                       const Sites<T>& S, SolverState* st, hipStream_t s) {
    TF_FL((k_cgf_update<T>), S.grid, s), n, x, r, z, p0, p1, q, D, S.D, S.B, st);
}
template <typename T>
void launch_cgf_check(bool mixed, const Sites<T>& S, const double* partE, SolverState* st, HostMirror* hm,
                      hipStream_t s) {
    if (mixed) TF_FL((k_cgf_check<T, true>), S.grid, s), S.B, partE, st, hm);
    else TF_FL((k_cgf_check<T, false>), S.grid, s), S.B, partE, st, hm);
}
template <typename T>
void launch_cgf_init_z(int n, const T* r, const T* D, T* z, const Sites<T>& S, const SolverState* st,
                       hipStream_t s) {
    TF_FL((k_cgf_init_z<T>), S.grid, s), n, r, D, z, S.B, st);
}

// 16-byte packs where every vector of the element-wise kernels allows them
// (TFLOWS_CG_PACK=0: one row per lane everywhere)
static bool packs16(std::initializer_list<const void*> ptrs) {
    const char* e = std::getenv("TFLOWS_CG_PACK");
    if (e && std::atoi(e) == 0) return false;
    for (const void* q : ptrs)
        if (reinterpret_cast<uintptr_t>(q) % 16 != 0) return false;
    return true;
}

template <typename T, int V>
static void launch_cgj_elementwise(bool mixed, bool carry, bool p_kernel, int n, T* x, T* r, T* z, T* p,
                                   const T* q, const T* D, const Sites<T>& S, const double* partE, SolverState* st,
                                   HostMirror* hm, hipStream_t s) {
    if (p_kernel) {
        if (carry) {
            if (mixed) TF_FL((k_cgj_p<T, true, true, V>), S.grid, s), (const T*)z, p, x, S.B, partE, n, st, hm);
            else TF_FL((k_cgj_p<T, false, true, V>), S.grid, s), (const T*)z, p, x, S.B, partE, n, st, hm);
        } else {
            if (mixed) TF_FL((k_cgj_p<T, true, false, V>), S.grid, s), (const T*)z, p, x, S.B, partE, n, st, hm);
            else TF_FL((k_cgj_p<T, false, false, V>), S.grid, s), (const T*)z, p, x, S.B, partE, n, st, hm);
        }
    } else {
        if (carry) TF_FL((k_cgj_update<T, true, V>), S.grid, s), n, x, r, z, (const T*)p, q, D, S.D, S.B, st);
        else TF_FL((k_cgj_update<T, false, V>), S.grid, s), n, x, r, z, (const T*)p, q, D, S.D, S.B, st);
    }
}

template <typename Op, typename T>
void launch_cgj(bool mixed, bool strided, const Op& A, int n, T* x, T* r, T* z, T* p, T* q, const T* D,
                const Sites<T>& S, const double* partE, SolverState* st, HostMirror* hm, hipStream_t s, bool carry) {
    constexpr int V = 16 / sizeof(T);
    const bool wide = packs16({x, r, z, p, q, D});
    auto elementwise = [&](bool p_kernel) {
        if (wide) launch_cgj_elementwise<T, V>(mixed, carry, p_kernel, n, x, r, z, p, q, D, S, partE, st, hm, s);
        else launch_cgj_elementwise<T, 1>(mixed, carry, p_kernel, n, x, r, z, p, q, D, S, partE, st, hm, s);
    };
    elementwise(true);
    TF_FL((k_cgj_spmv<Op>), S.grid, s), A, (const T*)p, q, S.D, st, strided ? 1 : 0);
    elementwise(false);
}
template void launch_cgj<Csr<double>, double>(bool, bool, const Csr<double>&, int, double*, double*, double*, double*,
                                              double*, const double*, const Sites<double>&, const double*,
                                              SolverState*, HostMirror*, hipStream_t, bool);
template void launch_cgj<Csr<float>, float>(bool, bool, const Csr<float>&, int, float*, float*, float*, float*, float*,
                                            const float*, const Sites<float>&, const double*, SolverState*,
                                            HostMirror*, hipStream_t, bool);
template void launch_cgj<DiffCsr32, float>(bool, bool, const DiffCsr32&, int, float*, float*, float*, float*, float*,
                                           const float*, const Sites<float>&, const double*, SolverState*,
                                           HostMirror*, hipStream_t, bool);

template void launch_cgj<Sell<double>, double>(bool, bool, const Sell<double>&, int, double*, double*, double*,
                                               double*, double*, const double*, const Sites<double>&, const double*,
                                               SolverState*, HostMirror*, hipStream_t, bool);
template void launch_cgf_spmv<Sell<double>, double>(bool, const Sell<double>&, const double*, double*, double*,
                                                    double*, const Sites<double>&, const double*, SolverState*,
                                                    HostMirror*, hipStream_t);
template void launch_cgj<DiffSell32, float>(bool, bool, const DiffSell32&, int, float*, float*, float*, float*,
                                            float*, const float*, const Sites<float>&, const double*, SolverState*,
                                            HostMirror*, hipStream_t, bool);
template void launch_cgf_spmv<DiffSell32, float>(bool, const DiffSell32&, const float*, float*, float*, float*,
                                                 const Sites<float>&, const double*, SolverState*, HostMirror*,
                                                 hipStream_t);
template void launch_cgf_spmv<Csr<double>, double>(bool, const Csr<double>&, const double*, double*, double*,
                                                   double*, const Sites<double>&, const double*, SolverState*,
                                                   HostMirror*, hipStream_t);
template void launch_cgf_spmv<Csr<float>, float>(bool, const Csr<float>&, const float*, float*, float*, float*,
                                                 const Sites<float>&, const double*, SolverState*, HostMirror*,
                                                 hipStream_t);
template void launch_cgf_spmv<DiffCsr32, float>(bool, const DiffCsr32&, const float*, float*, float*, float*,
                                                const Sites<float>&, const double*, SolverState*, HostMirror*,
                                                hipStream_t);
template void launch_cgf_update<double>(int, double*, double*, double*, const double*, const double*,
                                        const double*, const double*, const Sites<double>&, SolverState*,
                                        hipStream_t);
template void launch_cgf_update<float>(int, float*, float*, float*, const float*, const float*, const float*,
                                       const float*, const Sites<float>&, SolverState*, hipStream_t);
template void launch_cgf_check<double>(bool, const Sites<double>&, const double*, SolverState*, HostMirror*,
                                       hipStream_t);
template void launch_cgf_check<float>(bool, const Sites<float>&, const double*, SolverState*, HostMirror*,
                                      hipStream_t);
template void launch_cgf_init_z<double>(int, const double*, const double*, double*, const Sites<double>&,
                                        const SolverState*, hipStream_t);
template void launch_cgf_init_z<float>(int, const float*, const float*, float*, const Sites<float>&,
                                       const SolverState*, hipStream_t);


}  // namespace tflows
