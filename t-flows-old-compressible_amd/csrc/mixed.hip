// mixed.hip -- FP32 Krylov solves inside FP64 iterative refinement.
//
// The reference has only FP64 solves (every build line uses
// -fdefault-real-8, Process/Makefile:111, 130).  This component keeps the
// reference algorithm for the inner solve (the same CG / BiCG / CGS
// kernels, instantiated for float: SpMV, preconditioner, vector updates and
// dot products in FP32 with FP32 accumulation) and wraps it in an FP64
// refinement loop in which the true residual b - A x, the update of x and
// every convergence norm stay FP64:
//
//   r = b - A x                                   (FP64, Resid form)
//   while ||r|| >= tol:
//       s = ||r||;  r32 = fp32(r / s)             (scaled to unit norm)
//       solve A32 d = r32 with the FP32 Krylov loop to ||r32_k|| < eps
//       x = x + s * fp64(d)                       (FP64)
//       r = b - A x                               (FP64)
//
// The stopping test is the reference's absolute ||b - A x||_2 < tol, here
// evaluated on the true residual rather than on a recurrence.
//
// The FP32 SpMVs use the difference form of A (kernels.h, DiffCsr32): the
// pressure-correction matrix is singular up to one pinned cell, and the
// plain FP32 product loses the near-null mode (rounding of a_ii against
// sum a_ij exceeds its eigenvalue), which makes the refinement diverge.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <type_traits>
#include <vector>

#include "system.h"

namespace tflows {

template <typename T>
__global__ __launch_bounds__(kBlock) void k_norm2(int n, const T* __restrict__ v, T* __restrict__ partials) {
    const int ntiles = (n + kTileRows - 1) / kTileRows;
    int t0, t1;
    block_tiles(ntiles, t0, t1);
    const int rb = t0 * kTileRows, re = min(t1 * kTileRows, n);
    T a[1] = {T(0)};
    for (int i = rb + threadIdx.x; i < re; i += kBlock) a[0] += v[i] * v[i];
    site_write<T, 1>(a, partials);
}

// r32 = fp32(r64 * inv_s); x32 = 0
__global__ __launch_bounds__(kBlock) void k_scale_to_f32(int n, const double* __restrict__ r64, double inv_s,
                                                         float* __restrict__ r32, float* __restrict__ x32) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        r32[i] = (float)(r64[i] * inv_s);
        x32[i] = 0.0f;
    }
}

// x64 += s * fp64(d32);  r64 = b (ready for the Resid pass)
__global__ __launch_bounds__(kBlock) void k_refine(int n, double* __restrict__ x64, double s,
                                                   const float* __restrict__ d32, const double* __restrict__ b,
                                                   double* __restrict__ r64) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        x64[i] = x64[i] + s * (double)d32[i];
        r64[i] = b[i];
    }
}

// Build the FP32 difference-form operator from the FP64 CSR, one thread per
// row: off-diagonal entries keep their column order; rs_i is the FP64 sum of
// the row's values in column order, rounded once; diag32_i = fp32(a_ii).
__global__ __launch_bounds__(kBlock) void k_build_diff(int n, const int* __restrict__ rowptr,
                                                       const int* __restrict__ col, const int* __restrict__ dia,
                                                       const double* __restrict__ val, int* __restrict__ off_rowptr,
                                                       int* __restrict__ off_col, float* __restrict__ off_val,
                                                       float* __restrict__ rs, float* __restrict__ diag32) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int b = rowptr[i], e = rowptr[i + 1], d = dia[i];
    const int ob = b - i;                 // one diagonal entry per earlier row
    off_rowptr[i] = ob;
    if (i == n - 1) off_rowptr[n] = e - n;
    double sum = 0.0;
    int o = ob;
    for (int j = b; j < e; ++j) {
        const double a = val[j];
        sum += a;
        if (j == d) continue;
        off_col[o] = col[j];
        off_val[o] = (float)a;
        ++o;
    }
    rs[i] = (float)sum;
    diag32[i] = (float)val[d];
}

__global__ __launch_bounds__(kBlock) void k_f64_to_f32(int n, const double* __restrict__ a, float* __restrict__ b) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ __launch_bounds__(kBlock) void k_fill_f32(int n, float v, float* __restrict__ b) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_copy(int n, const double* __restrict__ a, double* __restrict__ b) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// FP64 refinement update of the reliable-update form: x64 += fp64(x32);
// x32 = 0.
__global__ __launch_bounds__(kBlock) void k_rel_merge(int n, double* __restrict__ x64, float* __restrict__ x32) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        x64[i] = x64[i] + (double)x32[i];
        x32[i] = 0.0f;
    }
}

// The same merge when the FP32 iterations carry x += alfa p into the next p
// kernel: the update of iteration `it`, if no p kernel has applied it yet,
// is applied here first (in FP32, as the iteration would have), and
// k_rel_resid then marks it applied.
__global__ __launch_bounds__(kBlock) void k_rel_merge_carry(int n, double* __restrict__ x64, float* __restrict__ x32,
                                                            const float* __restrict__ p,
                                                            const SolverState* __restrict__ st) {
    const bool carry = st->it > st->xupd;
    const float alfa = (float)st->alfa;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        float xs = x32[i];
        if (carry) xs = xs + alfa * p[i];
        x64[i] = x64[i] + (double)xs;
        x32[i] = 0.0f;
    }
}

// FP64 residual replacement: r = b - A x64 (Resid form, FP64), then
// r32 = fp32(r), z32 = r32/D32, with (r32,r32), (r32,z32) -> site B (FP32)
// and (r,r) -> site E (FP64).  The leader clears `suspend` and marks the
// next iteration as the one whose stopping test uses the FP64 norm.
struct RelEpi {
    float* __restrict__ r32;
    float* __restrict__ z32;
    const float* __restrict__ D32;
    double* __restrict__ r64;
    float a32[2];
    double a64[1];
    __device__ void operator()(int row, double (&y)[1]) {
        const double ri = y[0];
        if (r64) r64[row] = ri;
        const float rf = (float)ri;
        const float zf = rf / D32[row];
        r32[row] = rf;
        z32[row] = zf;
        a32[0] += rf * rf;
        a32[1] += rf * zf;
        a64[0] += ri * ri;
    }
};

template <class Op>
__global__ __launch_bounds__(kBlock) void k_rel_resid(Op A, const double* __restrict__ x64,
                                                      const double* __restrict__ b, float* __restrict__ r32,
                                                      float* __restrict__ z32, const float* __restrict__ D32,
                                                      double* __restrict__ r64, float* __restrict__ partB,
                                                      double* __restrict__ partE, SolverState* st) {
    // the SpMV's XCD-strided tile walk: rows a plane apart (the symmetric
    // operator's mirror values among them) are read by the same XCD while
    // still in its L2
    const TileSpan sp = block_span_strided((A.n + kTileRows - 1) / kTileRows);
    Gather1<double> g{x64};
    InitVec<double> init{b};
    RelEpi epi{r32, z32, D32, r64, {0.0f, 0.0f}, {0.0}};
    for (int t = sp.first; t < sp.end; t += sp.step)
        resid_tile<1>(A, t * kTileRows, min((t + 1) * kTileRows, A.n), g, init, epi);
    site_write<float, 2>(epi.a32, partB);
    site_write<double, 1>(epi.a64, partE);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->suspend = 0;
        st->replaced_before = st->it + 1;
        st->xupd = st->it;              // the merge before this pass applied x's pending update
    }
}

// Guards of the FP32 inner solves (their residual starts at norm 1):
//  * above 1e6 the rounding of FP32 recurrences (~1e-7 of the peak)
//    exceeds the unit right-hand side: the solve stops as diverged and its
//    correction is discarded;
//  * a solve whose smallest residual is older than max(100, the iteration
//    that reached it) has stalled: it stops and its correction is tried.
// A failed step (a diverged inner solve, or a correction that does not
// lower the FP64 residual) is undone and retried with a 10x looser inner
// target, up to 0.1.  CGS, whose residual polynomial is squared, needs
// this on transport matrices; a solve that still fails reports
// non-convergence.
constexpr double kInnerDivLimit = 1e6;
constexpr int kInnerStallMin = 100;
constexpr double kLoosestInnerRtol = 0.1;

// Sum of the block partials on the host, in block order.

static double host_site_sum(const double* d_partials, int grid, hipStream_t s) {
    std::vector<double> h(grid);
    TF_CHECK(hipMemcpyAsync(h.data(), d_partials, sizeof(double) * grid, hipMemcpyDeviceToHost, s));
    TF_CHECK(hipStreamSynchronize(s));
    double acc = 0.0;
    for (double v : h) acc += v;
    return acc;
}

DiffCsr32 System::diff32(hipStream_t s) {
    if (!diff32_valid_) {
        off_rowptr_.ensure(n_ + 1);
        off_col_.ensure(nnz_ - n_ > 0 ? nnz_ - n_ : 1);
        off_val_.ensure(nnz_ - n_ > 0 ? nnz_ - n_ : 1);
        rowsum32_.ensure(n_);
        diag32_.ensure(n_);
        hipLaunchKernelGGL(k_build_diff, dim3((n_ + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n_, rowptr_.p, col_.p,
                           dia_.p, (const double*)val_.p, off_rowptr_.p, off_col_.p, off_val_.p, rowsum32_.p,
                           diag32_.p);
        TF_CHECK(hipGetLastError());
        diff32_valid_ = true;
    }
    return DiffCsr32{n_, off_rowptr_.p, off_col_.p, off_val_.p, rowsum32_.p};
}

__global__ __launch_bounds__(kBlock) void k_sell_values(int64_t count, const int* __restrict__ map,
                                                        const float* __restrict__ src, float* __restrict__ dst) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e < count) dst[e] = map[e] < 0 ? 0.0f : src[map[e]];
}

DiffSell32 System::sell32(hipStream_t s) {
    const DiffCsr32 d = diff32(s);
    if (!sell_built_) {
        // slice layout of the off-diagonal pattern (once per pattern);
        // padding slots point at the row itself with a = 0
        const int64_t noff = nnz_ - n_;
        std::vector<int> orp(n_ + 1), ocol(noff > 0 ? noff : 0), off, col, map;
        TF_CHECK(hipMemcpyAsync(orp.data(), d.rowptr, sizeof(int) * (n_ + 1), hipMemcpyDeviceToHost, s));
        if (noff > 0)
            TF_CHECK(hipMemcpyAsync(ocol.data(), d.col, sizeof(int) * noff, hipMemcpyDeviceToHost, s));
        TF_CHECK(hipStreamSynchronize(s));
        std::vector<int> soff;
        std::vector<unsigned char> mask;
        sell_sidx_.release();
        sell_lbase_.release();
        sell_pairs_.release();
        sell_npairs_ = 0;
        if (shared_slots_from_env() && noff > 0 && shared_slot_layout(n_, orp, ocol, off, soff, mask, map)) {
            // one column offset per slot of a slice, a mask per row
            sell_soff_.ensure(soff.size());
            sell_mask_.ensure(mask.size());
            TF_CHECK(hipMemcpy(sell_soff_.p, soff.data(), sizeof(int) * soff.size(), hipMemcpyHostToDevice));
            TF_CHECK(hipMemcpy(sell_mask_.p, mask.data(), mask.size(), hipMemcpyHostToDevice));
            sell_col_.release();
            sell_dc_.release();
            // the symmetric form where the pattern allows it; the values are
            // checked at every change (below)
            SymSlots sym;
            if (sym_slots_from_env(true) && !sym32_refused_ &&
                symmetric_slot_layout(n_, orp, ocol, off, soff, mask, map, sym)) {
                off.swap(sym.off);
                map.swap(sym.map);
                sell_sidx_.ensure(sym.sidx.size());
                sell_lbase_.ensure(sym.lbase.size());
                sell_pairs_.ensure(sym.pairs.size());
                TF_CHECK(hipMemcpy(sell_sidx_.p, sym.sidx.data(), sizeof(int) * sym.sidx.size(),
                                   hipMemcpyHostToDevice));
                TF_CHECK(hipMemcpy(sell_lbase_.p, sym.lbase.data(), sizeof(SlotPair) * sym.lbase.size(),
                                   hipMemcpyHostToDevice));
                if (!sym.pairs.empty())
                    TF_CHECK(hipMemcpy(sell_pairs_.p, sym.pairs.data(), sizeof(int) * sym.pairs.size(),
                                       hipMemcpyHostToDevice));
                sell_npairs_ = (int64_t)sym.pairs.size() / 2;
            }
        } else {
            sell_soff_.release();
            sell_mask_.release();
            slice_layout(n_, orp, ocol, true, off, col, map);
            sell_col_.ensure(col.size());
            if (!col.empty())
                TF_CHECK(hipMemcpy(sell_col_.p, col.data(), sizeof(int) * col.size(), hipMemcpyHostToDevice));
            // the 16-bit column form where every slot's column lies within
            // 32767 rows of its row (TFLOWS_FP32_COL16=0 keeps 32-bit columns)
            std::vector<short> dc;
            if (col16_from_env("TFLOWS_FP32_COL16") && slot_offsets16(n_, off, col, true, dc)) {
                sell_dc_.ensure(dc.size());
                TF_CHECK(hipMemcpy(sell_dc_.p, dc.data(), sizeof(short) * dc.size(), hipMemcpyHostToDevice));
            } else {
                sell_dc_.release();
            }
        }
        sell_off_.ensure(off.size());
        sell_map_.ensure(map.size());
        sell_val_.ensure(map.size());
        TF_CHECK(hipMemcpy(sell_off_.p, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice));
        if (!map.empty())
            TF_CHECK(hipMemcpy(sell_map_.p, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice));
        sell_entries_ = (int64_t)map.size();
        sell_built_ = true;
        sell_valid_ = false;
    }
    if (!sell_valid_) {
        if (sell_lbase_.p && !pairs_symmetric(sell_pairs_, sell_npairs_, d.val, symflag_, s)) {
            // values that differ from their mirrors: the full form from now on
            sym32_refused_ = true;
            sell_built_ = false;
            return sell32(s);
        }
        if (sell_entries_ > 0)
            hipLaunchKernelGGL(k_sell_values, dim3((unsigned)((sell_entries_ + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                               s, sell_entries_, (const int*)sell_map_.p, d.val, sell_val_.p);
        TF_CHECK(hipGetLastError());
        sell_valid_ = true;
    }
    DiffSell32 op{n_,           sell_off_.p,  sell_col_.p,  sell_val_.p,   d.rs,
                  sell_dc_.p,   sell_soff_.p, sell_mask_.p, sell_sidx_.p, reinterpret_cast<const int2*>(sell_lbase_.p)};
    op.unroll8 = spmv_unroll_from_env() ? 1 : 0;
    return op;
}

const float* System::diag32(hipStream_t s) {
    diff32(s);
    return diag32_.p;
}

double System::true_residual(const double* x, const double* b, double* r, hipStream_t s) {
    const int nb = (n_ + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_copy, dim3(nb), dim3(kBlock), 0, s, n_, b, r);
    launch_resid(csr(), x, r, grid_, s, ctx());
    hipLaunchKernelGGL((k_norm2<double>), dim3(grid_), dim3(kBlock), 0, s, n_, (const double*)r, partials_.p);
    TF_CHECK(hipGetLastError());
    return std::sqrt(host_site_sum(partials_.p, grid_, s));
}

SolveResult System::solve_mixed(int method, int prec, double* x, const double* b, double* r_out, int cap,
                                double tol, double inner_rtol, int max_outer, int* outer_steps, hipStream_t s) {
    auto run = [&] {
        return solve_mixed_ext(method, prec, x, b, r_out, cap, tol, inner_rtol, max_outer, outer_steps, s);
    };
    if (prec != 2) return run();
    return dic_retry(&x, 1, s, run);
}

SolveResult System::solve_mixed_ext(int method, int prec, double* x, const double* b, double* r_out, int cap,
                                    double tol, double inner_rtol, int max_outer, int* outer_steps,
                                    hipStream_t s) {
    if (!reordered_) return solve_mixed_impl(method, prec, x, b, r_out, cap, tol, inner_rtol, max_outer, outer_steps, s);
    double *xi = ext_buffer(0), *bi = ext_buffer(1), *ri = ext_buffer(2);
    to_internal((const double*)x, xi, s);
    to_internal(b, bi, s);
    const SolveResult res = solve_mixed_impl(method, prec, xi, bi, ri, cap, tol, inner_rtol, max_outer, outer_steps, s);
    to_external((const double*)xi, x, s);
    if (r_out) to_external((const double*)ri, r_out, s);
    TF_CHECK(hipStreamSynchronize(s));
    return res;
}

SolveResult System::solve_mixed_impl(int method, int prec, double* x, const double* b, double* r_out, int cap,
                                     double tol, double inner_rtol, int max_outer, int* outer_steps,
                                     hipStream_t s) {
    if (method < 0 || method > 2) throw std::invalid_argument("method must be 0 (cg), 1 (bicg) or 2 (cgs)");
    if (!(inner_rtol > 0.0 && inner_rtol < 1.0)) throw std::invalid_argument("inner_rtol must be in (0, 1)");
    method = kernel_method(method);
    const size_t n = (size_t)n_;
    ws32_.ensure((V_COUNT + 2) * n);
    partials32_.ensure(Sites<float>::elems(grid_));
    state32_.ensure(1);
    r64_.ensure(n);
    if (!mirror32_)
        TF_CHECK(hipHostMalloc((void**)&mirror32_, sizeof(HostMirror), hipHostMallocMapped | hipHostMallocCoherent));
    ws_.ensure((size_t)V_COUNT * n);

    // the DIC sweeps of an FP32 inner solve read the plainly rounded
    // off-diagonal values; the operator of its SpMVs is the difference form
    const Csr<float> A32 = prec == 2 ? csr32(s) : Csr<float>{n_, rowptr_.p, col_.p, dia_.p, nullptr};
    KrylovVectors<float> w;
    float* base = ws32_.p;
    auto vec = [&](int k) { return base + (size_t)k * n; };
    w.D = vec(V_D); w.p = vec(V_P); w.q = vec(V_Q); w.r2 = vec(V_R2); w.p2 = vec(V_P2); w.q2 = vec(V_Q2);
    w.u1 = vec(V_U1); w.u2 = vec(V_U2); w.v2 = vec(V_V2);
    w.x = vec(V_COUNT);
    w.r1 = vec(V_COUNT + 1);
    TF_CHECK(hipMemsetAsync(ws32_.p, 0, sizeof(float) * (V_COUNT + 2) * n, s));
    Sites<float> S32 = Sites<float>::carve(partials32_.p, grid_);
    static const Levels kNoLevels;
    const Levels* L = prec == 2 ? &levels() : &kNoLevels;
    const int nb = (n_ + kBlock - 1) / kBlock;
    // the FP32 preconditioner: fp32(diag) for Jacobi, 1 for none, and the
    // DIC pivots computed in FP64 (Prec1.f90:51-59) then rounded
    if (prec == 1) {
        TF_CHECK(hipMemcpyAsync(w.D, diag32(s), sizeof(float) * n, hipMemcpyDeviceToDevice, s));
    } else if (prec == 2) {
        double* d64 = ws_.p;                        // FP64 scratch slot
        dic_setup(d64, s);
        prepare_dic(true, s);
        hipLaunchKernelGGL(k_f64_to_f32, dim3(nb), dim3(kBlock), 0, s, n_, (const double*)d64, w.D);
    } else {
        hipLaunchKernelGGL(k_fill_f32, dim3(nb), dim3(kBlock), 0, s, n_, 1.0f, w.D);
    }
    TF_CHECK(hipGetLastError());

    double* r64 = r_out ? r_out : r64_.p;
    double* xprev = ws_.p + (size_t)V_P * n;   // FP64 slot free during a mixed solve
    double err = true_residual(x, b, r64, s);
    SolveResult res;
    res.ini_res = err;
    res.ini_written = true;
    int total = 0, outer = 0;
    double rtol = inner_rtol;
    while (err >= tol && outer < max_outer && total < cap) {
        ++outer;
        hipLaunchKernelGGL(k_scale_to_f32, dim3(nb), dim3(kBlock), 0, s, n_, (const double*)r64, 1.0 / err, w.r1,
                           w.x);
        SolverState h{};
        h.tol = rtol;                  // ||r32|| = 1 up to rounding: relative target
        h.cap = cap - total;
        h.div_limit = kInnerDivLimit;
        h.stall_min = kInnerStallMin;
        h.err_min[0] = HUGE_VAL;
        TF_CHECK(hipMemcpyAsync(state32_.p, &h, sizeof(h), hipMemcpyHostToDevice, s));
        with_op32(s, [&](const auto& Aop) {
            run_solve<float, std::decay_t<decltype(Aop)>>(ctx(), method, prec, cap - total, Aop, A32, w, S32,
                                                          state32_.p, mirror32_, L, s, /*d_ready=*/true,
                                                          /*zero_x=*/true, &replay_);
        });
        SolverState h2;
        TF_CHECK(hipMemcpy(&h2, state32_.p, sizeof(h2), hipMemcpyDeviceToHost));
        // iterations actually executed: niter on convergence, cap on exhaustion
        const int done_its = h2.niter_out > h.cap ? h.cap : h2.niter_out;
        total += done_its;
        bool failed = h2.diverged != 0;   // a diverged correction is not applied
        if (!failed) {
            TF_CHECK(hipMemcpyAsync(xprev, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            hipLaunchKernelGGL(k_refine, dim3(nb), dim3(kBlock), 0, s, n_, x, err, (const float*)w.x, b, r64);
            launch_resid(csr(), (const double*)x, r64, grid_, s, ctx());
            hipLaunchKernelGGL((k_norm2<double>), dim3(grid_), dim3(kBlock), 0, s, n_, (const double*)r64,
                               partials_.p);
            TF_CHECK(hipGetLastError());
            const double err_new = std::sqrt(host_site_sum(partials_.p, grid_, s));
            if (!(err_new < err)) {    // a step that does not help is undone
                TF_CHECK(hipMemcpyAsync(x, xprev, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
                err = true_residual(x, b, r64, s);
                failed = true;
            } else {
                err = err_new;
            }
        }
        if (failed) {
            if (rtol < kLoosestInnerRtol) {
                rtol = std::min(kLoosestInnerRtol, rtol * 10.0);
                continue;
            }
            break;
        }
        if (done_its == 0) break;      // the inner solve could not make progress
    }
    if (outer_steps) *outer_steps = outer;
    if (prec == 2) check_dic(s);
    res.niter = err < tol ? total : total + 1;
    res.fin_res = err;
    return res;
}

// Mixed precision with reliable updates (residual replacement): one FP32
// CG recurrence runs across the whole solve, and every `interval`
// iterations (and whenever its recurrence residual reaches tol) an FP64
// refinement step merges the FP32 correction into x (FP64), recomputes the
// true residual b - A x in FP64 and restarts the FP32 residual from it.
// The solve stops when the FP64 true residual satisfies ||b - A x|| < tol.
// Unlike restarted refinement the Krylov directions p survive each
// refinement step, so the iteration count stays close to the FP64 CG's.
SolveResult System::solve_mixed_reliable(int prec, double* x, const double* b, double* r_out, int cap, double tol,
                                         int interval, int* outer_steps, hipStream_t s) {
    auto run = [&] {
        return solve_mixed_reliable_ext(prec, x, b, r_out, cap, tol, interval, outer_steps, s);
    };
    if (prec != 2) return run();
    return dic_retry(&x, 1, s, run);
}

SolveResult System::solve_mixed_reliable_ext(int prec, double* x, const double* b, double* r_out, int cap,
                                             double tol, int interval, int* outer_steps, hipStream_t s) {
    if (!reordered_) return solve_mixed_reliable_impl(prec, x, b, r_out, cap, tol, interval, outer_steps, s);
    double *xi = ext_buffer(0), *bi = ext_buffer(1), *ri = ext_buffer(2);
    to_internal((const double*)x, xi, s);
    to_internal(b, bi, s);
    const SolveResult res = solve_mixed_reliable_impl(prec, xi, bi, r_out ? ri : nullptr, cap, tol, interval,
                                                      outer_steps, s);
    to_external((const double*)xi, x, s);
    if (r_out) to_external((const double*)ri, r_out, s);
    TF_CHECK(hipStreamSynchronize(s));
    return res;
}

// The FP32 recurrence of the reliable-update forms pauses for a residual
// replacement once it has dropped this far below the last FP64 norm.
constexpr double kRelDrop = 1e-2;

// (r32, r32) of the replaced FP32 residual -> site B, so that the stopping
// test of the next iteration reads it (the second value, rho of the Jacobi
// forms, is not used by the DIC iteration).
__global__ __launch_bounds__(kBlock) void k_rel_site_b(int n, const float* __restrict__ r, float* __restrict__ partB) {
    float v[2] = {0.0f, 0.0f};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) v[0] += r[i] * r[i];
    site_write<float, 2>(v, partB);
}

// The reliable-update form with DIC (prec = 2).  The FP32 recurrence is
// the reference's preconditioned CG (CG.f90:113-259) with the DIC sweeps in
// FP32 (pivots computed in FP64, Prec1.f90:51-59, then rounded), run in
// chunks of at most `interval` iterations; after each chunk x64 +=
// fp64(x32), the FP64 true residual b - A x64 is formed and, unless it
// meets tol, replaces the FP32 residual while the search direction p
// carries on.  A chunk also ends once the FP32 recurrence norm has dropped
// below max(tol, 1e-2 * the last FP64 norm), as in the Jacobi form: the
// replacement then comes while the recurrence still tracks the true
// residual, so the next beta = rho_new / rho_old stays sound (replacing a
// recurrence that ran far below its true residual stalls the solve).  The
// solve stops when the FP64 true residual meets tol, or after `cap` FP32
// iterations (niter = cap + 1).
SolveResult System::solve_mixed_reliable_dic(double* x, const double* b, double* r_out, int cap, double tol,
                                             int interval, int* outer_steps, hipStream_t s) {
    const size_t n = (size_t)n_;
    ws32_.ensure((V_COUNT + 2) * n);
    partials32_.ensure(Sites<float>::elems(grid_));
    state32_.ensure(1);
    r64_.ensure(n);
    if (!mirror32_)
        TF_CHECK(hipHostMalloc((void**)&mirror32_, sizeof(HostMirror), hipHostMallocMapped | hipHostMallocCoherent));
    ws_.ensure((size_t)V_COUNT * n);
    const Csr<float> A32 = csr32(s);
    KrylovVectors<float> w;
    float* base = ws32_.p;
    auto vec = [&](int k) { return base + (size_t)k * n; };
    w.D = vec(V_D); w.p = vec(V_P); w.q = vec(V_Q); w.r2 = vec(V_R2); w.p2 = vec(V_P2); w.q2 = vec(V_Q2);
    w.u1 = vec(V_U1); w.u2 = vec(V_U2); w.v2 = vec(V_V2);
    w.x = vec(V_COUNT);
    w.r1 = vec(V_COUNT + 1);
    TF_CHECK(hipMemsetAsync(ws32_.p, 0, sizeof(float) * (V_COUNT + 2) * n, s));
    const Sites<float> S32 = Sites<float>::carve(partials32_.p, grid_);
    const Levels* L = &levels();
    const int nb = (n_ + kBlock - 1) / kBlock;
    double* d64 = ws_.p;                            // FP64 scratch slot
    dic_setup(d64, s);
    prepare_dic(true, s);
    hipLaunchKernelGGL(k_f64_to_f32, dim3(nb), dim3(kBlock), 0, s, n_, (const double*)d64, w.D);

    double* r64 = r_out ? r_out : r64_.p;
    double err = true_residual(x, b, r64, s);
    SolveResult res;
    res.ini_res = err;
    res.ini_written = true;
    int replacements = 0, k_next = 1, ran = 0;
    if (err >= tol && cap > 0) {
        hipLaunchKernelGGL(k_scale_to_f32, dim3(nb), dim3(kBlock), 0, s, n_, (const double*)r64, 1.0, w.r1, w.x);
        SolverState h{};
        h.tol = tol;
        h.cap = cap;
        TF_CHECK(hipMemcpyAsync(state32_.p, &h, sizeof(h), hipMemcpyHostToDevice, s));
        with_op32(s, [&](const auto& Aop) {
            using Op = std::decay_t<decltype(Aop)>;
            start_solve<float, Op>(ctx(), 0, 2, Aop, A32, w, S32, state32_.p, mirror32_, L, s, true, true);
            for (;;) {
                // this chunk's pause target for the FP32 recurrence
                const double target = std::max(tol, kRelDrop * err);
                // written on s, behind the chunk before it (a copy on the
                // null stream is not ordered after a non-blocking stream)
                launch_set_double(&state32_.p->tol, target, s);
                for (int i = 0; i < interval && k_next + i <= cap + 1; ++i)
                    enqueue_iteration<float, Op>(ctx(), 0, false, k_next + i, Aop, A32, w, S32, state32_.p,
                                                 mirror32_, L, s);
                const int k_end = std::min(k_next + interval - 1, cap + 1);
                SolverState h2;
                TF_CHECK(hipMemcpyAsync(&h2, state32_.p, sizeof(h2), hipMemcpyDeviceToHost, s));
                TF_CHECK(hipStreamSynchronize(s));
                // iterations that ran: up to the stopping test that fired, or the chunk
                const int last = h2.done ? h2.niter_out : k_end;
                ran = std::min(last, cap);
                hipLaunchKernelGGL(k_rel_merge, dim3(nb), dim3(kBlock), 0, s, n_, x, w.x);
                err = true_residual(x, b, r64, s);
                ++replacements;
                if (err < tol || ran >= cap) break;
                // the FP32 residual restarts from the FP64 one; the next
                // iteration's stopping test reads its norm
                hipLaunchKernelGGL(k_scale_to_f32, dim3(nb), dim3(kBlock), 0, s, n_, (const double*)r64, 1.0, w.r1,
                                   w.x);
                hipLaunchKernelGGL(k_rel_site_b, dim3(grid_), dim3(kBlock), 0, s, n_, (const float*)w.r1, S32.B);
                // a pause that ran no iteration cannot be resumed: the FP32
                // norm of the replaced residual is below tol, its FP64 norm
                // is not (they differ by FP32 rounding); report as it stands
                if (h2.done && last < k_next) break;
                if (h2.done) launch_set_int(&state32_.p->done, 0, s);
                k_next = last + 1;
            }
        });
        check_dic(s);
    }
    TF_CHECK(hipGetLastError());
    if (outer_steps) *outer_steps = replacements;
    res.niter = err < tol ? ran : ran + 1;
    res.fin_res = err;
    return res;
}

SolveResult System::solve_mixed_reliable_impl(int prec, double* x, const double* b, double* r_out, int cap,
                                              double tol, int interval, int* outer_steps, hipStream_t s) {
    if (interval < 1) interval = 1;
    if (prec == 2) return solve_mixed_reliable_dic(x, b, r_out, cap, tol, interval, outer_steps, s);
    const size_t n = (size_t)n_;
    ws32_.ensure((V_COUNT + 2) * n);
    partials32_.ensure(Sites<float>::elems(grid_));
    state32_.ensure(1);
    if (!mirror32_)
        TF_CHECK(hipHostMalloc((void**)&mirror32_, sizeof(HostMirror), hipHostMallocMapped | hipHostMallocCoherent));
    float* base = ws32_.p;
    auto vec = [&](int k) { return base + (size_t)k * n; };
    float *D32 = vec(V_D), *p0 = vec(V_P), *p1 = vec(V_P2), *q = vec(V_Q), *z = vec(V_U1);
    float *x32 = vec(V_COUNT), *r32 = vec(V_COUNT + 1);
    const int nb = (n_ + kBlock - 1) / kBlock;
    if (prec == 1) TF_CHECK(hipMemcpyAsync(D32, diag32(s), sizeof(float) * n, hipMemcpyDeviceToDevice, s));
    else hipLaunchKernelGGL(k_fill_f32, dim3(nb), dim3(kBlock), 0, s, n_, 1.0f, D32);
    hipLaunchKernelGGL(k_fill_f32, dim3(nb), dim3(kBlock), 0, s, n_, 0.0f, x32);
    Sites<float> S32 = Sites<float>::carve(partials32_.p, grid_);
    double* partE = partials_.p + Sites<double>::elems(grid_);
    SolverState h{};
    h.tol = tol;
    h.cap = cap;
    TF_CHECK(hipMemcpyAsync(state32_.p, &h, sizeof(h), hipMemcpyHostToDevice, s));
    mirror32_->done = 0;
    mirror32_->iter = 0;
    // the FP64 residual on the operator of the FP64 iterations (sliced or
    // CSR: the same bits)
    const bool carry = cg_form() != 2 && cg_carry();
    auto replace = [&](double* rout) {
        if (carry)
            hipLaunchKernelGGL(k_rel_merge_carry, dim3(nb), dim3(kBlock), 0, s, n_, x, x32, (const float*)p0,
                               (const SolverState*)state32_.p);
        else
            hipLaunchKernelGGL(k_rel_merge, dim3(nb), dim3(kBlock), 0, s, n_, x, x32);
        with_op64(s, [&](const auto& A) {
            hipLaunchKernelGGL((k_rel_resid<std::decay_t<decltype(A)>>), dim3(grid_), dim3(kBlock), 0, s, A,
                               (const double*)x, b, r32, z, D32, rout, S32.B, partE, state32_.p);
        });
    };
    replace(nullptr);
    TF_CHECK(hipGetLastError());

    hipEvent_t ev[2];
    TF_CHECK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    TF_CHECK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    bool pending[2] = {false, false};
    int slot = 0, issued = 0, replacements = 1;
    bool done = false;
    const long long guard = 2LL * (cap > 0 ? cap : 0) + 2LL * interval;
    auto iterations = [&](const auto& Aop) {
        for (int i = 0; i < interval; ++i) {
            if (cg_form() == 2) {
                launch_cgf_spmv(true, Aop, (const float*)z, p0, p1, q, S32, (const double*)partE, state32_.p,
                                mirror32_, s);
                launch_cgf_update(n_, x32, r32, z, (const float*)p0, (const float*)p1, (const float*)q,
                                  (const float*)D32, S32, state32_.p, s);
            } else {
                launch_cgj(true, ctx().strided != 0, Aop, n_, x32, r32, z, p0, q, (const float*)D32, S32,
                           (const double*)partE, state32_.p, mirror32_, s, carry);
            }
        }
    };
    while (!done && issued <= guard) {
        with_op32(s, iterations);
        issued += interval;
        replace(nullptr);
        ++replacements;
        launch_cgf_check(true, S32, (const double*)partE, state32_.p, mirror32_, s);
        TF_CHECK(hipGetLastError());
        TF_CHECK(hipEventRecord(ev[slot], s));
        pending[slot] = true;
        const int other = slot ^ 1;
        if (pending[other]) {
            TF_CHECK(hipEventSynchronize(ev[other]));
            pending[other] = false;
            done = mirror32_->done != 0;
        }
        slot = other;
    }
    // final merge and FP64 residual (also the caller's r_out)
    replace(r_out);
    TF_CHECK(hipGetLastError());
    SolverState h2;
    TF_CHECK(hipMemcpyAsync(&h2, state32_.p, sizeof(h2), hipMemcpyDeviceToHost, s));
    const double fin = std::sqrt(host_site_sum(partE, grid_, s));
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    SolveResult res;
    res.ini_res = h2.ini_res;
    res.ini_written = h2.ini_written != 0;
    res.fin_res = fin;
    res.niter = fin < tol ? h2.it : h2.it + 1;
    if (outer_steps) *outer_steps = replacements;
    return res;
}

}  // namespace tflows
