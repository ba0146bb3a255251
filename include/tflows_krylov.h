/*
 * tflows_krylov.h -- C ABI of the MI355X (gfx950) Krylov solver library
 * libtflows_krylov.so.
 *
 * Two groups of entry points:
 *
 * 1. Drop-in replacements of the reference's external Fortran subroutines.
 *    They keep the Fortran calling convention exactly (every argument by
 *    reference, lowercase name with a trailing underscore, no hidden
 *    arguments), take HOST arrays in the reference layout and return when
 *    the results are back in host memory, so a Fortran host links them in
 *    place of the reference objects and its call sites compile unchanged.
 *    Array bases follow the Fortran declarations: x is x(-NB:N) and Ab is
 *    Ab(-NB:-1), so x(i) is x[NB + i]; A, Acol, Arow, Adia, r1 are 1-based
 *    and the stored indices are 1-based.  The single-domain build is
 *    assumed (NPro = 0): Ab and x(-NB:-1) are never read or written.
 *
 * 2. Device-resident entry points (tflows_*): the matrix pattern and values
 *    and the x / r vectors live in HBM, the caller passes device pointers
 *    and a HIP stream; used by batching, mixed precision, the adjoint and
 *    the benchmarks.  They return 0 on success and a negative code on
 *    failure; tflows_last_error() gives the message.
 *
 * All FP arithmetic of the drop-ins is IEEE FP64; indices are int32.
 */
#ifndef TFLOWS_KRYLOV_H
#define TFLOWS_KRYLOV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * 1. Drop-in Fortran-ABI entry points
 * ---------------------------------------------------------------------- */

/* Preconditioned CG.  Replaces SUBROUTINE cg, Process/CG.f90:2-3.
 * r1 enters as the right-hand side b and leaves as the final residual;
 * x enters as the initial guess and leaves as the solution; niter enters as
 * the cap and leaves as the iteration count (cap+1 if not converged);
 * the stopping test is ||r||_2 < tol (absolute).  prec: 0 none, 1 diagonal,
 * 2 diagonal incomplete Cholesky.  On the early exit ||b|| < tol, IniRes is
 * not written and FinRes receives the residual left by the previous solve,
 * as the reference's module variable `error` does. */
void cg_(const int32_t* N, const int32_t* NB, const int32_t* NONZ, const double* A, const int32_t* Acol,
         const int32_t* Arow, const int32_t* Adia, const double* Ab, double* x, double* r1, const int32_t* prec,
         int32_t* niter, const double* tol, double* IniRes, double* FinRes);

/* "BiCG" with the reference's semantics: the shadow system is multiplied by
 * A (not A^T) and starts from r2 = r1.  Replaces SUBROUTINE bicg,
 * Process/BiCG.f90:2-3.  Arguments as cg_. */
void bicg_(const int32_t* N, const int32_t* NB, const int32_t* NONZ, const double* A, const int32_t* Acol,
           const int32_t* Arow, const int32_t* Adia, const double* Ab, double* x, double* r1,
           const int32_t* prec, int32_t* niter, const double* tol, double* IniRes, double* FinRes);

/* Preconditioned CGS with the fixed shadow r~ = r0.  Replaces SUBROUTINE
 * CGS, Process/CGS.f90:2-3 (Fortran names are case-insensitive: cgs_).
 * Arguments as cg_. */
void cgs_(const int32_t* N, const int32_t* NB, const int32_t* NONZ, const double* A, const int32_t* Acol,
          const int32_t* Arow, const int32_t* Adia, const double* Ab, double* x, double* r1, const int32_t* prec,
          int32_t* niter, const double* tol, double* IniRes, double* FinRes);

/* Preconditioner set-up D(1:N).  Replaces SUBROUTINE Prec1,
 * Process/Prec1.f90:2: prec 1 -> D = diag(A); 2 -> DIC pivots
 * D(i) = 1/(a_ii - sum_{k<i} D(k) a_ik^2); otherwise D = 1. */
void prec1_(const int32_t* N, const int32_t* NONZ, const double* A, const int32_t* Acol, const int32_t* Arow,
            const int32_t* Adia, double* D, const int32_t* prec);

/* r1 <- r1 - A x.  Replaces SUBROUTINE Resid, Process/Resid.f90:2. */
void resid_(const int32_t* N, const int32_t* NB, const int32_t* NONZ, const double* A, const int32_t* Acol,
            const int32_t* Arow, const double* Ab, const double* x, double* r1);

/* The five drop-ins share one device copy of the last pattern they were
 * given, as the reference fixes its pattern once (TopolM,
 * Process/Processor.f90:100) and calls the solvers on it every outer
 * iteration.  A call whose N, NONZ, Acol, Arow and Adia equal the cached
 * copy's (a host comparison) uploads only A and the vectors, and keeps the
 * pattern's locality order, sliced operators and DIC level analysis; any
 * difference rebuilds the copy.  No counterpart in the reference:
 * tflows_dropin_stats reports the pattern builds and drop-in calls so far
 * (either pointer may be NULL); tflows_dropin_release frees the cached
 * copy (the next drop-in call builds it again). */
int tflows_dropin_stats(int64_t* pattern_builds, int64_t* calls);
int tflows_dropin_release(void);

/* ------------------------------------------------------------------------
 * 2. Device-resident API (no counterpart in the reference)
 * ---------------------------------------------------------------------- */

enum { TFLOWS_CG = 0, TFLOWS_BICG = 1, TFLOWS_CGS = 2 };
enum { TFLOWS_MEM_HOST = 0, TFLOWS_MEM_DEVICE = 1 };

typedef struct tflows_system_s* tflows_system;

/* Message of the last failed call on this thread ("" if none). */
const char* tflows_last_error(void);

/* Library version string. */
const char* tflows_version(void);

/* Create a system for an N x N matrix in the T-FlowS CSR layout
 * (Acol = 1-based row pointer of N+1, Arow = 1-based column index of NONZ,
 * Adia = 1-based diagonal position, NONZ = nnz + 1), from host or device
 * arrays according to `mem`.  The pattern is copied; values are set with
 * tflows_system_set_values.  `stream` is a hipStream_t (NULL = default). */
int tflows_system_create(int32_t n, int32_t nonz, const int32_t* acol, const int32_t* arow, const int32_t* adia,
                         int32_t mem, void* stream, tflows_system* out);
int tflows_system_destroy(tflows_system s);

/* Copy the NONZ values A(1:NONZ) (host or device). */
int tflows_system_set_values(tflows_system s, const double* aval, int32_t mem, void* stream);

/* Solve on the device.  x (N doubles, x(1:N)) and r1 (N doubles, in: b,
 * out: final residual) are device pointers; *niter is cap in / count out;
 * *ini_res is written only when the reference would write IniRes;
 * *fin_res follows the reference's FinRes.  Synchronous with respect to
 * the host on return. */
int tflows_solve(tflows_system s, int32_t method, int32_t prec, double* x, double* r1, int32_t* niter,
                 double tol, double* ini_res, double* fin_res, void* stream);

/* tflows_solve without any allocation: the solver state, the reduction
 * partials and the work vectors (D, p, q, z, and x / r1 in the internal
 * order of a renumbered system) live in one fixed device array of the
 * library (each GPU holds its own copy, about 48 MiB, until the process
 * ends), so
 * the call makes no hipMalloc / hipFree / hipHostMalloc and no allocating
 * library call.  No counterpart in the reference: it is for callers that
 * cannot allocate during a solve -- tight solve loops and harnesses that
 * forbid device allocation.  Same arguments, results and bits as
 * tflows_solve for what it takes: method 0 (CG) or 1 (BiCG, which runs
 * through the CG kernels, as in tflows_solve), prec 0 or 1, N <=
 * TFLOWS_NOALLOC_MAX_ROWS; anything else returns an error.  All calls on
 * one GPU share that GPU's array (each call clears what must start at
 * zero), so calls that overlap on one GPU -- from two host threads or on
 * two streams -- race and corrupt each other's results. */
#define TFLOWS_NOALLOC_MAX_ROWS 1048576
int tflows_solve_noalloc(tflows_system s, int32_t method, int32_t prec, double* x, double* r1, int32_t* niter,
                         double tol, double* ini_res, double* fin_res, void* stream);

/* Mixed precision: the CG / BiCG / CGS loop of `method` in FP32 (SpMV,
 * preconditioner, vector updates and dot products in FP32) inside an FP64
 * iterative-refinement loop.  x (device, N doubles) is the initial guess in
 * and the solution out; b (device) is not modified; r_out (device, may be
 * NULL) receives the final true residual b - A x.  Stops when the FP64 true
 * residual satisfies ||b - A x||_2 < tol.  Each inner solve runs to a
 * relative reduction inner_rtol of its (unit-scaled) right-hand side;
 * *niter is the cap on the total inner iterations in and their count out
 * (count + 1 if not converged); at most max_outer refinement steps;
 * *outer_steps (may be NULL) receives the number taken.  IniRes / FinRes
 * are the FP64 true-residual norms before and after. */
int tflows_solve_mixed(tflows_system s, int32_t method, int32_t prec, double* x, const double* b, double* r_out,
                       int32_t* niter, double tol, double inner_rtol, int32_t max_outer, double* ini_res,
                       double* fin_res, int32_t* outer_steps, void* stream);

/* Mixed precision, reliable-update form (CG, prec 0, 1 or 2): a single FP32
 * CG recurrence (SpMV in difference form, preconditioner -- DIC sweeps in
 * FP32 for prec 2 -- vector updates and dot products in FP32) with an FP64
 * refinement step every `interval` iterations and whenever the FP32
 * recurrence residual reaches tol or falls 100x below the last FP64 norm: the
 * FP32 correction is merged into x in FP64, the true residual b - A x is
 * recomputed in FP64 and replaces the FP32 residual; the search direction
 * is kept.  Stops when the FP64 true residual satisfies ||b - A x|| < tol,
 * checked at the refinement steps.  Arguments as tflows_solve_mixed;
 * *niter is the cap on FP32 iterations in and the count out (+1 if not
 * converged); *replacements (may be NULL) receives the number of FP64
 * refinement steps. */
int tflows_solve_mixed_reliable(tflows_system s, int32_t prec, double* x, const double* b, double* r_out,
                                int32_t* niter, double tol, int32_t interval, double* ini_res, double* fin_res,
                                int32_t* replacements, void* stream);

/* Adjoint (transposed) solve A^T lambda = g with the Krylov method and
 * preconditioner of the forward solve, for a structurally symmetric pattern
 * (the values of A^T are gathered from the mirror positions; when every
 * value equals its mirror, as for the pressure-correction matrix, the solve
 * runs on the system itself).  lambda (device,
 * N doubles): initial guess in, solution out; r1 (device): g in, final
 * residual out; other arguments as tflows_solve.  No counterpart in the
 * reference. */
int tflows_solve_adjoint(tflows_system s, int32_t method, int32_t prec, double* lambda, double* r1, int32_t* niter,
                         double tol, double* ini_res, double* fin_res, void* stream);

/* Sparse outer-product gradient on the pattern: grad[p] = -lambda_i x_j for
 * every stored entry p = (i, j), i.e. dL/dA given lambda = A^-T dL/dx;
 * grad (device) has NONZ doubles, the last one set to 0. */
int tflows_outer_grad(tflows_system s, const double* lambda, const double* x, double* grad, void* stream);

/* Values of A^T on the same pattern (device, nnz doubles). */
int tflows_transpose_values(tflows_system s, double* aval_t, void* stream);

/* y = A x on the device, FP64 (x, y: N doubles). */
int tflows_spmv(tflows_system s, const double* x, double* y, void* stream);
/* y = A x with the FP32 copy of the values (x, y: N floats). */
int tflows_spmv32(tflows_system s, const float* x, float* y, void* stream);
/* Preconditioner set-up on the device: D (N doubles, device). */
int tflows_prec1(tflows_system s, int32_t prec, double* d, void* stream);
/* r <- r - A x on the device (Resid form, FP64). */
int tflows_resid(tflows_system s, const double* x, double* r, void* stream);

/* Number of DIC levels of the lower-triangular dependency graph
 * (0 before the first prec = 2 use); for reporting. */
int32_t tflows_system_levels(tflows_system s);

/* Form of the system's DIC sweeps (builds the level analysis): 3 = one
 * launch per triangle (and one for the pivots) in which each wavefront
 * sweeps compact tiles of up to 8 levels with the tile's values in LDS,
 * waiting only on values of other tiles (the default where the pattern's
 * rows are sorted around the diagonal, as TopolM makes them, with at most 4
 * entries per triangle); 2 = one launch per triangle with every row
 * waiting on its dependencies (the default elsewhere); 1 = one launch per
 * level over a level-major copy; 0 = one launch per level over the matrix
 * rows; -1 on error.  TFLOWS_DIC_SWEEP picks a form when the levels are
 * built.  All four give the same bits.  A one-launch sweep whose waiting
 * row gives up (it waited beyond TFLOWS_DIC_WAIT_MS, e.g. while other work
 * held the CUs) makes the solve restore its inputs and run again with one
 * launch per level (form 1 after form 2, form 0 after form 3), which the
 * system then keeps.  No counterpart in the reference. */
int32_t tflows_system_dic_sweep(tflows_system s);

/* SpMV form of the system's solves and products: lanes per matrix row
 * (1 = one row per thread with LDS-staged products, 64 rows per wavefront;
 * 2, 4, 8 = lane groups streaming the row, 32 / 16 / 8 rows per
 * wavefront).  Every form sums each row in column order (the reference's
 * order, CG.f90:193-199), so results are identical.  No counterpart in the
 * reference. */
int tflows_system_set_lanes(tflows_system s, int32_t lanes);

/* 1 when the system stores its matrix in a locality (reverse
 * Cuthill-McKee) order internally, 0 when in the caller's order, -1 on
 * error.  The order is chosen at creation (environment TFLOWS_REORDER:
 * 0 never, 1 always, unset: when it shortens the mean column offset at
 * least 4x).  Entries keep their order within each row, so every row sum,
 * and every result of tflows_spmv / tflows_resid / tflows_prec1, is the
 * same bit for bit; vectors are always exchanged in the caller's order.
 * No counterpart in the reference. */
int32_t tflows_system_reordered(tflows_system s);

/* Layout of the FP32 inner operator of the mixed-precision solves (the
 * difference form of A: off-diagonal a_ij and FP64 row sums rounded to
 * FP32).  *sliced = 1: sliced ELL, 64-row slices stored slot-major and
 * padded to the slice's longest row; 2: the same with each column held as
 * a 16-bit offset from its row (chosen where every offset fits, unless
 * TFLOWS_FP32_COL16=0); 3: the same slices with the column offsets shared
 * by each slice -- one ascending list of at most 8 offsets per slice and a
 * slot mask per row, no column stored per entry (chosen where each row's
 * columns ascend and the slices' offset lists stay short, as on a
 * naturally numbered structured mesh, unless TFLOWS_SHARED_SLOTS=0); 4:
 * layout 3 storing only the slots of non-negative offsets, each lower
 * entry read from its stored mirror (the FP32 operator with
 * TFLOWS_SYM_SLOTS32=1, the FP64 one unless TFLOWS_SYM_SLOTS=0, where the
 * pattern allows it and every value so far has equalled its mirror); 0:
 * CSR (environment TFLOWS_FP32_OP=csr when the system was created, or
 * slower by the first use's timing).  *stored_entries: off-diagonal
 * entries held, padding included.  Every layout gives the same bits.
 * Builds the operator if no solve has yet.  No counterpart in the
 * reference. */
int tflows_system_fp32_layout(tflows_system s, int32_t* sliced, int64_t* stored_entries, void* stream);

/* Layout of the FP64 operator of the solver iterations (A itself, the
 * diagonal where it stands in each row): *sliced = 1 sliced ELL, 2 the
 * same with 16-bit column offsets (where every offset fits, unless
 * TFLOWS_FP64_COL16=0), 3 the same with column offsets shared by each
 * slice, 4 layout 3 with the lower entries read from their mirrors (both
 * as for the FP32 operator), 0 the CSR arrays; *stored_entries:
 * entries held, padding included.  No counterpart in the reference. */
int tflows_system_fp64_layout(tflows_system s, int32_t* sliced, int64_t* stored_entries, void* stream);

/* ------------------------------------------------------------------------
 * Batches of independent systems (no counterpart in the reference, whose
 * callers solve one system per call: NewUVW.f90:506 once per velocity
 * component, CalcZmix.f90:394 per scalar).  K systems of any sizes and
 * patterns are solved together: every kernel of an iteration runs once for
 * all of them, each system follows the reference loop on its own matrix
 * with its own stopping test, cap and result, and the iterations are
 * replayed from captured HIP graphs.  Vectors are device arrays holding the
 * K systems' rows back to back (system k starts at row_offsets[k]).
 * ------------------------------------------------------------------------ */
typedef struct tflows_batch_s* tflows_batch;

/* K systems, each in the T-FlowS layout (host arrays acol[k], arow[k],
 * adia[k] of system k, sizes n[k], nonz[k]). */
int tflows_batch_create(int32_t k, const int32_t* n, const int32_t* nonz, const int32_t* const* acol,
                        const int32_t* const* arow, const int32_t* const* adia, tflows_batch* out);
int tflows_batch_destroy(tflows_batch b);
/* Values A(1:NONZ-1) of system `sys`, or of all systems back to back
 * (sys = -1); host or device memory (mem). */
int tflows_batch_set_values(tflows_batch b, int32_t sys, const double* aval, int32_t mem, void* stream);
/* Total rows and entries, and the K+1 row offsets (any output may be NULL). */
int tflows_batch_layout(tflows_batch b, int32_t* rows, int64_t* entries, int32_t* row_offsets);
/* Launch configuration: blocks per launch, lanes per row, iterations per
 * graph replay (0 = chosen by the autotuner / default), and whether blocks
 * of finished systems are handed to running ones (rebalance != 0). */
int tflows_batch_configure(tflows_batch b, int32_t grid, int32_t lanes, int32_t chunk, int32_t rebalance);
/* Configuration in use for (bits = 64 or 32, method, prec) and the number
 * of rebalances so far. */
int tflows_batch_config(tflows_batch b, int32_t bits, int32_t method, int32_t prec, int32_t* grid, int32_t* lanes,
                        int32_t* chunk, int32_t* rebalances);
/* Launch form of the iterations.  1: one launch per kernel of an
 * iteration, every system in each launch, chunks of iterations replayed
 * from a captured graph.  2: one persistent launch per chunk of
 * iterations, also replayed from a captured graph, in which the kernels of
 * an iteration become phases separated by barriers over each system's own
 * blocks, so every system advances at its own pace; its grid is the blocks
 * the device holds at once.  Form 2 exists for CG and BiCG (which runs
 * through the CG kernels) with prec 0 or 1, in FP64 and for the FP32
 * inner solves; other solves run form 1.  Both forms give the same bits.
 * 0 (the default) lets the autotuner time both on the caller's systems
 * and keep the faster.  A persistent launch whose barrier wait exceeds
 * TFLOWS_BARRIER_WAIT_MS (default 2000 ms; other work holding the CUs) is
 * abandoned and the solve runs again in form 1, which the batch keeps. */
int tflows_batch_set_form(tflows_batch b, int32_t form);
/* The form (1 or 2) in use for (bits, method, prec); -1 on error. */
int32_t tflows_batch_form(tflows_batch b, int32_t bits, int32_t method, int32_t prec);
/* How the systems share the iterations (form 1, prec 0 or 1).  1: every
 * system from the first iteration.  2: consecutive systems packed into
 * groups whose operators and vectors fit 192 MB (3/4 of the 256 MB MALL),
 * each group on the whole grid, one group after another; a system's
 * iterations count from its group's start.  0 (default): the autotuner times
 * both on the caller's systems and keeps the faster.  Bits may differ
 * between the two, as between grids.  No counterpart in the reference. */
int tflows_batch_set_groups(tflows_batch b, int32_t mode);
/* The number of groups the solves of (bits, method, prec) run in (1: all
 * at once); -1 on error. */
int32_t tflows_batch_groups(tflows_batch b, int32_t bits, int32_t method, int32_t prec);
/* FP64 solve of all K systems with `method` / `prec` (as tflows_solve):
 * x, r1 device arrays (r1 in: b, out: final residual); cap[k], tol[k] per
 * system (host); niter[k], ini_res[k] (written unless the system took the
 * ||b|| < tol exit), fin_res[k] (in: the value that exit returns) out. */
int tflows_batch_solve(tflows_batch b, int32_t method, int32_t prec, double* x, double* r1, const int32_t* cap,
                       const double* tol, int32_t* niter, double* ini_res, double* fin_res, void* stream);
/* FP32 inner solves inside FP64 iterative refinement, per system (as
 * tflows_solve_mixed); rhs is not modified; r_out may be NULL. */
int tflows_batch_solve_mixed(tflows_batch b, int32_t method, int32_t prec, double* x, const double* rhs,
                             double* r_out, const int32_t* cap, const double* tol, double inner_rtol,
                             int32_t max_outer, int32_t* niter, double* ini_res, double* fin_res,
                             int32_t* outer_steps, void* stream);

/* ------------------------------------------------------------------------
 * Pressure-correction assembly on the device: the step before the pressure
 * solve (Process/CalcPS.f90:51-202), which fills Aval, the mass source b and
 * the face fluxes from the face list, the geometry and the cell fields.
 * The reference keeps these in modules (all_mod, pro_mod); here they are
 * passed with the reference's array bases:
 *   sidec   SideC(0:2,NS), column-major: sidec[3*(s-1) + k] = SideC(k,s)
 *   sidaij  SidAij(1:2,NS): sidaij[2*(s-1) + k-1] (1-based positions in
 *           Aval; only interior faces are read)
 *   adia    Adia(NC)
 *   typebc  TypeBC(-NbC:NC): typebc[NbC + c]   (codes of allp_mod.f90:16-24)
 *   sx..dz, f   (NS)
 *   volume and every cell field: (-NbC:NC), x[NbC + c]
 *   Abou    (-NbC:-1): abou[NbC + c], NbC entries
 * tflows_calcps_create keeps the mesh part on the device (once, after
 * TopolM); tflows_calcps_assemble takes the fields of one outer iteration
 * and writes Aval(NONZ), b(NC), Flux(NS) and Abou like CalcPS.f90:51-202,
 * with errmax (:172-175) and the removed mean bc_sum/vol (:201-204).
 * Diagonals, b before the mean, off-diagonals, Flux, Abou and errmax are
 * the reference's bit for bit; bc_sum is summed as a tree, so the mean and
 * b differ from a sequential sum by its rounding.  With mem =
 * TFLOWS_MEM_DEVICE every array argument is a device pointer, so the
 * values go straight to tflows_system_set_values and the solve; with
 * TFLOWS_MEM_HOST they are copied.  Replaces the face and cell loops of
 * CalcPS; the solve (CalcPS.f90:214-219) stays tflows_solve / cg_. */
typedef struct tflows_calcps_s* tflows_calcps;
int tflows_calcps_create(int32_t nc, int32_t nbc, int32_t ns, int32_t nonz, const int32_t* sidec,
                         const int32_t* sidaij, const int32_t* adia, const int32_t* typebc, const double* sx,
                         const double* sy, const double* sz, const double* dx, const double* dy, const double* dz,
                         const double* f, const double* volume, int32_t mem, void* stream, tflows_calcps* out);
int tflows_calcps_destroy(tflows_calcps h);
/* How the momentum and mixture-fraction cell kernels visit the cells of
 * this mesh: 2 in blocks of 64 cells x 2 lines x 2 planes (a naturally
 * numbered box whose lines are whole multiples of 64 cells, with an even
 * number of lines and of planes), 1 in 256-cell tiles brick by brick (a box
 * whose planes are whole tiles), 0 in 256-cell tiles in order; -1 on
 * error.  TFLOWS_UVW_ORDER=0 or 1 caps it when the handle is created.  All
 * give the same bits.  No counterpart in the reference. */
int32_t tflows_calcps_visit_order(tflows_calcps h);
int tflows_calcps_assemble(tflows_calcps h, int32_t nmat, double dt, const double* asave, const double* rho_n,
                           const double* rho_o, const double* rho_oo, const double* u, const double* v,
                           const double* w, const double* p, const double* px, const double* py, const double* pz,
                           double* aval, double* b, double* flux, double* abou, double* errmax, double* mean,
                           int32_t mem, void* stream);

/* The correction after the pressure solve, Process/CorUVW.f90:36-121 (the
 * CFL / Peclet monitor of :126-147 is not computed): U, V, W (-NbC:NC) are
 * corrected in place with the pressure gradient Px, Py, Pz and copied to
 * PRESSURE boundary cells; Flux(NS) is corrected in place with the pressure
 * correction pp (-NbC:NC) through Aval(SidAij(1,s)) and Abou; b(NC)
 * receives the mass error per unit volume and *errmax its maximum (CorUVW's
 * value, :183).  Bitwise the reference's.  Same bases and mem rule as
 * tflows_calcps_assemble. */
int tflows_calcps_correct(tflows_calcps h, double dt, const double* pp, const double* aval, const double* abou,
                          const double* asave, const double* rho_n, const double* rho_o, const double* rho_oo,
                          const double* px, const double* py, const double* pz, double* u, double* v, double* w,
                          double* flux, double* b, double* errmax, int32_t mem, void* stream);

/* The CFL and Peclet monitor of Process/CorUVW.f90:124-147, per material m
 * (m = 1..Nmat of tflows_calcps_uvw_setup, which must come first: it brings
 * Scoef and the materials): over the interior and BUFFER faces with a cell
 * of material m on either side, cflmax[m-1] = max |dt Flux / (RhoS Scoef
 * (Dx^2 + Dy^2 + Dz^2))| with RhoS = f rho%n(c1) + (1 - f) rho%n(c2), and
 * pemax[m-1] = max |Flux / Scoef / (VISc + TINY)|; zero for a material
 * with no such face.  rho_n (-NbC:NC) and the corrected Flux (NS) as mem
 * says; cflmax and pemax are host arrays of Nmat values, the reference's
 * bit for bit (a maximum does not depend on the order).  Call it after
 * tflows_calcps_correct, as CorUVW does. */
int tflows_calcps_cfl_peclet(tflows_calcps h, double dt, double visc, const double* rho_n, const double* flux,
                             double* cflmax, double* pemax, int32_t mem, void* stream);

/* The momentum assembly of one velocity component, Process/NewUVW.f90:137-499
 * (the reference's `newuvw` up to its BiCG call at :505-507), on the mesh of
 * a tflows_calcps handle.  Covered: CONVEC, DIFFUS (CN or FI), CROSS = FI or
 * not; INERT = LIN or PAR; BLEND = YES, NO or one of the face schemes of
 * Process/ConvScheme.f90 (listed below); BLEND_TEM upwind part; every boundary type of :383-408; Asave (var = 1);
 * under-relaxation with U%URF.  VISeff = VISc (the default of :287-295).
 * Diagonal, off-diagonals, b, Abou, Asave and rhoUi are the reference's bit
 * for bit.  The caller keeps the old-level shift (:154-157) and supplies
 * the source (:180-199); the solve is tflows_solve / bicg_ with rhoUi(1:NC)
 * as the starting point, then tflows_calcps_uvw_update.
 *
 * tflows_calcps_uvw_setup: Scoef(NS), fF(NS) (Process/Calc3.f90:23-70) and
 * material(-NbC:NC) with values 1..nmaterials (at most 8), once per mesh.
 * tflows_calcps_uvw_assemble: flags and constants in *p (codes of
 * Process/IniPar.f90:22-32; per-material arrays at [m - 1]); var selects
 * Si..Dk from the handle's Sx..Dz as Process/Processor.f90:314-340 does; the
 * cell arrays (-NbC:NC) rho_n/o/oo, Ui n/o/oo, source, Pi and the seven
 * gradients in newuvw's argument order (NewUVW.f90:1-41); Flux_u(NS); flux,
 * the mass flux Flux(NS) whose sign picks ConvScheme's upwind cell
 * (Process/ConvScheme.f90:55-65), or NULL for Flux_u.  BLEND may be YES, NO
 * or a ConvScheme scheme (CDS 20 .. GAMMA 27, with PHImax/PHImin of
 * Process/CalMinMax.f90); a face whose materials put YES next to NO is
 * refused.  Writes Aval(NONZ), b(NC), Abou(-NbC:-1), Asave (cells 1..NC of a
 * (-NbC:NC) array, var = 1, may be NULL) and rhoUi(-NbC:NC).
 * tflows_calcps_uvw_update: Ui%n(1:NC) = rhoUi / rho%n (:526). */
typedef struct tflows_uvw_params {
    int32_t var, convec, diffus, cross, inert;
    double visc, dt, urf;
    int32_t blend[8], blend_tem[8];
    double urfc[8];
    int32_t visc_rho;   /* 1: VISeff = visc * RhoS, RhoS = fF rho%n(c1) + (1 - fF) rho%n(c2) (the
                           Moin_Problem() == 4 branch, NewUVW.f90:287-295, CalcZmix.f90:232-242);
                           0: VISeff = visc */
} tflows_uvw_params;
int tflows_calcps_uvw_setup(tflows_calcps h, const double* scoef, const double* fF, const int32_t* material,
                            int32_t nmaterials, int32_t mem, void* stream);
int tflows_calcps_uvw_assemble(tflows_calcps h, const tflows_uvw_params* p, const double* rho_n,
                               const double* rho_o, const double* rho_oo, const double* ui_n, const double* ui_o,
                               const double* ui_oo, const double* source, const double* pi, const double* dUidi,
                               const double* dUidj, const double* dUidk, const double* dUjdi, const double* dUkdi,
                               const double* dUjdj, const double* dUkdk, const double* flux_u, const double* flux,
                               double* aval, double* b, double* abou, double* asave, double* rhoui, int32_t mem,
                               void* stream);
int tflows_calcps_uvw_update(tflows_calcps h, const double* rhoui, const double* rho_n, double* ui_n, int32_t mem,
                             void* stream);

/* The mixture-fraction assembly, Process/CalcZmix.f90:100-387 (the
 * reference's `CalcZmix` up to its CG call at :394), on the same handle
 * after tflows_calcps_uvw_setup.  In *p: blend_tem = BLEND_TEM (it drives
 * the face scheme, the upwind matrix and the upwind part), urfc = URFC_Tem,
 * visc = VIScZmix, urf = Zmix%URF, convec/diffus/cross/inert/dt as for the
 * momentum (var and blend are not read).  Cell arrays (-NbC:NC): rho_n/o/oo,
 * Z n/o/oo, the source and dZ/dx, dZ/dy, dZ/dz; Flux_u(NS); flux as for the
 * momentum.  Writes Aval(NONZ), b(NC), Abou(-NbC:-1) and rhoZ(-NbC:NC), bit
 * for bit the reference's; BLEND_TEM takes the codes BLEND does.
 * tflows_calcps_zmix_update: Zmix%n(1:NC) = rhoZ / rho%n limited to [0, 1]
 * (:401-405). */
int tflows_calcps_zmix_assemble(tflows_calcps h, const tflows_uvw_params* p, const double* rho_n,
                                const double* rho_o, const double* rho_oo, const double* z_n, const double* z_o,
                                const double* z_oo, const double* source, const double* dZdx, const double* dZdy,
                                const double* dZdz, const double* flux_u, const double* flux, double* aval,
                                double* b, double* abou, double* rhoz, int32_t mem, void* stream);
int tflows_calcps_zmix_update(tflows_calcps h, const double* rhoz, const double* rho_n, double* z_n, int32_t mem,
                              void* stream);

/* T-FlowS mesh files (host side; no GPU needed).
 * tflows_mesh_load reads NAME.cns as Process/CnsLoa.f90:38-62 does and, when
 * geo_path is not NULL, NAME.geo as Library/GeoLoa.f90:22-49 does (Fortran
 * sequential unformatted records), then builds the system pattern as
 * Process/TopolM.f90 does and the centre distances D and Scoef as
 * Process/Calc3.f90:9-30 does.
 * tflows_mesh_int_array / tflows_mesh_real_array copy one array out by name;
 * count must equal its length.  Integer arrays: "material", "bcmark",
 * "copyc" (NbC + NC + 1, cell c at [NbC + c]), "sidec" (3 NS, SideC(0:2,NS)
 * column-major), "acol" (NC + 1), "arow" (NONZERO), "adia" (NC), "sidaij"
 * (2 NS, SidAij(1:2,NS) column-major, 0 on boundary faces).  Real arrays:
 * "xc", "yc", "zc", "volume", "delta", "wallds" (NbC + NC + 1; volume, delta,
 * WallDs zero outside 1..NC), "sx", "sy", "sz", "f", "xsp", "ysp", "zsp",
 * "dx_file", "dy_file", "dz_file" (the file's Dx: periodic shifts), "dx",
 * "dy", "dz" (Calc3's centre distances), "scoef" (NS).
 * tflows_mesh_typebc: TypeBC(-NbC:NC) from bcmark and the boundary-type
 * table typBou(1:nb) of the .b file (Process/BouLoa.f90:167-175, 292-294).
 * tflows_mesh_ff: fF(NS) as Calc3.f90:53-71, with StateMat(1:Nmat) of the
 * .b file (FLUID 7, SOLID 8; NULL: every material fluid).
 * The arrays feed tflows_system_create (acol, arow, adia) and
 * tflows_calcps_create / tflows_calcps_uvw_setup directly. */
typedef struct tflows_mesh_s* tflows_mesh;
/* NAME.b, the boundary-condition file, read as Process/BouLoa.f90:30-160 does
 * (comment and token rules of Library/ReadC.f90).  hot: HOT = YES (the
 * materials then carry CONc and CAPc, the conditions a Zmix value).
 * Materials: StateMat (FLUID 7, SOLID 8), DENc, CONc, CAPc (0 unless hot);
 * VISc is one value, the last material's, as the reference keeps it.
 * Boundary conditions 1..NB in file order: typBou (allp_mod codes), a FILE
 * flag with the profile name (tflows_bfile_name kind 0), else U, V, W, P
 * (PRESSURE only) and Zmix (hot only), 0 where absent.  Initial conditions
 * 1..Ninit: a FILE flag and name (kind 1), else U, V, W, Zmix.  TypeBC of
 * the mesh's boundary cells is then typBou(bcmark) (tflows_mesh_typebc).
 * An unknown state or type, a short line or a bad number is an error.
 * Host code, no GPU: the reading part of BouLoa, not its profile files. */
typedef struct tflows_bfile_s* tflows_bfile;
int tflows_bfile_load(const char* path, int32_t hot, tflows_bfile* out);
int tflows_bfile_destroy(tflows_bfile b);
int tflows_bfile_info(tflows_bfile b, int32_t* nmat, int32_t* nb, int32_t* ninit, double* visc);
int tflows_bfile_materials(tflows_bfile b, int32_t* statemat, double* denc, double* conc, double* capc);
int tflows_bfile_bounds(tflows_bfile b, int32_t* typbou, int32_t* from_file, double* u, double* v, double* w,
                        double* p, double* zmix);
int tflows_bfile_inits(tflows_bfile b, int32_t* from_file, double* u, double* v, double* w, double* zmix);
int tflows_bfile_name(tflows_bfile b, int32_t kind, int32_t n, char* buf, int32_t len);

int tflows_mesh_load(const char* cns_path, const char* geo_path, tflows_mesh* out);
int tflows_mesh_destroy(tflows_mesh m);
int tflows_mesh_info(tflows_mesh m, int32_t* nc, int32_t* nbc, int32_t* ns, int32_t* nmat, int32_t* nonz,
                     int32_t* has_geo);
int tflows_mesh_int_array(tflows_mesh m, const char* name, int32_t* dst, int64_t count);
int tflows_mesh_real_array(tflows_mesh m, const char* name, double* dst, int64_t count);
int tflows_mesh_typebc(tflows_mesh m, const int32_t* typbou, int32_t nb, int32_t* typebc, int64_t count);
int tflows_mesh_ff(tflows_mesh m, const int32_t* typebc, const int32_t* statemat, double* ff, int64_t count);

#ifdef __cplusplus
}
#endif

#endif /* TFLOWS_KRYLOV_H */
