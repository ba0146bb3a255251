// system.h -- host-side owner of one matrix on the device and of the
// workspace of its solves.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "krylov_kernels.h"
#include "slot_layout.h"

namespace tflows {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
// A one-launch DIC sweep gave up waiting on its dependencies (the rows it
// left hold NaN); raised by System::check_dic.
struct DicWaitFault : HipError {
    using HipError::HipError;
};

#define TF_CHECK(expr)                                                                            \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            throw ::tflows::HipError(std::string(#expr) + ": " + hipGetErrorString(e_) + " at " + \
                                     __FILE__ + ":" + std::to_string(__LINE__));                  \
    } while (0)

// Device allocation that frees itself.
template <typename T>
struct DevArray {
    T* p = nullptr;
    size_t count = 0;
    DevArray() = default;
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    ~DevArray() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        count = 0;
    }
    void ensure(size_t n) {
        if (n <= count && p) return;
        release();
        TF_CHECK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
        count = n;
    }
};

struct SolveResult {
    int niter = 0;
    double ini_res = 0.0;
    bool ini_written = false;
    double fin_res = 0.0;
};

// Vectors of one Krylov solve (x and r1 belong to the caller; the rest is
// workspace).  Names follow the reference's sol_mod (Library/sol_mod.f90:7-10).
template <typename T>
struct KrylovVectors {
    T *x = nullptr, *r1 = nullptr, *D = nullptr, *p = nullptr, *q = nullptr, *r2 = nullptr, *p2 = nullptr,
      *q2 = nullptr, *u1 = nullptr, *u2 = nullptr, *v2 = nullptr;
};
enum { V_D = 0, V_P, V_Q, V_R2, V_P2, V_Q2, V_U1, V_U2, V_V2, V_COUNT };

// Level sets of the lower (forward) and upper (backward) triangular
// dependency graphs of the pattern, for the DIC sweeps.
struct Levels {
    int nfwd = 0, nbwd = 0;
    std::vector<int> fwd_ptr, bwd_ptr;   // host offsets into the row lists
    DevArray<int> fwd_rows, bwd_rows;    // device row lists, level by level
    std::vector<int> fwd_rows_h, bwd_rows_h;

    // Level-major copy of the triangles for the sweeps (DIC apply): rows in
    // forward-level order, so each forward level is a contiguous range of
    // positions and its loads are coalesced; the lower and upper entries of
    // each row keep their order (the same products summed in the same
    // order as the sweeps over the matrix itself).  The values are gathered
    // from the matrix through the entry maps whenever it changes.
    //
    // mode 1 sweeps one level per launch.  mode 2 sweeps each triangle in
    // one launch of a persistent grid that walks the positions in order,
    // each row waiting for the values of its dependencies, which start as
    // an "empty" NaN pattern that no arithmetic produces (a
    // synchronisation-free triangular solve).  Every wavefront holds rows
    // of one forward level, so no row waits on a row of its own
    // wavefront; the backward sweep runs over the positions in reverse.
    // A row that waits longer than wait_ticks gives up, raises the fault
    // word and leaves NaN; the host then reruns the solve in mode 1
    // (System::dic_retry).
    //
    // mode 3 (tiles, tile_plan.h) sweeps each triangle, and computes the
    // pivots, in one launch of one-wavefront blocks that each walk compact
    // tiles of up to kTilePhases levels with the tile's values in LDS; only
    // dependencies between tiles wait on device-coherent slots.  Its
    // positions are tile by tile, its entries in slot-major ELL form
    // (lcol / lmap / ucol / umap hold tslots entries per position, coded as
    // tile_plan.h describes).  A mode-3 sweep that gives up waiting sends
    // the system to mode 0.
    bool plan = false;
    int mode = 0;                        // 0 matrix rows by level lists, 1 / 2 / 3 as above
    int ntiles = 0, tslots = 0;          // mode 3: tiles and entry slots per position
    int twaves = 0;                      // mode 3: wavefronts of a tiled launch
    DevArray<int> tph;                   // mode 3: phase starts of every tile
    DevArray<int> wstart, wtiles;        // mode 3: each wavefront's tiles (tile_schedule)
    DevArray<int> dmap;                  // mode 3: the diagonal entry of every position
    mutable DevArray<double> dgp;        // mode 3: the diagonal in position order (pivots)
    int npos = 0;                        // positions (rows and padding)
    std::vector<int> pos_ptr;            // host: position range of each forward level
    size_t nlow = 0, nup = 0;            // lower / upper entries
    DevArray<int> fault;                 // raised by a row that gave up waiting
    uint64_t wait_ticks = 0;
    DevArray<int> order;                 // position -> row (-1: padding)
    DevArray<int> lptr, lcol, lmap;      // lower entries: position-based CSR, source entry
    DevArray<int> uptr, ucol, umap;      // upper entries
    DevArray<int> bwd_pos;               // backward levels as positions
    mutable DevArray<double> lval, uval, up, zp, zb, dp;
    mutable DevArray<float> lval32, uval32, up32, zp32, zb32, dp32;
    mutable bool vals64 = false, vals32 = false;
};

// Sliced ELL layout (kernels.h, Sell / DiffSell32) of a CSR pattern with
// row pointer rp and columns cl: slice offsets, the column of every slot
// and the CSR entry it holds (-1 for padding).  Padding slots take the
// row's own index when pad_own, else -1.
// 16-bit column offsets (col - row) of a sliced layout; false when one
// does not fit.  pad_own: padding slots hold the row itself; otherwise
// column -1, which becomes kPad16.
bool slot_offsets16(int n, const std::vector<int>& off, const std::vector<int>& col, bool pad_own,
                    std::vector<short>& dc);
// false when the environment variable `name` is "0"
bool col16_from_env(const char* name);
void slice_layout(int n, const std::vector<int>& rp, const std::vector<int>& cl, bool pad_own, std::vector<int>& off,
                  std::vector<int>& col, std::vector<int>& map);
bool shared_slots_from_env();

// Whether the FP64 (TFLOWS_SYM_SLOTS, default on) or the FP32
// (TFLOWS_SYM_SLOTS32, default off) operator tries the symmetric form of
// the shared-offset layout (slot_layout.h).
bool sym_slots_from_env(bool fp32);
bool spmv_unroll_from_env();
template <typename T>
bool pairs_symmetric(const DevArray<int>& pairs, int64_t npairs, const T* a, DevArray<int>& flag, hipStream_t s);

// TFLOWS_DIC_SWEEP (read when a system first builds its levels): 0 sweeps
// the matrix rows level by level, 1 the level-major copy level by level,
// 2 (default) the level-major copy one launch per triangle.  All three
// give the same bits.  TFLOWS_DIC_WAIT_MS bounds a row's wait in form 2.
int dic_sweep_mode();
uint64_t dic_wait_ms();

// Every vector argument of the public solve / product entry points is in
// the caller's numbering.  The matrix itself may be stored permuted for
// locality (reorder.hip): then the pattern arrays returned by csr() /
// rowptr() / col() / dia() and every workspace vector are in the internal
// order, and the entry points gather their inputs in and scatter their
// outputs back.
// Graph replay of one system's DIC iterations.  A DIC iteration is one
// launch per level per sweep (some 1500 dependent launches at 256^3), so
// the host, not the GPU, limits the rate at which they are issued; one
// iteration is captured into a HIP graph on a stream of the replayer's
// own (the caller's may be the legacy stream, which cannot be captured)
// and replayed with its iteration index read from a device counter.
// Graphs are kept per set of buffers they were captured with.
struct Replay {
    hipStream_t stream = nullptr;
    hipEvent_t ev[3] = {};
    DevArray<int> kbase;
    struct Entry {
        std::string key;
        hipGraphExec_t exec;
    };
    std::vector<Entry> cache;
    Replay() = default;
    Replay(const Replay&) = delete;
    Replay& operator=(const Replay&) = delete;
    ~Replay();
    void ensure();
};
bool dic_graph_enabled();

// The reference's BiCG multiplies its shadow by A and starts it at r
// (BiCG.f90:93-95, 196-200), so its shadow vectors equal the primal ones bit
// for bit at every step (r2 = r1, p2 = p1, q2 = q1; its rho = (z1, r2) is
// CG's (r, z) with the factors of each product swapped, which IEEE
// multiplication does not distinguish).  Its results are CG's results, and
// method 1 runs the CG kernels, which move half the bytes.
// TFLOWS_BICG_SHADOW=1 (read at every solve) runs the two-vector kernels.
int kernel_method(int method);

// Locality (reverse Cuthill-McKee) order of a 0-based pattern, new -> old,
// when the environment (TFLOWS_REORDER) and the pattern call for one (it
// must shorten the mean column offset at least 4x); empty otherwise.
std::vector<int> locality_order(int n, const std::vector<int>& rowptr, const std::vector<int>& col, bool allow);
// dst[i] = src[perm[i]]  /  dst[perm[i]] = src[i]
void launch_gather(int n, const int* perm, const double* src, double* dst, hipStream_t s);
void launch_scatter(int n, const int* perm, const double* src, double* dst, hipStream_t s);

class System {
   public:
    System(int n, int nonz, const int* acol, const int* arow, const int* adia, bool device_ptrs, hipStream_t s,
           bool allow_reorder = true);
    ~System();

    int n() const { return n_; }
    int nnz() const { return nnz_; }

    void set_values(const double* aval, bool device_ptr, hipStream_t s);
    // values of entries [offset, offset + count) only (0-based)
    void set_values_range(int64_t offset, int64_t count, const double* aval, bool device_ptr, hipStream_t s);
    // values in the caller's entry order
    const double* values_original() const { return reordered_ ? val_orig_.p : val_.p; }
    bool reordered() const { return reordered_; }
    // caller numbering <-> internal numbering of an N-vector
    template <typename T>
    void to_internal(const T* ext, T* in, hipStream_t s) const;
    template <typename T>
    void to_external(const T* in, T* ext, hipStream_t s) const;
    const float* values32(hipStream_t s);   // FP32 copy, made on first use

    Csr<double> csr() const { return Csr<double>{n_, rowptr_.p, col_.p, dia_.p, val_.p}; }
    Csr<float> csr32(hipStream_t s) { return Csr<float>{n_, rowptr_.p, col_.p, dia_.p, values32(s)}; }
    // FP32 difference-form operator (off-diagonal CSR + FP64 row sums) and
    // the FP32 diagonal, made on first use after set_values (mixed.hip)
    DiffCsr32 diff32(hipStream_t s);
    // the same operator in sliced ELL layout (built on first use, values
    // refreshed after set_values)
    DiffSell32 sell32(hipStream_t s);
    // Calls f with the FP32 inner operator of this system: the sliced or
    // the CSR form.  TFLOWS_FP32_OP=csr / =sliced (read when the system is
    // created) fixes the form; otherwise the first use times both and keeps
    // the faster (they give the same values).
    template <class F>
    void with_op32(hipStream_t s, F&& f) {
        if (op32_sliced(s)) f(sell32(s));
        else f(diff32(s));
    }
    bool op32_sliced(hipStream_t s);
    // The FP64 operator of the solver iterations: A in sliced ELL layout or
    // the CSR arrays themselves (the same row sums bit for bit).
    // TFLOWS_FP64_OP=csr / =sliced fixes the form; otherwise the first use
    // times both and keeps the faster.
    Sell<double> sell64(hipStream_t s);
    template <class F>
    void with_op64(hipStream_t s, F&& f) {
        if (op64_sliced(s)) f(sell64(s));
        else f(csr());
    }
    bool op64_sliced(hipStream_t s);
    // bytes per column index of the sliced FP32 operator: 2 when it holds
    // 16-bit column offsets, else 4
    int sell32_col_bytes(hipStream_t s) {
        sell32(s);
        return sell_soff_.p ? 0 : (sell_dc_.p ? 2 : 4);
    }
    int64_t sell32_entries(hipStream_t s) {
        sell32(s);
        return sell_entries_;
    }
    // the same for the sliced FP64 operator
    int sell64_col_bytes(hipStream_t s) {
        sell64(s);
        return sell64_soff_.p ? 0 : (sell64_dc_.p ? 2 : 4);
    }
    int64_t sell64_entries(hipStream_t s) {
        sell64(s);
        return sell64_entries_;
    }
    // whether the sliced operators are in the symmetric form
    bool sell64_symmetric(hipStream_t s) {
        sell64(s);
        return sell64_lbase_.p != nullptr;
    }
    bool sell32_symmetric(hipStream_t s) {
        sell32(s);
        return sell_lbase_.p != nullptr;
    }
    const float* diag32(hipStream_t s);

    // Solve with x, r1 device pointers (x(1:N), r1(1:N)); `stale_error` is
    // the FinRes the early exit hands back.
    SolveResult solve(int method, int prec, double* x, double* r1, int cap, double tol, double stale_error,
                      hipStream_t s);

    // FP32 Krylov inside FP64 iterative refinement (mixed.hip).  b is not
    // modified; r_out (may be null) receives the final true residual.
    SolveResult solve_mixed(int method, int prec, double* x, const double* b, double* r_out, int cap, double tol,
                            double inner_rtol, int max_outer, int* outer_steps, hipStream_t s);
    // the same with reliable updates (residual replacement every `interval`
    // FP32 CG iterations); CG with prec 0 / 1
    SolveResult solve_mixed_reliable(int prec, double* x, const double* b, double* r_out, int cap, double tol,
                                     int interval, int* outer_steps, hipStream_t s);
    // r = b - A x (FP64) and its norm; internal numbering
    double true_residual(const double* x, const double* b, double* r, hipStream_t s);

    // adjoint support (adjoint.hip): mirror positions of the structurally
    // symmetric pattern, a system holding A^T on the same pattern (values
    // refreshed after set_values; this system itself when its values equal
    // their mirrors), and grad[p] = -lambda_i x_j
    const int* transpose_map();
    System& transposed(hipStream_t s);
    void outer_grad(const double* lambda, const double* x, double* grad, hipStream_t s);

    void prec1(int prec, double* d, hipStream_t s);
    void spmv(const double* x, double* y, hipStream_t s);
    void spmv32(const float* x, float* y, hipStream_t s);
    void resid(const double* x, double* r, hipStream_t s);

    const Levels& levels();
    int grid() const { return grid_; }
    // SpMV form of the solves (lanes per row: 1, 2, 4 or 8)
    int lanes() const { return lanes_; }
    void set_lanes(int lanes);
    // tile order of the SpMV kernels (0 contiguous per block, 1 strided
    // inside each XCD's chunk)
    void set_tile_order(int strided) { strided_ = strided ? 1 : 0; }
    BatchCtx ctx() const {
        BatchCtx c;
        c.lanes = lanes_;
        c.strided = strided_;
        return c;
    }

    // raw pattern access for other components (batching, adjoint)
    const int* rowptr() const { return rowptr_.p; }
    const int* col() const { return col_.p; }
    const int* dia() const { return dia_.p; }

    Sites<double> sites();
    KrylovVectors<double> vectors(double* x, double* r1);
    SolverState* state() { return state_.p; }
    HostMirror* mirror() { return mirror_; }
    // D = the DIC pivots of the current values (Prec1.f90:51-59), in the
    // internal numbering: one tiled launch in mode 3, else one per level
    void dic_setup(double* D, hipStream_t s);
    // values of the level-major triangles for the next DIC sweeps (FP64,
    // and FP32 rounded from them when fp32)
    void prepare_dic(bool fp32, hipStream_t s);
    // throws DicWaitFault when a one-launch sweep of the last solve gave up
    // waiting
    void check_dic(hipStream_t s) const;
    // the DIC sweep form this system runs (0 .. 3; builds the levels)
    int dic_form() { return levels().plan ? levels_.mode : 0; }
    // From now on sweep one level per launch, which needs no co-resident
    // grid: form 1 after form 2, form 0 (the matrix rows) after form 3,
    // whose positions are not level-major; the one-launch form stays off
    // for this system.
    void latch_dic_form1();
    // Runs a DIC solve `f`, whose in/out vectors are the `count` device
    // arrays `vecs` of n rows each.  When a one-launch sweep in it gave up
    // waiting (another process or stream held the CUs past the wait bound),
    // the vectors are restored, the system latches a form with one launch
    // per level and f runs again;
    // a second failure propagates.
    template <class F>
    SolveResult dic_retry(double* const* vecs, int count, hipStream_t s, F&& f) {
        if (dic_form() < 2) return f();
        dic_backup_.ensure((size_t)count * n_);
        for (int k = 0; k < count; ++k)
            TF_CHECK(hipMemcpyAsync(dic_backup_.p + (size_t)k * n_, vecs[k], sizeof(double) * n_,
                                    hipMemcpyDeviceToDevice, s));
        try {
            return f();
        } catch (const DicWaitFault&) {
            for (int k = 0; k < count; ++k)
                TF_CHECK(hipMemcpyAsync(vecs[k], dic_backup_.p + (size_t)k * n_, sizeof(double) * n_,
                                        hipMemcpyDeviceToDevice, s));
            latch_dic_form1();
            return f();
        }
    }
    template <typename T>
    void dic_apply(const Csr<T>& A, const T* D, const T* u, T* z, const SolverState* st, hipStream_t s);

   private:
    SolveResult solve_ext(int method, int prec, double* x, double* r1, int cap, double tol, double stale_error,
                          hipStream_t s);
    SolveResult solve_impl(int method, int prec, double* x, double* r1, int cap, double tol, double stale_error,
                           hipStream_t s);
    SolveResult solve_mixed_ext(int method, int prec, double* x, const double* b, double* r_out, int cap,
                                double tol, double inner_rtol, int max_outer, int* outer_steps, hipStream_t s);
    SolveResult solve_mixed_reliable_ext(int prec, double* x, const double* b, double* r_out, int cap, double tol,
                                         int interval, int* outer_steps, hipStream_t s);
    SolveResult solve_mixed_impl(int method, int prec, double* x, const double* b, double* r_out, int cap,
                                 double tol, double inner_rtol, int max_outer, int* outer_steps, hipStream_t s);
    SolveResult solve_mixed_reliable_dic(double* x, const double* b, double* r_out, int cap, double tol,
                                         int interval, int* outer_steps, hipStream_t s);
    SolveResult solve_mixed_reliable_impl(int prec, double* x, const double* b, double* r_out, int cap, double tol,
                                          int interval, int* outer_steps, hipStream_t s);
    void build_storage(bool allow_reorder, hipStream_t s);
    void build_dic_plan(const std::vector<int>& flev, bool sorted);
    void permute_values(hipStream_t s);
    double* ext_buffer(int k);   // k-th N-vector of the gather / scatter buffer

    int n_, nnz_;
    DevArray<int> rowptr_, col_, dia_;
    DevArray<double> val_;
    DevArray<float> val32_;
    bool val32_valid_ = false;
    DevArray<double> ws_;          // workspace vectors
    DevArray<double> partials_;
    DevArray<SolverState> state_;
    HostMirror* mirror_ = nullptr;  // coherent pinned host memory
    DevArray<float> ws32_;          // FP32 workspace of the mixed solver
    DevArray<float> partials32_;
    DevArray<SolverState> state32_;
    HostMirror* mirror32_ = nullptr;
    DevArray<double> r64_;
    DevArray<int> off_rowptr_, off_col_;
    DevArray<int> sell_off_, sell_col_, sell_map_;
    DevArray<short> sell_dc_;       // col - row, 16-bit, where every slot fits
    DevArray<int> sell_soff_;       // shared slot offsets, where the pattern allows
    DevArray<unsigned char> sell_mask_;
    DevArray<float> sell_val_;
    int64_t sell_entries_ = 0;
    bool sell_built_ = false, sell_valid_ = false;
    int op32_mode_ = 2;                  // 0 CSR, 1 sliced, 2 decide by timing on first use
    DevArray<int> sell64_off_, sell64_col_, sell64_map_;
    DevArray<short> sell64_dc_;     // col - row, 16-bit, where every slot fits
    DevArray<int> sell64_soff_;     // shared slot offsets, where the pattern allows
    DevArray<unsigned char> sell64_mask_;
    DevArray<double> sell64_val_;
    int64_t sell64_entries_ = 0;
    bool sell64_built_ = false, sell64_valid_ = false;
    // symmetric forms of the two sliced operators (where the pattern and the
    // values allow): slice offset starts, mirror bases, entry pairs
    DevArray<int> sell64_sidx_, sell64_pairs_, sell_sidx_, sell_pairs_;
    DevArray<SlotPair> sell64_lbase_, sell_lbase_;   // read by the kernels as int2
    int64_t sell64_npairs_ = 0, sell_npairs_ = 0;
    bool sym64_refused_ = false, sym32_refused_ = false;   // values seen that are not symmetric
    DevArray<int> symflag_;
    int op64_mode_ = 2;
    DevArray<float> off_val_, rowsum32_, diag32_;
    bool diff32_valid_ = false;
    DevArray<int> tpos_;
    bool tpos_ready_ = false;
    std::unique_ptr<System> transposed_;
    bool transposed_valid_ = false;
    bool symmetric_ = false;         // the values equal their mirrors: A is its own transpose
    DevArray<int> sym_flag_;
    std::vector<int> h_rowptr_, h_col_, h_dia_;
    Levels levels_;
    bool levels_ready_ = false;
    DevArray<double> dic_backup_;   // inputs of a DIC solve, for dic_retry
    int grid_;
    int lanes_ = 1;
    int strided_ = 0;
    Replay replay_;
    // locality reordering (reorder.hip): perm_ new -> old rows, eperm_
    // new -> old entries, the original pattern for the adjoint outer
    // product, the values in the caller's order
    bool reordered_ = false;
    std::vector<int> iperm_h_;
    DevArray<int> perm_, eperm_, orowptr_, ocol_;
    DevArray<double> val_orig_, extbuf_, tvals_;
    DevArray<float> extbuf32_;
};

int max_blocks_from_env();
bool fused_cg_enabled();
int cg_form();
// whether one system's Jacobi CG iterations carry x += alfa p into the next
// p kernel (then the solve ends with launch_cg_xfinish)
bool cg_carry();
bool carries_x(const BatchCtx& c, int method, bool jac);
template <typename T, typename Op>
void start_solve(const BatchCtx& c, int method, int prec, const Op& Aop, const Csr<T>& A,
                 const KrylovVectors<T>& w, const Sites<T>& S, SolverState* st, HostMirror* hm, const Levels* L,
                 hipStream_t s, bool d_ready, bool zero_x);
template <typename T, typename Op>
void run_solve(const BatchCtx& c, int method, int prec, int cap, const Op& Aop, const Csr<T>& A,
               const KrylovVectors<T>& w, const Sites<T>& S, SolverState* st, HostMirror* hm, const Levels* L,
               hipStream_t s, bool d_ready, bool zero_x, Replay* rp = nullptr);
template <typename T, typename Op>
void pace_iterations(const BatchCtx& c, int method, bool jac, int cap, const Op& Aop, const Csr<T>& A,
                     const KrylovVectors<T>& w, const Sites<T>& S, SolverState* st, HostMirror* hm,
                     const Levels* L, hipStream_t s, Replay* rp = nullptr);
template <typename T, typename Op>
void enqueue_iteration(const BatchCtx& c, int method, bool jac, int k, const Op& Aop, const Csr<T>& A,
                       const KrylovVectors<T>& w, const Sites<T>& S, SolverState* st, HostMirror* hm,
                       const Levels* L, hipStream_t s);
void dic_pivots_levels(const Levels& L, const Csr<double>& A, double* D, hipStream_t s);
template <typename T>
void dic_apply_levels(const Levels& L, const Csr<T>& A, const T* D, const T* u, T* z, const SolverState* st,
                      hipStream_t s);

}  // namespace tflows
