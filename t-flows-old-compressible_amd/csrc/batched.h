// batched.h -- many independent systems solved together.
//
// K systems (any sizes, any patterns) are stored as one block-diagonal
// matrix.  Every solver kernel of an iteration runs once over all of them:
// each block belongs to one system and owns a range of its 256-row tiles,
// every system keeps its own SolverState, dot products sum over its own
// blocks only, and each system stops on its own test.  Iterations are
// replayed from captured HIP graphs (chunks of iterations whose index is
// read relative to a device counter), and when systems finish, their
// blocks are handed to the systems still running.
#pragma once
#include <functional>
#include <map>
#include <tuple>

#include "system.h"

namespace tflows {

struct BatchConfig {
    int grid = 0;    // blocks per launch (multiple of 8)
    int lanes = 0;   // SpMV form: lanes per row (1, 2, 4, 8)
    int chunk = 0;   // iterations per graph replay
    int form = 0;    // 1: one launch per kernel; 2: one persistent launch per chunk (CG, Jacobi / none)
    int groups = 0;  // 1: every system from the start; 2: in groups that fit the MALL, one after another
};
enum { kFormKernels = 1, kFormPersistent = 2 };

class Batch {
   public:
    // Host arrays of each system in the T-FlowS layout (1-based).
    Batch(int k, const int* n, const int* nonz, const int* const* acol, const int* const* arow,
          const int* const* adia);
    ~Batch();

    int count() const { return k_; }
    int rows() const { return n_tot_; }
    int64_t entries() const { return nnz_tot_; }
    const std::vector<int>& row_offsets() const { return roff_; }
    const std::vector<int>& entry_offsets() const { return eoff_; }

    // values A(1:NONZ_k - 1) of system `sys`, or of all systems (sys = -1,
    // concatenated in system order) -- host or device memory
    void set_values(int sys, const double* aval, bool device_ptr, hipStream_t s);

    // FP64 solves of every system: x, r1 device arrays of rows() doubles
    // (x in/out, r1 in: b, out: final residual); cap and tol per system;
    // res[k].fin_res on input is the value returned on the ||b|| < tol exit.
    void solve(int method, int prec, double* x, double* r1, const int* cap, const double* tol, SolveResult* res,
               hipStream_t s);
    // FP32 inner solves inside FP64 iterative refinement (restart form), per
    // system; b is not modified, r_out (may be null) receives b - A x.
    void solve_mixed(int method, int prec, double* x, const double* b, double* r_out, const int* cap,
                     const double* tol, double inner_rtol, int max_outer, SolveResult* res, int* outer_steps,
                     hipStream_t s);

    // Fixed launch configuration (0 = let the autotuner choose), and the
    // configuration in use for (precision bits, method, prec).
    void configure(const BatchConfig& c) { fixed_ = c; }
    BatchConfig config_for(int bits, int method, int prec) const;
    void set_rebalance(bool on) { rebalance_ = on; }
    int rebalances() const { return rebalances_; }
    // 0: the autotuner picks the form; 1 / 2 fix it (2 applies where the
    // persistent form exists: CG with prec 0 / 1; elsewhere 1 runs)
    void set_form(int form) { form_ = form; }
    // 0: the autotuner decides; 1: every system from the first iteration;
    // 2: systems in groups whose working sets fit the 256 MB MALL, one group
    // after another (form 1, Jacobi or no preconditioner)
    void set_groups(int mode) { groups_ = mode; }
    // the number of groups the solves of (bits, method, prec) run in
    int groups_in_use(int bits, int method, int prec) const;
    // false once a persistent launch gave up waiting on its barriers; the
    // batch then keeps the per-kernel form
    bool persistent_ok() const { return persistent_ok_; }

   private:
    struct Key {
        int bits, method, prec;
        bool operator<(const Key& o) const {
            return std::tie(bits, method, prec) < std::tie(o.bits, o.method, o.prec);
        }
    };
    struct GraphKey {
        int bits, method, prec, grid, lanes, chunk, form;
        bool operator<(const GraphKey& o) const {
            return std::tie(bits, method, prec, grid, lanes, chunk, form) <
                   std::tie(o.bits, o.method, o.prec, o.grid, o.lanes, o.chunk, o.form);
        }
    };

    // block maps of the running systems for a grid of G blocks, built in
    // the pinned staging buffer, then copied to the device
    void build_maps(const std::vector<char>& active, int G);
    void upload_maps(hipStream_t s);
    BatchCtx ctx(int lanes, bool with_kbase) const;
    void join(hipStream_t s);   // batch stream waits for the caller's stream
    KrylovVectors<double> vectors64();
    KrylovVectors<float> vectors32();

    template <typename T, typename Op>
    void run(int bits, int method, int prec, const Op& Aop, const Csr<T>& A, const KrylovVectors<T>& w,
             T* partials, SolverState* st, HostMirror* hm, const std::vector<SolverState>& init, bool d_ready,
             bool zero_x, const std::function<void()>& prepare, std::vector<SolverState>& out);
    // whether the persistent form can run (method, preconditioner, at
    // least one resident block per system) and its largest grid
    template <typename Op>
    int persistent_grid(int method, int prec, int lanes);
    template <typename T, typename Op>
    void enqueue_chunk(const GraphKey& gk, const BatchCtx& c, const Op& Aop, const Csr<T>& A,
                       const KrylovVectors<T>& w, T* partials, SolverState* st, HostMirror* hm, hipStream_t s);
    template <typename T, typename Op>
    BatchConfig tune(int bits, int method, int prec, const Op& Aop, const Csr<T>& A, const KrylovVectors<T>& w,
                     T* partials, SolverState* st, HostMirror* hm, const std::vector<SolverState>& init,
                     bool d_ready, bool zero_x, const std::function<void()>& prepare);
    template <typename T, typename Op>
    hipGraphExec_t graph(const GraphKey& gk, const Op& Aop, const Csr<T>& A, const KrylovVectors<T>& w,
                         T* partials, SolverState* st, HostMirror* hm);
    template <typename T>
    void rebalance(const std::vector<char>& active, int grid, T* partials, SolverState* st);
    void tile_norms(const double* v, std::vector<double>& out);
    // consecutive systems packed while their working sets (the operator and
    // the method's vectors) stay within mall_budget_ bytes
    std::vector<std::vector<int>> mall_groups(int bits, int method) const;
    // caller order <-> stored order of a batch vector (on the batch stream)
    void gather_in(const double* ext, double* in);
    void scatter_out(const double* in, double* ext);

    int k_, n_tot_;
    int64_t nnz_tot_;
    std::vector<int> roff_, eoff_, tiles_, tile0_;
    std::unique_ptr<System> sys_;      // the block-diagonal matrix
    hipStream_t bs_ = nullptr;         // stream of every batch launch (capturable)
    DevArray<int> tile_r0_, tile_sys_, maps_, kbase_, koff_;
    int* h_maps_ = nullptr;            // pinned staging of the block maps
    int* h_koff_ = nullptr;            // pinned staging of the per-system iteration offsets
    int map_grid_ = 0;
    int gmax_ = 0;                     // largest grid (size of the map arrays)
    DevArray<double> ws_, partials_, tnorm_, scale_;
    DevArray<float> ws32_, partials32_;
    DevArray<SolverState> st_, st32_;
    HostMirror* hm_ = nullptr;         // K mirrors each, coherent pinned memory
    HostMirror* hm32_ = nullptr;
    // the `done` flags of every system at the end of each replayed chunk,
    // two slots by chunk parity (coherent pinned memory): the host's
    // rebalancing decisions read the state at a fixed chunk boundary, so
    // they (and every result) do not depend on how far the device has run
    int* snap_ = nullptr;
    BatchConfig fixed_;
    std::map<Key, BatchConfig> tuned_;
    std::map<GraphKey, hipGraphExec_t> graphs_;
    bool rebalance_ = true;
    // locality order of renumbered systems: perm_ stored row -> caller row,
    // eperm_ stored entry -> caller entry (whole batch); val_orig_ the
    // values in the callers' order
    bool reordered_ = false;
    DevArray<int> perm_, eperm_;
    DevArray<double> val_orig_, val_tmp_;
    int rebalances_ = 0;
    int form_ = 0;                     // fixed form (0: tuned)
    int groups_ = 0;                   // fixed grouping (0: tuned)
    // working-set budget of a group: the MALL (Infinity Cache) holds 256 MB,
    // a group's operator and vectors are kept within 3/4 of it so that they
    // stay resident from one iteration to the next while the group runs
    double mall_budget_ = 192.0 * 1024 * 1024;
    bool persistent_ok_ = true;
    DevArray<int> bar_, pfault_;       // barrier counters (one per system) and the fault word of the persistent form
    uint64_t wait_ticks_ = 0;
    hipEvent_t ev_[5] = {};            // pacing (2), tuner timing (2), join
};

}  // namespace tflows
