// system.h -- host-side owner of one matrix on the device and of the
// workspace of its solves.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "krylov_kernels.h"

namespace tflows {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
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
};

class System {
   public:
    System(int n, int nonz, const int* acol, const int* arow, const int* adia, bool device_ptrs, hipStream_t s);
    ~System();

    int n() const { return n_; }
    int nnz() const { return nnz_; }

    void set_values(const double* aval, bool device_ptr, hipStream_t s);
    const double* values() const { return val_.p; }
    const float* values32(hipStream_t s);   // FP32 copy, made on first use

    Csr<double> csr() const { return Csr<double>{n_, rowptr_.p, col_.p, dia_.p, val_.p}; }
    Csr<float> csr32(hipStream_t s) { return Csr<float>{n_, rowptr_.p, col_.p, dia_.p, values32(s)}; }

    // Solve with x, r1 device pointers (x(1:N), r1(1:N)); `stale_error` is
    // the FinRes the early exit hands back.
    SolveResult solve(int method, int prec, double* x, double* r1, int cap, double tol, double stale_error,
                      hipStream_t s);

    void prec1(int prec, double* d, hipStream_t s);
    void spmv(const double* x, double* y, hipStream_t s);
    void spmv32(const float* x, float* y, hipStream_t s);
    void resid(const double* x, double* r, hipStream_t s);

    const Levels& levels();
    int grid() const { return grid_; }

    // raw pattern access for other components (batching, adjoint)
    const int* rowptr() const { return rowptr_.p; }
    const int* col() const { return col_.p; }
    const int* dia() const { return dia_.p; }

    Sites<double> sites();
    KrylovVectors<double> vectors(double* x, double* r1);
    SolverState* state() { return state_.p; }
    HostMirror* mirror() { return mirror_; }
    void dic_setup(double* D, hipStream_t s);
    template <typename T>
    void dic_apply(const Csr<T>& A, const T* D, const T* u, T* z, const SolverState* st, hipStream_t s);

   private:

    int n_, nnz_;
    DevArray<int> rowptr_, col_, dia_;
    DevArray<double> val_;
    DevArray<float> val32_;
    bool val32_valid_ = false;
    DevArray<double> ws_;          // workspace vectors
    DevArray<double> partials_;
    DevArray<SolverState> state_;
    HostMirror* mirror_ = nullptr;  // coherent pinned host memory
    std::vector<int> h_rowptr_, h_col_, h_dia_;
    Levels levels_;
    bool levels_ready_ = false;
    int grid_;
};

template <typename T>
void dic_apply_levels(const Levels& L, const Csr<T>& A, const T* D, const T* u, T* z, const SolverState* st,
                      hipStream_t s);
template <typename T>
void enqueue_iteration(int method, bool jac, int k, const Csr<T>& A, const KrylovVectors<T>& w, const Sites<T>& S,
                       SolverState* st, HostMirror* hm, const Levels* L, hipStream_t s);
template <typename T>
void pace_iterations(int method, bool jac, int cap, const Csr<T>& A, const KrylovVectors<T>& w, const Sites<T>& S,
                     SolverState* st, HostMirror* hm, const Levels* L, hipStream_t s);

}  // namespace tflows
