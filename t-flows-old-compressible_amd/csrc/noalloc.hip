// noalloc.hip -- a CG / BiCG solve on an existing system that allocates
// nothing: its workspace (the solver state, the reduction partials and the
// D, p, q vectors, plus the caller-order gather buffers of a renumbered
// system) lives in one file-scope device array, `scratch`, of fixed size.
//
// The kernels are the ones System::solve runs for CG with Jacobi or no
// preconditioning (the three-kernel iteration: stopping test and p update,
// SpMV with (p, q), x / r update), on the system's CSR arrays; the sliced
// operator System::solve may pick gives the same row sums bit for bit, and
// both sum the dot products over the same tiles, so the results are the
// same bits.  BiCG runs through the CG kernels, as System::solve does: the
// reference's shadow system reproduces the primal one (BiCG.f90:93-95,
// 196-200).
//
// Each GPU the library runs on holds its own copy of `scratch` until the
// process ends; its address is looked up once per device, with that
// device current, on the first call there.  Every call on one GPU shares
// that GPU's array: calls that overlap on one GPU (two host threads, or two
// streams) race and corrupt each other's results.  Whatever must start at
// zero is cleared by kernels on the caller's stream at the start of a call.
#include <algorithm>
#include <mutex>

#include "../../include/tflows_krylov.h"
#include "system.h"

namespace tflows {

// rows of the largest system a call accepts
constexpr int kNoAllocMaxRows = 1 << 20;
static_assert(kNoAllocMaxRows == TFLOWS_NOALLOC_MAX_ROWS, "the header states the row limit");
// blocks of the largest reduction grid (stream_grid's default cap is 2048)
constexpr int kNoAllocMaxGrid = 4096;
constexpr size_t kStateDoubles = 64;                              // 512 B, one SolverState
constexpr size_t kPartialDoubles = 8 * (size_t)kNoAllocMaxGrid;   // Sites<double>::elems(grid)
constexpr size_t kVectorDoubles = 6 * (size_t)kNoAllocMaxRows;    // D, p, q, z, x and r1 in internal order
constexpr size_t kScratchDoubles = kStateDoubles + kPartialDoubles + kVectorDoubles;
static_assert(sizeof(SolverState) <= kStateDoubles * sizeof(double), "SolverState outgrew its scratch slot");

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of the current device's `scratch`.  In host code the
// array's name is only a host-side stand-in; the address comes from
// hipGetSymbolAddress, once per device.
double* device_scratch() {
    int dev = 0;
    TF_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        TF_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

__global__ __launch_bounds__(kBlock) void k_zero(size_t count, double* __restrict__ a) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += (size_t)gridDim.x * kBlock)
        a[i] = 0.0;
}

// The initial state of a solve: every field zero, then tol, cap and the
// stale FinRes of the early exit (what System::solve copies from the host).
__global__ void k_state_init(SolverState* st, double tol, int cap, double stale_error) {
    if (blockIdx.x != 0) return;
    double* w = reinterpret_cast<double*>(st);
    for (size_t i = threadIdx.x; i < kStateDoubles; i += blockDim.x) w[i] = 0.0;
    __syncthreads();
    if (threadIdx.x == 0) {
        st->tol = tol;
        st->cap = cap;
        st->error = stale_error;
    }
}

}  // namespace

SolveResult solve_noalloc(System& S, int method, int prec, double* x, double* r1, int cap, double tol,
                          double stale_error, hipStream_t s) {
    if (method != 0 && method != 1)
        throw std::invalid_argument("the no-allocation solve runs CG (0) and BiCG (1) only");
    if (prec != 0 && prec != 1)
        throw std::invalid_argument("the no-allocation solve takes prec 0 (none) or 1 (diagonal) only");
    if (cg_form() == 2)
        throw std::invalid_argument("TFLOWS_CG_FORM=2 selects CG kernels whose workspace the no-allocation "
                                    "solve does not hold; unset it");
    const int n = S.n();
    if (n > kNoAllocMaxRows)
        throw std::invalid_argument("N = " + std::to_string(n) + " exceeds the no-allocation maximum of " +
                                    std::to_string(kNoAllocMaxRows) + " rows");
    if (S.grid() > kNoAllocMaxGrid)
        throw std::invalid_argument("the system's grid of " + std::to_string(S.grid()) +
                                    " blocks exceeds the no-allocation maximum of " +
                                    std::to_string(kNoAllocMaxGrid));
    if (!x || !r1) throw std::invalid_argument("null vector");
    double* base = device_scratch();
    SolverState* st = reinterpret_cast<SolverState*>(base);
    double* partials = base + kStateDoubles;
    double* vec = partials + kPartialDoubles;
    const size_t nn = (size_t)n;
    KrylovVectors<double> w;
    w.D = vec;
    w.p = vec + nn;
    w.q = vec + 2 * nn;
    w.u1 = vec + 3 * nn;            // z = r / D of the device-indexed CG kernels
    w.x = x;
    w.r1 = r1;
    const bool reord = S.reordered();
    if (reord) {   // the kernels run in the system's internal numbering
        w.x = vec + 4 * nn;
        w.r1 = vec + 5 * nn;
    }
    const unsigned zb = (unsigned)std::min<size_t>(2048, (4 * nn + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_zero, dim3(zb), dim3(kBlock), 0, s, 4 * nn, vec);
    hipLaunchKernelGGL(k_state_init, dim3(1), dim3(kWave), 0, s, st, tol, cap, stale_error);
    TF_CHECK(hipGetLastError());
    if (reord) {
        S.to_internal((const double*)x, w.x, s);
        S.to_internal((const double*)r1, w.r1, s);
    }
    const Sites<double> sites = Sites<double>::carve(partials, S.grid());
    const Csr<double> A = S.csr();
    const BatchCtx c = S.ctx();
    static const Levels kNoLevels;
    start_solve<double, Csr<double>>(c, 0, prec, A, A, w, sites, st, nullptr, &kNoLevels, s, false, false);
    TF_CHECK(hipGetLastError());
    // Iterations in chunks (4, doubling to 64); after each chunk the host
    // waits for the stream and reads the device's `done` flag.  Kernels
    // queued past convergence return at once.
    int issued = 0, chunk = 4;
    bool done = false;
    while (issued < cap && !done) {
        const int k = std::min(chunk, cap - issued);
        for (int i = 0; i < k; ++i) {
            ++issued;
            enqueue_iteration<double, Csr<double>>(c, 0, true, issued, A, A, w, sites, st, nullptr, &kNoLevels, s);
        }
        TF_CHECK(hipGetLastError());
        int flag = 0;
        TF_CHECK(hipMemcpyAsync(&flag, &st->done, sizeof(int), hipMemcpyDeviceToHost, s));
        TF_CHECK(hipStreamSynchronize(s));
        done = flag != 0;
        chunk = std::min(64, chunk * 2);
    }
    if (!done) launch_check(c, std::max(cap, 0) + 1, sites, st, nullptr, s);
    // the x update of the last iteration, which no p kernel applied
    if (carries_x(c, 0, true)) launch_cg_xfinish(A.n, w.x, (const double*)w.p, st, s);
    TF_CHECK(hipGetLastError());
    SolverState h;
    TF_CHECK(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s));
    TF_CHECK(hipStreamSynchronize(s));
    if (!h.done) throw HipError("solver finished without reaching a stopping state");
    if (reord) {
        S.to_external((const double*)w.x, x, s);
        S.to_external((const double*)w.r1, r1, s);
        TF_CHECK(hipStreamSynchronize(s));
    }
    SolveResult res;
    res.niter = h.niter_out;
    res.ini_written = h.ini_written != 0;
    res.ini_res = h.ini_res;
    res.fin_res = h.error;
    return res;
}

}  // namespace tflows
