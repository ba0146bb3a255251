// capi.hip -- the extern "C" boundary of libtflows_krylov.so
// (declared in include/tflows_krylov.h).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tflows_krylov.h"
#include "batched.h"
#include "calcps.h"
#include "system.h"

using tflows::System;

struct tflows_system_s {
    std::unique_ptr<System> sys;
};
struct tflows_batch_s {
    std::unique_ptr<tflows::Batch> batch;
};
struct tflows_calcps_s {
    std::unique_ptr<tflows::CalcPS> cps;
};

namespace {

thread_local std::string g_last_error;

// The reference keeps `error` in a module (Library/sol_mod.f90:11) shared by
// cg, bicg and cgs; its value survives between calls and is what FinRes
// receives on the ||b|| < tol early exit.  One process-wide copy here.
double g_sol_error = 0.0;
std::mutex g_mutex;

int fail(const std::string& msg) {
    g_last_error = msg;
    return -1;
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        return fail(e.what());
    }
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

[[noreturn]] void die(const char* where) {
    std::fprintf(stderr, "libtflows_krylov: %s failed: %s\n", where, g_last_error.c_str());
    std::fflush(stderr);
    std::abort();
}

// Host-array drop-ins.  The reference fixes the pattern once (TopolM,
// Process/Processor.f90:100) and then calls cg / bicg with new values on it
// every outer iteration (CalcPS.f90:217, NewUVW.f90:506, CalcZmix.f90:394).
// The device System (validated pattern, locality order, sliced operators,
// DIC levels, the operator choice) is therefore kept between calls: a call
// whose N, NONZ, Acol, Arow and Adia equal the cached copy's (compared on
// the host) uploads only the values and the vectors; any difference
// rebuilds it.  prec1_ and resid_ share the cache.
struct HostCache {
    std::unique_ptr<System> sys;
    int n = -1, nonz = -1;
    std::vector<int32_t> acol, arow, adia;   // 1-based, as the caller passed them
    tflows::DevArray<double> x, r;
    int64_t builds = 0, calls = 0;
};
HostCache g_cache;

bool same(const std::vector<int32_t>& v, const int32_t* p, size_t count) {
    return v.size() == count && (count == 0 || std::memcmp(v.data(), p, sizeof(int32_t) * count) == 0);
}

// The cached system for this pattern; adia may be null (resid_), then any
// cached diagonal positions of the same Acol / Arow serve, and a new
// system takes the diagonal found by scanning each row.
System& dropin_system(int n, int nonz, const int32_t* Acol, const int32_t* Arow, const int32_t* Adia,
                      hipStream_t s) {
    ++g_cache.calls;
    if (n <= 0) throw std::invalid_argument("N must be positive");
    if (nonz < n + 1) throw std::invalid_argument("NONZ must be at least N + 1 (diagonal entries + 1)");
    const size_t nnz = (size_t)nonz - 1;
    if (g_cache.sys && g_cache.n == n && g_cache.nonz == nonz && same(g_cache.acol, Acol, (size_t)n + 1) &&
        same(g_cache.arow, Arow, nnz) && (!Adia || same(g_cache.adia, Adia, (size_t)n)))
        return *g_cache.sys;
    std::vector<int32_t> adia_scan;
    if (!Adia) {
        // Resid has no Adia argument; the pattern is validated with the
        // diagonal found by scanning each row
        if (Acol[0] != 1 || Acol[n] != nonz)
            throw std::invalid_argument("Acol must start at 1 and end at NONZ (= nnz + 1)");
        for (int i = 1; i <= n; ++i)
            if (Acol[i] < Acol[i - 1]) throw std::invalid_argument("Acol is not non-decreasing");
        adia_scan.resize(n);
        for (int i = 1; i <= n; ++i) {
            adia_scan[i - 1] = Acol[i - 1];
            for (int j = Acol[i - 1]; j < Acol[i]; ++j)
                if (Arow[j - 1] == i) adia_scan[i - 1] = j;
        }
        Adia = adia_scan.data();
    }
    g_cache.sys.reset();
    g_cache.n = g_cache.nonz = -1;
    g_cache.sys.reset(new System(n, nonz, Acol, Arow, Adia, false, s));
    g_cache.acol.assign(Acol, Acol + n + 1);
    g_cache.arow.assign(Arow, Arow + nnz);
    g_cache.adia.assign(Adia, Adia + n);
    g_cache.n = n;
    g_cache.nonz = nonz;
    ++g_cache.builds;
    return *g_cache.sys;
}

void drop_in(int method, const int32_t* N, const int32_t* NB, const int32_t* NONZ, const double* A,
             const int32_t* Acol, const int32_t* Arow, const int32_t* Adia, double* x, double* r1,
             const int32_t* prec, int32_t* niter, const double* tol, double* IniRes, double* FinRes) {
    std::lock_guard<std::mutex> lock(g_mutex);
    const int n = *N, nb = *NB;
    hipStream_t s = nullptr;
    System& S = dropin_system(n, *NONZ, Acol, Arow, Adia, s);
    S.set_values(A, false, s);
    g_cache.x.ensure(n);
    g_cache.r.ensure(n);
    TF_CHECK(hipMemcpyAsync(g_cache.x.p, x + nb + 1, sizeof(double) * n, hipMemcpyHostToDevice, s));
    TF_CHECK(hipMemcpyAsync(g_cache.r.p, r1, sizeof(double) * n, hipMemcpyHostToDevice, s));
    tflows::SolveResult res = S.solve(method, *prec, g_cache.x.p, g_cache.r.p, *niter, *tol, g_sol_error, s);
    TF_CHECK(hipMemcpyAsync(x + nb + 1, g_cache.x.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    TF_CHECK(hipMemcpyAsync(r1, g_cache.r.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    TF_CHECK(hipStreamSynchronize(s));
    *niter = res.niter;
    if (res.ini_written) *IniRes = res.ini_res;
    *FinRes = res.fin_res;
    g_sol_error = res.fin_res;
}

}  // namespace

namespace tflows {
// errors of host-side entry points compiled elsewhere (meshio.cpp)
int set_last_error(const std::string& msg) { return fail(msg); }
// the no-allocation solve (noalloc.hip)
SolveResult solve_noalloc(System& S, int method, int prec, double* x, double* r1, int cap, double tol,
                          double stale_error, hipStream_t s);
}  // namespace tflows


extern "C" {

const char* tflows_last_error(void) { return g_last_error.c_str(); }
const char* tflows_version(void) { return "tflows-krylov 0.1 gfx950"; }

#define TF_DROPIN(name, method)                                                                                 \
    void name(const int32_t* N, const int32_t* NB, const int32_t* NONZ, const double* A, const int32_t* Acol,   \
              const int32_t* Arow, const int32_t* Adia, const double* Ab, double* x, double* r1,               \
              const int32_t* prec, int32_t* niter, const double* tol, double* IniRes, double* FinRes) {        \
        (void)Ab;                                                                                               \
        if (guarded([&] { drop_in(method, N, NB, NONZ, A, Acol, Arow, Adia, x, r1, prec, niter, tol, IniRes,  \
                                  FinRes); }))                                                                  \
            die(#name);                                                                                         \
    }

TF_DROPIN(cg_, 0)
TF_DROPIN(bicg_, 1)
TF_DROPIN(cgs_, 2)

void prec1_(const int32_t* N, const int32_t* NONZ, const double* A, const int32_t* Acol, const int32_t* Arow,
            const int32_t* Adia, double* D, const int32_t* prec) {
    if (guarded([&] {
            std::lock_guard<std::mutex> lock(g_mutex);
            hipStream_t s = nullptr;
            const int n = *N;
            System& S = dropin_system(n, *NONZ, Acol, Arow, Adia, s);
            S.set_values(A, false, s);
            g_cache.x.ensure(n);
            S.prec1(*prec, g_cache.x.p, s);
            TF_CHECK(hipMemcpyAsync(D, g_cache.x.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
            TF_CHECK(hipStreamSynchronize(s));
        }))
        die("prec1_");
}

void resid_(const int32_t* N, const int32_t* NB, const int32_t* NONZ, const double* A, const int32_t* Acol,
            const int32_t* Arow, const double* Ab, const double* x, double* r1) {
    (void)Ab;
    if (guarded([&] {
            std::lock_guard<std::mutex> lock(g_mutex);
            hipStream_t s = nullptr;
            const int n = *N, nb = *NB;
            System& S = dropin_system(n, *NONZ, Acol, Arow, nullptr, s);
            S.set_values(A, false, s);
            g_cache.x.ensure(n);
            g_cache.r.ensure(n);
            TF_CHECK(hipMemcpyAsync(g_cache.x.p, x + nb + 1, sizeof(double) * n, hipMemcpyHostToDevice, s));
            TF_CHECK(hipMemcpyAsync(g_cache.r.p, r1, sizeof(double) * n, hipMemcpyHostToDevice, s));
            S.resid(g_cache.x.p, g_cache.r.p, s);
            TF_CHECK(hipMemcpyAsync(r1, g_cache.r.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
            TF_CHECK(hipStreamSynchronize(s));
        }))
        die("resid_");
}

int tflows_dropin_stats(int64_t* pattern_builds, int64_t* calls) {
    return guarded([&] {
        std::lock_guard<std::mutex> lock(g_mutex);
        if (pattern_builds) *pattern_builds = g_cache.builds;
        if (calls) *calls = g_cache.calls;
    });
}

int tflows_dropin_release(void) {
    return guarded([&] {
        std::lock_guard<std::mutex> lock(g_mutex);
        g_cache.sys.reset();
        g_cache.n = g_cache.nonz = -1;
        g_cache.acol.clear();
        g_cache.arow.clear();
        g_cache.adia.clear();
        g_cache.x.release();
        g_cache.r.release();
    });
}

int32_t tflows_system_dic_sweep(tflows_system s) {
    int32_t v = 0;
    if (guarded([&] { v = s->sys->dic_form(); })) return -1;
    return v;
}

int tflows_system_create(int32_t n, int32_t nonz, const int32_t* acol, const int32_t* arow, const int32_t* adia,
                         int32_t mem, void* stream, tflows_system* out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("out is NULL");
        auto h = std::make_unique<tflows_system_s>();
        h->sys.reset(new System(n, nonz, acol, arow, adia, mem == TFLOWS_MEM_DEVICE, as_stream(stream)));
        *out = h.release();
    });
}

int tflows_system_destroy(tflows_system s) {
    return guarded([&] { delete s; });
}

int tflows_system_set_values(tflows_system s, const double* aval, int32_t mem, void* stream) {
    return guarded([&] { s->sys->set_values(aval, mem == TFLOWS_MEM_DEVICE, as_stream(stream)); });
}

int tflows_solve(tflows_system s, int32_t method, int32_t prec, double* x, double* r1, int32_t* niter,
                 double tol, double* ini_res, double* fin_res, void* stream) {
    return guarded([&] {
        std::lock_guard<std::mutex> lock(g_mutex);
        tflows::SolveResult r = s->sys->solve(method, prec, x, r1, *niter, tol, g_sol_error, as_stream(stream));
        *niter = r.niter;
        if (r.ini_written && ini_res) *ini_res = r.ini_res;
        if (fin_res) *fin_res = r.fin_res;
        g_sol_error = r.fin_res;
    });
}

int tflows_solve_noalloc(tflows_system s, int32_t method, int32_t prec, double* x, double* r1, int32_t* niter,
                         double tol, double* ini_res, double* fin_res, void* stream) {
    return guarded([&] {
        if (!s || !niter) throw std::invalid_argument("null system or niter");
        std::lock_guard<std::mutex> lock(g_mutex);
        tflows::SolveResult r =
            tflows::solve_noalloc(*s->sys, method, prec, x, r1, *niter, tol, g_sol_error, as_stream(stream));
        *niter = r.niter;
        if (r.ini_written && ini_res) *ini_res = r.ini_res;
        if (fin_res) *fin_res = r.fin_res;
        g_sol_error = r.fin_res;
    });
}

int tflows_solve_mixed(tflows_system s, int32_t method, int32_t prec, double* x, const double* b, double* r_out,
                       int32_t* niter, double tol, double inner_rtol, int32_t max_outer, double* ini_res,
                       double* fin_res, int32_t* outer_steps, void* stream) {
    return guarded([&] {
        std::lock_guard<std::mutex> lock(g_mutex);
        int outer = 0;
        tflows::SolveResult r = s->sys->solve_mixed(method, prec, x, b, r_out, *niter, tol, inner_rtol, max_outer,
                                                    &outer, as_stream(stream));
        *niter = r.niter;
        if (ini_res) *ini_res = r.ini_res;
        if (fin_res) *fin_res = r.fin_res;
        if (outer_steps) *outer_steps = outer;
    });
}

int tflows_solve_mixed_reliable(tflows_system s, int32_t prec, double* x, const double* b, double* r_out,
                                int32_t* niter, double tol, int32_t interval, double* ini_res, double* fin_res,
                                int32_t* replacements, void* stream) {
    return guarded([&] {
        std::lock_guard<std::mutex> lock(g_mutex);
        int rep = 0;
        tflows::SolveResult r =
            s->sys->solve_mixed_reliable(prec, x, b, r_out, *niter, tol, interval, &rep, as_stream(stream));
        *niter = r.niter;
        if (ini_res) *ini_res = r.ini_res;
        if (fin_res) *fin_res = r.fin_res;
        if (replacements) *replacements = rep;
    });
}

int tflows_solve_adjoint(tflows_system s, int32_t method, int32_t prec, double* lambda, double* r1, int32_t* niter,
                         double tol, double* ini_res, double* fin_res, void* stream) {
    return guarded([&] {
        std::lock_guard<std::mutex> lock(g_mutex);
        System& T = s->sys->transposed(as_stream(stream));
        tflows::SolveResult r = T.solve(method, prec, lambda, r1, *niter, tol, g_sol_error, as_stream(stream));
        *niter = r.niter;
        if (r.ini_written && ini_res) *ini_res = r.ini_res;
        if (fin_res) *fin_res = r.fin_res;
        g_sol_error = r.fin_res;
    });
}

int tflows_outer_grad(tflows_system s, const double* lambda, const double* x, double* grad, void* stream) {
    return guarded([&] { s->sys->outer_grad(lambda, x, grad, as_stream(stream)); });
}

int tflows_transpose_values(tflows_system s, double* aval_t, void* stream) {
    return guarded([&] {
        System& T = s->sys->transposed(as_stream(stream));
        TF_CHECK(hipMemcpyAsync(aval_t, T.values_original(), sizeof(double) * T.nnz(), hipMemcpyDeviceToDevice,
                                as_stream(stream)));
    });
}

int tflows_spmv(tflows_system s, const double* x, double* y, void* stream) {
    return guarded([&] { s->sys->spmv(x, y, as_stream(stream)); });
}
int tflows_spmv32(tflows_system s, const float* x, float* y, void* stream) {
    return guarded([&] { s->sys->spmv32(x, y, as_stream(stream)); });
}
int tflows_prec1(tflows_system s, int32_t prec, double* d, void* stream) {
    return guarded([&] { s->sys->prec1(prec, d, as_stream(stream)); });
}
int tflows_resid(tflows_system s, const double* x, double* r, void* stream) {
    return guarded([&] { s->sys->resid(x, r, as_stream(stream)); });
}
int32_t tflows_system_levels(tflows_system s) {
    int32_t v = 0;
    if (guarded([&] { v = s->sys->levels().nfwd; })) return -1;
    return v;
}

int32_t tflows_system_reordered(tflows_system s) {
    int32_t v = 0;
    if (guarded([&] { v = s->sys->reordered() ? 1 : 0; })) return -1;
    return v;
}

int tflows_system_fp32_layout(tflows_system s, int32_t* sliced, int64_t* stored_entries, void* stream) {
    return guarded([&] {
        if (!sliced || !stored_entries) throw std::invalid_argument("null output pointer");
        const hipStream_t st = as_stream(stream);
        const bool sl = s->sys->op32_sliced(st);
        const int cb = sl ? s->sys->sell32_col_bytes(st) : 4;
        *sliced = sl ? (cb == 0 ? (s->sys->sell32_symmetric(st) ? 4 : 3) : (cb == 2 ? 2 : 1)) : 0;
        *stored_entries = sl ? s->sys->sell32_entries(st) : s->sys->nnz() - s->sys->n();
        TF_CHECK(hipStreamSynchronize(st));
    });
}

int tflows_system_fp64_layout(tflows_system s, int32_t* sliced, int64_t* stored_entries, void* stream) {
    return guarded([&] {
        if (!sliced || !stored_entries) throw std::invalid_argument("null output pointer");
        const hipStream_t st = as_stream(stream);
        const bool sl = s->sys->op64_sliced(st);
        const int cb = sl ? s->sys->sell64_col_bytes(st) : 4;
        *sliced = sl ? (cb == 0 ? (s->sys->sell64_symmetric(st) ? 4 : 3) : (cb == 2 ? 2 : 1)) : 0;
        *stored_entries = sl ? s->sys->sell64_entries(st) : s->sys->nnz();
        TF_CHECK(hipStreamSynchronize(st));
    });
}

int tflows_system_set_lanes(tflows_system s, int32_t lanes) {
    return guarded([&] { s->sys->set_lanes(lanes); });
}

// ---- batches of independent systems ---------------------------------------
int tflows_batch_create(int32_t k, const int32_t* n, const int32_t* nonz, const int32_t* const* acol,
                        const int32_t* const* arow, const int32_t* const* adia, tflows_batch* out) {
    return guarded([&] {
        if (!out || !n || !nonz || !acol || !arow || !adia) throw std::invalid_argument("NULL argument");
        auto h = std::make_unique<tflows_batch_s>();
        h->batch.reset(new tflows::Batch(k, n, nonz, acol, arow, adia));
        *out = h.release();
    });
}

int tflows_batch_destroy(tflows_batch b) {
    return guarded([&] { delete b; });
}

int tflows_batch_set_values(tflows_batch b, int32_t sys, const double* aval, int32_t mem, void* stream) {
    return guarded([&] { b->batch->set_values(sys, aval, mem == TFLOWS_MEM_DEVICE, as_stream(stream)); });
}

int tflows_batch_layout(tflows_batch b, int32_t* rows, int64_t* entries, int32_t* row_offsets) {
    return guarded([&] {
        if (rows) *rows = b->batch->rows();
        if (entries) *entries = b->batch->entries();
        if (row_offsets) {
            const auto& r = b->batch->row_offsets();
            std::copy(r.begin(), r.end(), row_offsets);
        }
    });
}

int tflows_batch_configure(tflows_batch b, int32_t grid, int32_t lanes, int32_t chunk, int32_t rebalance) {
    return guarded([&] {
        if (lanes != 0 && lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8)
            throw std::invalid_argument("lanes per row must be 0 (auto), 1, 2, 4 or 8");
        if (grid < 0 || chunk < 0) throw std::invalid_argument("grid and chunk must be >= 0");
        tflows::BatchConfig c;
        c.grid = grid;
        c.lanes = lanes;
        c.chunk = chunk;
        b->batch->configure(c);
        b->batch->set_rebalance(rebalance != 0);
    });
}

int tflows_batch_set_form(tflows_batch b, int32_t form) {
    return guarded([&] {
        if (form < 0 || form > 2) throw std::invalid_argument("form must be 0 (tuned), 1 (kernels) or 2 (persistent)");
        b->batch->set_form(form);
    });
}

int tflows_batch_set_groups(tflows_batch b, int32_t mode) {
    return guarded([&] {
        if (mode < 0 || mode > 2)
            throw std::invalid_argument("groups must be 0 (tuned), 1 (all systems at once) or 2 (MALL-sized groups)");
        b->batch->set_groups(mode);
    });
}

int32_t tflows_batch_groups(tflows_batch b, int32_t bits, int32_t method, int32_t prec) {
    int32_t v = -1;
    if (guarded([&] { v = b->batch->groups_in_use(bits, tflows::kernel_method(method), prec); })) return -1;
    return v;
}

int32_t tflows_batch_form(tflows_batch b, int32_t bits, int32_t method, int32_t prec) {
    int32_t v = -1;
    if (guarded([&] { v = b->batch->config_for(bits, tflows::kernel_method(method), prec).form; })) return -1;
    return v;
}

int tflows_batch_config(tflows_batch b, int32_t bits, int32_t method, int32_t prec, int32_t* grid, int32_t* lanes,
                        int32_t* chunk, int32_t* rebalances) {
    return guarded([&] {
        const tflows::BatchConfig c = b->batch->config_for(bits, tflows::kernel_method(method), prec);
        if (grid) *grid = c.grid;
        if (lanes) *lanes = c.lanes;
        if (chunk) *chunk = c.chunk;
        if (rebalances) *rebalances = b->batch->rebalances();
    });
}

int tflows_batch_solve(tflows_batch b, int32_t method, int32_t prec, double* x, double* r1, const int32_t* cap,
                       const double* tol, int32_t* niter, double* ini_res, double* fin_res, void* stream) {
    return guarded([&] {
        const int k = b->batch->count();
        if (!cap || !tol || !niter || !fin_res) throw std::invalid_argument("NULL argument");
        std::vector<tflows::SolveResult> res(k);
        for (int i = 0; i < k; ++i) res[i].fin_res = fin_res[i];
        b->batch->solve(method, prec, x, r1, cap, tol, res.data(), as_stream(stream));
        for (int i = 0; i < k; ++i) {
            niter[i] = res[i].niter;
            if (res[i].ini_written && ini_res) ini_res[i] = res[i].ini_res;
            fin_res[i] = res[i].fin_res;
        }
    });
}

int tflows_batch_solve_mixed(tflows_batch b, int32_t method, int32_t prec, double* x, const double* rhs,
                             double* r_out, const int32_t* cap, const double* tol, double inner_rtol,
                             int32_t max_outer, int32_t* niter, double* ini_res, double* fin_res,
                             int32_t* outer_steps, void* stream) {
    return guarded([&] {
        const int k = b->batch->count();
        if (!cap || !tol || !niter || !fin_res) throw std::invalid_argument("NULL argument");
        std::vector<tflows::SolveResult> res(k);
        int outer = 0;
        b->batch->solve_mixed(method, prec, x, rhs, r_out, cap, tol, inner_rtol, max_outer, res.data(), &outer,
                              as_stream(stream));
        for (int i = 0; i < k; ++i) {
            niter[i] = res[i].niter;
            if (ini_res) ini_res[i] = res[i].ini_res;
            fin_res[i] = res[i].fin_res;
        }
        if (outer_steps) *outer_steps = outer;
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// pressure-correction assembly
// ---------------------------------------------------------------------------
int tflows_calcps_create(int32_t nc, int32_t nbc, int32_t ns, int32_t nonz, const int32_t* sidec,
                         const int32_t* sidaij, const int32_t* adia, const int32_t* typebc, const double* sx,
                         const double* sy, const double* sz, const double* dx, const double* dy, const double* dz,
                         const double* f, const double* volume, int32_t mem, void* stream, tflows_calcps* out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("out is NULL");
        if (!sidec || !adia || !typebc || !volume || (ns > 0 && (!sidaij || !sx || !sy || !sz || !dx || !dy ||
                                                                 !dz || !f)))
            throw std::invalid_argument("null array argument");
        const double* geom[7] = {sx, sy, sz, dx, dy, dz, f};
        auto h = std::make_unique<tflows_calcps_s>();
        h->cps = std::make_unique<tflows::CalcPS>(nc, nbc, ns, nonz, sidec, sidaij, adia, typebc, geom, volume,
                                                  mem == TFLOWS_MEM_DEVICE, as_stream(stream));
        *out = h.release();
    });
}

int tflows_calcps_destroy(tflows_calcps h) {
    return guarded([&] { delete h; });
}

int32_t tflows_calcps_visit_order(tflows_calcps h) {
    int32_t v = 0;
    if (guarded([&] {
            if (!h) throw std::invalid_argument("null handle");
            v = h->cps->visit_order();
        }))
        return -1;
    return v;
}

int tflows_calcps_assemble(tflows_calcps h, int32_t nmat, double dt, const double* asave, const double* rho_n,
                           const double* rho_o, const double* rho_oo, const double* u, const double* v,
                           const double* w, const double* p, const double* px, const double* py, const double* pz,
                           double* aval, double* b, double* flux, double* abou, double* errmax, double* mean,
                           int32_t mem, void* stream) {
    return guarded([&] {
        if (!h) throw std::invalid_argument("null handle");
        const double* in[11] = {asave, rho_n, rho_o, rho_oo, u, v, w, p, px, py, pz};
        for (const double* q : in)
            if (!q) throw std::invalid_argument("null field");
        tflows::CalcPS& C = *h->cps;
        if (!aval || !b || (C.ns() > 0 && !flux) || (C.nbc() > 0 && !abou))
            throw std::invalid_argument("null output");
        const hipStream_t st = as_stream(stream);
        const size_t ncell = (size_t)C.nbc() + C.nc() + 1;
        if (mem == TFLOWS_MEM_DEVICE) {
            const tflows::CalcPSFields F{asave, rho_n, rho_o, rho_oo, u, v, w, p, px, py, pz};
            C.assemble(nmat, dt, F, aval, b, flux, abou, errmax, mean, st);
            return;
        }
        // host arrays: one device staging area, fields in, results out;
        // Flux and Abou enter with the caller's values (kept where the
        // reference leaves them)
        const size_t total = 11 * ncell + (size_t)C.nonz() + C.nc() + C.ns() + C.nbc();
        double* d = C.staging(total);
        const double* dev[11];
        for (int q = 0; q < 11; ++q) {
            TF_CHECK(hipMemcpyAsync(d + q * ncell, in[q], sizeof(double) * ncell, hipMemcpyHostToDevice, st));
            dev[q] = d + q * ncell;
        }
        double* daval = d + 11 * ncell;
        double* db = daval + C.nonz();
        double* dflux = db + C.nc();
        double* dabou = dflux + C.ns();
        if (C.ns() > 0) TF_CHECK(hipMemcpyAsync(dflux, flux, sizeof(double) * C.ns(), hipMemcpyHostToDevice, st));
        if (C.nbc() > 0) TF_CHECK(hipMemcpyAsync(dabou, abou, sizeof(double) * C.nbc(), hipMemcpyHostToDevice, st));
        const tflows::CalcPSFields F{dev[0], dev[1], dev[2], dev[3], dev[4], dev[5], dev[6], dev[7], dev[8], dev[9],
                                     dev[10]};
        C.assemble(nmat, dt, F, daval, db, dflux, dabou, errmax, mean, st);
        TF_CHECK(hipMemcpyAsync(aval, daval, sizeof(double) * C.nonz(), hipMemcpyDeviceToHost, st));
        TF_CHECK(hipMemcpyAsync(b, db, sizeof(double) * C.nc(), hipMemcpyDeviceToHost, st));
        if (C.ns() > 0) TF_CHECK(hipMemcpyAsync(flux, dflux, sizeof(double) * C.ns(), hipMemcpyDeviceToHost, st));
        if (C.nbc() > 0) TF_CHECK(hipMemcpyAsync(abou, dabou, sizeof(double) * C.nbc(), hipMemcpyDeviceToHost, st));
        TF_CHECK(hipStreamSynchronize(st));
    });
}

int tflows_calcps_correct(tflows_calcps h, double dt, const double* pp, const double* aval, const double* abou,
                          const double* asave, const double* rho_n, const double* rho_o, const double* rho_oo,
                          const double* px, const double* py, const double* pz, double* u, double* v, double* w,
                          double* flux, double* b, double* errmax, int32_t mem, void* stream) {
    return guarded([&] {
        if (!h) throw std::invalid_argument("null handle");
        const double* in[8] = {pp, asave, rho_n, rho_o, rho_oo, px, py, pz};
        for (const double* q : in)
            if (!q) throw std::invalid_argument("null field");
        tflows::CalcPS& C = *h->cps;
        if (!aval || !u || !v || !w || !b || (C.ns() > 0 && !flux) || (C.nbc() > 0 && !abou))
            throw std::invalid_argument("null array");
        const hipStream_t st = as_stream(stream);
        const size_t ncell = (size_t)C.nbc() + C.nc() + 1;
        double em = 0.0;
        if (mem == TFLOWS_MEM_DEVICE) {
            const tflows::CalcPSFields F{asave, rho_n, rho_o, rho_oo, nullptr, nullptr, nullptr, nullptr, px, py, pz};
            em = C.correct(dt, F, pp, aval, abou, u, v, w, flux, b, st);
        } else {
            const size_t total = 11 * ncell + (size_t)C.nonz() + C.nbc() + C.ns() + C.nc();
            double* d = C.staging(total);
            double* dv[11];
            const double* src[11] = {pp, asave, rho_n, rho_o, rho_oo, px, py, pz, u, v, w};
            for (int q = 0; q < 11; ++q) {
                dv[q] = d + q * ncell;
                TF_CHECK(hipMemcpyAsync(dv[q], src[q], sizeof(double) * ncell, hipMemcpyHostToDevice, st));
            }
            double* daval = d + 11 * ncell;
            double* dabou = daval + C.nonz();
            double* dflux = dabou + C.nbc();
            double* db = dflux + C.ns();
            TF_CHECK(hipMemcpyAsync(daval, aval, sizeof(double) * C.nonz(), hipMemcpyHostToDevice, st));
            if (C.nbc() > 0) TF_CHECK(hipMemcpyAsync(dabou, abou, sizeof(double) * C.nbc(), hipMemcpyHostToDevice, st));
            if (C.ns() > 0) TF_CHECK(hipMemcpyAsync(dflux, flux, sizeof(double) * C.ns(), hipMemcpyHostToDevice, st));
            const tflows::CalcPSFields F{dv[1], dv[2], dv[3], dv[4], nullptr, nullptr, nullptr, nullptr, dv[5], dv[6],
                                         dv[7]};
            em = C.correct(dt, F, dv[0], daval, dabou, dv[8], dv[9], dv[10], dflux, db, st);
            double* outs[3] = {u, v, w};
            for (int q = 0; q < 3; ++q)
                TF_CHECK(hipMemcpyAsync(outs[q], dv[8 + q], sizeof(double) * ncell, hipMemcpyDeviceToHost, st));
            if (C.ns() > 0) TF_CHECK(hipMemcpyAsync(flux, dflux, sizeof(double) * C.ns(), hipMemcpyDeviceToHost, st));
            TF_CHECK(hipMemcpyAsync(b, db, sizeof(double) * C.nc(), hipMemcpyDeviceToHost, st));
            TF_CHECK(hipStreamSynchronize(st));
        }
        if (errmax) *errmax = em;
    });
}

int tflows_calcps_cfl_peclet(tflows_calcps h, double dt, double visc, const double* rho_n, const double* flux,
                             double* cflmax, double* pemax, int32_t mem, void* stream) {
    return guarded([&] {
        if (!h) throw std::invalid_argument("null handle");
        tflows::CalcPS& C = *h->cps;
        if (!rho_n || (C.ns() > 0 && !flux) || !cflmax || !pemax) throw std::invalid_argument("null array");
        const hipStream_t st = as_stream(stream);
        if (mem == TFLOWS_MEM_DEVICE) {
            C.cfl_peclet(dt, visc, rho_n, flux, cflmax, pemax, st);
        } else {
            const size_t ncell = (size_t)C.nbc() + C.nc() + 1;
            double* d = C.staging(ncell + C.ns());
            TF_CHECK(hipMemcpyAsync(d, rho_n, sizeof(double) * ncell, hipMemcpyHostToDevice, st));
            if (C.ns() > 0)
                TF_CHECK(hipMemcpyAsync(d + ncell, flux, sizeof(double) * C.ns(), hipMemcpyHostToDevice, st));
            C.cfl_peclet(dt, visc, d, d + ncell, cflmax, pemax, st);
        }
    });
}

// ---------------------------------------------------------------------------
// momentum assembly
// ---------------------------------------------------------------------------
int tflows_calcps_uvw_setup(tflows_calcps h, const double* scoef, const double* fF, const int32_t* material,
                            int32_t nmaterials, int32_t mem, void* stream) {
    return guarded([&] {
        if (!h) throw std::invalid_argument("null handle");
        h->cps->uvw_setup(scoef, fF, material, nmaterials, mem == TFLOWS_MEM_DEVICE, as_stream(stream));
    });
}

int tflows_calcps_uvw_assemble(tflows_calcps h, const tflows_uvw_params* p, const double* rho_n,
                               const double* rho_o, const double* rho_oo, const double* ui_n, const double* ui_o,
                               const double* ui_oo, const double* source, const double* pi, const double* dUidi,
                               const double* dUidj, const double* dUidk, const double* dUjdi, const double* dUkdi,
                               const double* dUjdj, const double* dUkdk, const double* flux_u, const double* flux,
                               double* aval, double* b, double* abou, double* asave, double* rhoui, int32_t mem,
                               void* stream) {
    return guarded([&] {
        if (!h || !p) throw std::invalid_argument("null handle or parameters");
        const double* in[15] = {rho_n, rho_o, rho_oo, ui_n,  ui_o,  ui_oo, source, pi,
                                dUidi, dUidj, dUidk,  dUjdi, dUkdi, dUjdj, dUkdk};
        for (const double* q : in)
            if (!q) throw std::invalid_argument("null field");
        tflows::CalcPS& C = *h->cps;
        if (!aval || !b || !rhoui || (C.ns() > 0 && !flux_u) || (C.nbc() > 0 && !abou))
            throw std::invalid_argument("null array");
        tflows::UvwParams P;
        P.var = p->var;
        P.convec = p->convec;
        P.diffus = p->diffus;
        P.cross = p->cross;
        P.inert = p->inert;
        P.visc = p->visc;
        P.visc_rho = p->visc_rho;
        P.dt = p->dt;
        P.urf = p->urf;
        for (int m = 0; m < tflows::kMaxMaterials; ++m) {
            P.blend[m] = p->blend[m];
            P.blend_tem[m] = p->blend_tem[m];
            P.urfc[m] = p->urfc[m];
        }
        const hipStream_t st = as_stream(stream);
        const size_t ncell = (size_t)C.nbc() + C.nc() + 1;
        if (mem == TFLOWS_MEM_DEVICE) {
            const tflows::UvwFields F{rho_n, rho_o, rho_oo, ui_n, ui_o, ui_oo, source, pi,
                                      {dUidi, dUidj, dUidk, dUjdi, dUkdi, dUjdj, dUkdk}, flux_u, flux};
            C.uvw_assemble(P, F, aval, b, abou, asave, rhoui, st);
            return;
        }
        const size_t total = 17 * ncell + 2 * (size_t)C.ns() + C.nonz() + C.nc() + C.nbc();
        double* d = C.staging(total);
        const double* dv[15];
        for (int q = 0; q < 15; ++q) {
            TF_CHECK(hipMemcpyAsync(d + q * ncell, in[q], sizeof(double) * ncell, hipMemcpyHostToDevice, st));
            dv[q] = d + q * ncell;
        }
        double* dasave = d + 15 * ncell;
        double* drhoui = dasave + ncell;
        double* dflux = drhoui + ncell;
        double* daval = dflux + C.ns();
        double* db = daval + C.nonz();
        double* dabou = db + C.nc();
        double* dmass = dabou + C.nbc();
        if (C.ns() > 0) TF_CHECK(hipMemcpyAsync(dflux, flux_u, sizeof(double) * C.ns(), hipMemcpyHostToDevice, st));
        if (C.ns() > 0 && flux)
            TF_CHECK(hipMemcpyAsync(dmass, flux, sizeof(double) * C.ns(), hipMemcpyHostToDevice, st));
        // Asave keeps the caller's values outside 1..NC
        if (asave) TF_CHECK(hipMemcpyAsync(dasave, asave, sizeof(double) * ncell, hipMemcpyHostToDevice, st));
        const tflows::UvwFields F{dv[0], dv[1], dv[2], dv[3], dv[4], dv[5], dv[6], dv[7],
                                  {dv[8], dv[9], dv[10], dv[11], dv[12], dv[13], dv[14]}, dflux,
                                  flux ? dmass : nullptr};
        C.uvw_assemble(P, F, daval, db, dabou, asave ? dasave : nullptr, drhoui, st);
        TF_CHECK(hipMemcpyAsync(aval, daval, sizeof(double) * C.nonz(), hipMemcpyDeviceToHost, st));
        TF_CHECK(hipMemcpyAsync(b, db, sizeof(double) * C.nc(), hipMemcpyDeviceToHost, st));
        TF_CHECK(hipMemcpyAsync(rhoui, drhoui, sizeof(double) * ncell, hipMemcpyDeviceToHost, st));
        if (C.nbc() > 0) TF_CHECK(hipMemcpyAsync(abou, dabou, sizeof(double) * C.nbc(), hipMemcpyDeviceToHost, st));
        if (asave) TF_CHECK(hipMemcpyAsync(asave, dasave, sizeof(double) * ncell, hipMemcpyDeviceToHost, st));
        TF_CHECK(hipStreamSynchronize(st));
    });
}

int tflows_calcps_uvw_update(tflows_calcps h, const double* rhoui, const double* rho_n, double* ui_n, int32_t mem,
                             void* stream) {
    return guarded([&] {
        if (!h || !rhoui || !rho_n || !ui_n) throw std::invalid_argument("null argument");
        tflows::CalcPS& C = *h->cps;
        const hipStream_t st = as_stream(stream);
        if (mem == TFLOWS_MEM_DEVICE) {
            C.uvw_update(rhoui, rho_n, ui_n, st);
            return;
        }
        const size_t ncell = (size_t)C.nbc() + C.nc() + 1;
        double* d = C.staging(3 * ncell);
        TF_CHECK(hipMemcpyAsync(d, rhoui, sizeof(double) * ncell, hipMemcpyHostToDevice, st));
        TF_CHECK(hipMemcpyAsync(d + ncell, rho_n, sizeof(double) * ncell, hipMemcpyHostToDevice, st));
        TF_CHECK(hipMemcpyAsync(d + 2 * ncell, ui_n, sizeof(double) * ncell, hipMemcpyHostToDevice, st));
        C.uvw_update(d, d + ncell, d + 2 * ncell, st);
        TF_CHECK(hipMemcpyAsync(ui_n, d + 2 * ncell, sizeof(double) * ncell, hipMemcpyDeviceToHost, st));
        TF_CHECK(hipStreamSynchronize(st));
    });
}

// ---------------------------------------------------------------------------
// mixture-fraction assembly
// ---------------------------------------------------------------------------
int tflows_calcps_zmix_assemble(tflows_calcps h, const tflows_uvw_params* p, const double* rho_n,
                                const double* rho_o, const double* rho_oo, const double* z_n, const double* z_o,
                                const double* z_oo, const double* source, const double* dZdx, const double* dZdy,
                                const double* dZdz, const double* flux_u, const double* flux, double* aval,
                                double* b, double* abou, double* rhoz, int32_t mem, void* stream) {
    return guarded([&] {
        if (!h || !p) throw std::invalid_argument("null handle or parameters");
        const double* in[10] = {rho_n, rho_o, rho_oo, z_n, z_o, z_oo, source, dZdx, dZdy, dZdz};
        for (const double* q : in)
            if (!q) throw std::invalid_argument("null field");
        tflows::CalcPS& C = *h->cps;
        if (!aval || !b || !rhoz || (C.ns() > 0 && !flux_u) || (C.nbc() > 0 && !abou))
            throw std::invalid_argument("null array");
        tflows::UvwParams P;
        P.convec = p->convec;
        P.diffus = p->diffus;
        P.cross = p->cross;
        P.inert = p->inert;
        P.visc = p->visc;
        P.visc_rho = p->visc_rho;
        P.dt = p->dt;
        P.urf = p->urf;
        for (int m = 0; m < tflows::kMaxMaterials; ++m) {
            P.blend_tem[m] = p->blend_tem[m];
            P.urfc[m] = p->urfc[m];
        }
        const hipStream_t st = as_stream(stream);
        const size_t ncell = (size_t)C.nbc() + C.nc() + 1;
        if (mem == TFLOWS_MEM_DEVICE) {
            const tflows::UvwFields F{rho_n, rho_o, rho_oo, z_n, z_o, z_oo, source, nullptr,
                                      {dZdx, dZdy, dZdz, nullptr, nullptr, nullptr, nullptr}, flux_u, flux};
            C.zmix_assemble(P, F, aval, b, abou, rhoz, st);
            return;
        }
        const size_t total = 11 * ncell + 2 * (size_t)C.ns() + C.nonz() + C.nc() + C.nbc();
        double* d = C.staging(total);
        const double* dv[10];
        for (int q = 0; q < 10; ++q) {
            TF_CHECK(hipMemcpyAsync(d + q * ncell, in[q], sizeof(double) * ncell, hipMemcpyHostToDevice, st));
            dv[q] = d + q * ncell;
        }
        double* drhoz = d + 10 * ncell;
        double* dflux = drhoz + ncell;
        double* daval = dflux + C.ns();
        double* db = daval + C.nonz();
        double* dabou = db + C.nc();
        double* dmass = dabou + C.nbc();
        if (C.ns() > 0) TF_CHECK(hipMemcpyAsync(dflux, flux_u, sizeof(double) * C.ns(), hipMemcpyHostToDevice, st));
        if (C.ns() > 0 && flux)
            TF_CHECK(hipMemcpyAsync(dmass, flux, sizeof(double) * C.ns(), hipMemcpyHostToDevice, st));
        const tflows::UvwFields F{dv[0], dv[1], dv[2], dv[3], dv[4], dv[5], dv[6], nullptr,
                                  {dv[7], dv[8], dv[9], nullptr, nullptr, nullptr, nullptr}, dflux,
                                  flux ? dmass : nullptr};
        C.zmix_assemble(P, F, daval, db, dabou, drhoz, st);
        TF_CHECK(hipMemcpyAsync(aval, daval, sizeof(double) * C.nonz(), hipMemcpyDeviceToHost, st));
        TF_CHECK(hipMemcpyAsync(b, db, sizeof(double) * C.nc(), hipMemcpyDeviceToHost, st));
        TF_CHECK(hipMemcpyAsync(rhoz, drhoz, sizeof(double) * ncell, hipMemcpyDeviceToHost, st));
        if (C.nbc() > 0) TF_CHECK(hipMemcpyAsync(abou, dabou, sizeof(double) * C.nbc(), hipMemcpyDeviceToHost, st));
        TF_CHECK(hipStreamSynchronize(st));
    });
}

int tflows_calcps_zmix_update(tflows_calcps h, const double* rhoz, const double* rho_n, double* z_n, int32_t mem,
                              void* stream) {
    return guarded([&] {
        if (!h || !rhoz || !rho_n || !z_n) throw std::invalid_argument("null argument");
        tflows::CalcPS& C = *h->cps;
        const hipStream_t st = as_stream(stream);
        if (mem == TFLOWS_MEM_DEVICE) {
            C.zmix_update(rhoz, rho_n, z_n, st);
            return;
        }
        const size_t ncell = (size_t)C.nbc() + C.nc() + 1;
        double* d = C.staging(3 * ncell);
        TF_CHECK(hipMemcpyAsync(d, rhoz, sizeof(double) * ncell, hipMemcpyHostToDevice, st));
        TF_CHECK(hipMemcpyAsync(d + ncell, rho_n, sizeof(double) * ncell, hipMemcpyHostToDevice, st));
        TF_CHECK(hipMemcpyAsync(d + 2 * ncell, z_n, sizeof(double) * ncell, hipMemcpyHostToDevice, st));
        C.zmix_update(d, d + ncell, d + 2 * ncell, st);
        TF_CHECK(hipMemcpyAsync(z_n, d + 2 * ncell, sizeof(double) * ncell, hipMemcpyDeviceToHost, st));
        TF_CHECK(hipStreamSynchronize(st));
    });
}
