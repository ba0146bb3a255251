// calcps.hip -- the pressure-correction assembly of T-FlowS
// (Process/CalcPS.f90:51-202) on the GPU: the matrix values, the mass
// source b and the face fluxes, written into device memory where the
// pressure solve that follows (CalcPS.f90:214-219) reads them.
//
// The reference runs one loop over the faces (:68-168) that scatters into
// the diagonal and b of both cells, then a loop over the cells (:188-199).
// Here:
//   k_calcps_faces  one thread per face: SMDPN, the interpolated mass
//                   velocities, A12 and the flux, in the reference's
//                   operation order; writes the off-diagonal pair
//                   Aval(SidAij(1:2,s)) = -A12, Abou, Flux(s) and the face's
//                   diagonal term;
//   k_calcps_cells  one thread per cell: the diagonal and b summed over the
//                   cell's faces in ascending face order -- the order in
//                   which the reference's face loop reaches them, so both
//                   are the reference's bit for bit (Nmat passes included);
//                   max |b| (:172-175); the time term (:189) and the
//                   partial sums of b V and V per block;
//   k_calcps_total  one block totals the partials in a fixed order: max |b|
//                   and bc_sum / vol;
//   k_calcps_mean   b -= bc_sum / vol (:201-204).
// The reference sums b V and V over the cells one after another; here the
// sum is a tree over blocks, so the mean and b differ from the reference by
// the rounding of that sum (deterministic from run to run).
//
// After the pressure solve, CorUVW (Process/CorUVW.f90:36-121) corrects the
// velocities (k_cor_velocity, one thread per cell), copies them to PRESSURE
// boundary cells and corrects the face fluxes with the pressure correction
// (k_cor_faces, one thread per face), and forms the mass error per unit
// volume over the same cell-to-face lists (k_cor_cells) and its maximum,
// all bitwise the reference's.
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "calcps.h"
#include "device_common.h"

namespace tflows {

namespace {

constexpr int kBcInflow = 1, kBcOutflow = 3, kBcConvect = 5, kBcBuffer = 11, kBcPressure = 12;  // allp_mod.f90

// what a face does to its cells (fixed by the face list and TypeBC)
enum : unsigned char {
    kNone = 0,      // SYMMETRY, walls: Flux = 0 (:160-161)
    kInterior = 1,  // diagonal of c1 and c2, b(c1) -= F, b(c2) += F (:86-94, :114-115)
    kBuffer = 2,    // diagonal of c1, Abou(c2), b(c1) -= F (:96-101)
    kInflow = 3,    // b(c1) -= F only (:120-133)
    kDiagOut = 4,   // OUTFLOW / CONVECT / PRESSURE: diagonal of c1, b(c1) -= F (:134-159)
    kPressure = 5,  // as kDiagOut, with the cell's own mass velocity (:150-152)
};

__device__ __forceinline__ bool adds_diag(unsigned char k) {
    return k == kInterior || k == kBuffer || k == kDiagOut || k == kPressure;
}

__global__ __launch_bounds__(kBlock) void k_calcps_faces(CalcPSMesh M, CalcPSFields F, double* __restrict__ aval,
                                                          double* __restrict__ flux, double* __restrict__ abou,
                                                          double* __restrict__ fa12) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= M.ns) return;
    const unsigned char k = M.kind[s];
    if (k == kNone) {
        flux[s] = 0.0;
        fa12[s] = 0.0;
        return;
    }
    const int nb = M.nbc;
    const int c1 = M.c1[s], c2 = M.c2[s];
    const double Sx = M.sx[s], Sy = M.sy[s], Sz = M.sz[s];
    const double fs = M.f[s];
    double Urs, Vrs, Wrs;
    if (k == kPressure) {                                              // :150-152
        Urs = F.rho_n[nb + c1] * F.u[nb + c1];
        Vrs = F.rho_n[nb + c1] * F.v[nb + c1];
        Wrs = F.rho_n[nb + c1] * F.w[nb + c1];
    } else {                                                           // :81-83
        const double r1 = fs * F.rho_n[nb + c1], r2 = (1.0 - fs) * F.rho_n[nb + c2];
        Urs = r1 * F.u[nb + c1] + r2 * F.u[nb + c2];
        Vrs = r1 * F.v[nb + c1] + r2 * F.v[nb + c2];
        Wrs = r1 * F.w[nb + c1] + r2 * F.w[nb + c2];
    }
    const double mass = (Urs * Sx + Vrs * Sy) + Wrs * Sz;
    if (k == kInflow) {                                                // :131-132
        flux[s] = mass;
        fa12[s] = 0.0;
        return;
    }
    const double Dx = M.dx[s], Dy = M.dy[s], Dz = M.dz[s];
    const double SMDPN = ((Sx * Sx + Sy * Sy) + Sz * Sz) / ((Sx * Dx + Sy * Dy) + Sz * Dz);   // :77-78
    if (k == kDiagOut || k == kPressure) {                             // :143-147, :156-158
        const double A12 = (SMDPN * M.volume[nb + c1]) / F.asave[nb + c1];
        flux[s] = mass;
        fa12[s] = A12;
        return;
    }
    // interior and BUFFER faces (:86-112)
    const double A12 = (0.5 * SMDPN) * (M.volume[nb + c1] / F.asave[nb + c1] + M.volume[nb + c2] / F.asave[nb + c2]);
    if (k == kInterior) {
        aval[M.sid1[s]] = -A12;
        aval[M.sid2[s]] = -A12;
    } else {
        abou[nb + c2] = -A12;                                          // Abou(-NbC:-1)
    }
    const double dPxi = (0.5 * (F.px[nb + c1] + F.px[nb + c2])) * Dx;
    const double dPyi = (0.5 * (F.py[nb + c1] + F.py[nb + c2])) * Dy;
    const double dPzi = (0.5 * (F.pz[nb + c1] + F.pz[nb + c2])) * Dz;
    flux[s] = (mass + A12 * (F.p[nb + c1] - F.p[nb + c2])) + A12 * ((dPxi + dPyi) + dPzi);
    fa12[s] = A12;
}

// Partials per block: max |b| before the time term, sum of b V, sum of V.
__global__ __launch_bounds__(kBlock) void k_calcps_cells(CalcPSMesh M, CalcPSFields F, int nmat, double dt,
                                                          const double* __restrict__ flux,
                                                          const double* __restrict__ fa12,
                                                          double* __restrict__ aval, double* __restrict__ b,
                                                          double* __restrict__ part) {
    const int nb = M.nbc;
    double vmax = 0.0, sbv = 0.0, sv = 0.0;
    const int rb = blockIdx.x * kBlock + threadIdx.x;
    for (int i = rb; i < M.nc; i += gridDim.x * kBlock) {
        const int e0 = M.cptr[i], e1 = M.cptr[i + 1];
        double diag = 0.0, bb = 0.0;                                   // :51, :56
        for (int m = 0; m < nmat; ++m) {
            for (int e = e0; e < e1; ++e) {
                const int code = M.cface[e];
                const int s = code >> 1;
                const unsigned char k = M.kind[s];
                if (k == kNone) continue;                              // no flux in the assembly
                if (adds_diag(k)) diag = diag + fa12[s];
                if (code & 1) bb = bb + flux[s];                       // the cell is c2
                else bb = bb - flux[s];
            }
        }
        aval[M.adia[i]] = diag;
        vmax = fmax(vmax, fabs(bb));                                   // :173-175
        const int c = i + 1;
        const double V = M.volume[nb + c];
        const double term = (V / dt) * ((1.5 * F.rho_n[nb + c] - 2.0 * F.rho_o[nb + c]) + 0.5 * F.rho_oo[nb + c]);
        for (int m = 0; m < nmat; ++m) {                               // :188-197
            bb = bb - term;
            sbv += bb * V;
            sv += V;
        }
        b[i] = bb;
    }
    __shared__ double sm[3][kWavesPerBlock];
    // block max and sums, in a fixed order
    double v3[3] = {vmax, sbv, sv};
    for (int off = kWave / 2; off > 0; off >>= 1) {
        v3[0] = fmax(v3[0], __shfl_down(v3[0], off, kWave));
        v3[1] += __shfl_down(v3[1], off, kWave);
        v3[2] += __shfl_down(v3[2], off, kWave);
    }
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (lane == 0)
        for (int q = 0; q < 3; ++q) sm[q][wave] = v3[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = sm[0][0], t1 = sm[1][0], t2 = sm[2][0];
        for (int wv = 1; wv < kWavesPerBlock; ++wv) {
            t0 = fmax(t0, sm[0][wv]);
            t1 += sm[1][wv];
            t2 += sm[2][wv];
        }
        part[blockIdx.x] = t0;
        part[gridDim.x + blockIdx.x] = t1;
        part[2 * gridDim.x + blockIdx.x] = t2;
    }
}

// The trailing NONZ slot and any pattern entry no face writes stay zero (:51).
__global__ __launch_bounds__(kBlock) void k_calcps_zero(int count, const int* __restrict__ idx,
                                                         double* __restrict__ aval) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < count) aval[idx[t]] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_calcps_zero_all(int count, double* __restrict__ aval) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < count) aval[t] = 0.0;
}

// One block: max |b| and bc_sum / vol from the block partials, summed in a
// fixed order (thread-strided, then a fixed tree).
__global__ __launch_bounds__(kBlock) void k_calcps_total(int parts, const double* __restrict__ part,
                                                          double* __restrict__ out2) {
    double v3[3] = {0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < parts; q += kBlock) {
        v3[0] = fmax(v3[0], part[q]);
        v3[1] += part[parts + q];
        v3[2] += part[2 * parts + q];
    }
    __shared__ double sm[3][kBlock];
    for (int k = 0; k < 3; ++k) sm[k][threadIdx.x] = v3[k];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sm[0][threadIdx.x] = fmax(sm[0][threadIdx.x], sm[0][threadIdx.x + w]);
            sm[1][threadIdx.x] += sm[1][threadIdx.x + w];
            sm[2][threadIdx.x] += sm[2][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out2[0] = sm[0][0];
        out2[1] = sm[1][0] / sm[2][0];                                 // bc_sum / vol
    }
}

__global__ __launch_bounds__(kBlock) void k_calcps_mean(int nc, const double* __restrict__ out2,
                                                         double* __restrict__ b) {
    const double mean = out2[1];
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < nc; i += gridDim.x * kBlock) b[i] = b[i] - mean;
}

// ---------------------------------------------------------------------------
// The correction after the pressure solve (Process/CorUVW.f90:36-121).
// ---------------------------------------------------------------------------
// :36-46  U = U - Px V / Asave / rho, cells 1..NC (V, W likewise)
__global__ __launch_bounds__(kBlock) void k_cor_velocity(CalcPSMesh M, CalcPSFields F, double* __restrict__ u,
                                                          double* __restrict__ v, double* __restrict__ w) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= M.nc) return;
    const int c = M.nbc + i + 1;
    const double V = M.volume[c], As = F.asave[c], rho = F.rho_n[c];
    u[c] = u[c] - ((F.px[c] * V) / As) / rho;
    v[c] = v[c] - ((F.py[c] * V) / As) / rho;
    w[c] = w[c] - ((F.pz[c] * V) / As) / rho;
}

// :48-59 PRESSURE boundary cells take the corrected velocity of c1;
// :76-87 Flux += (PP(c2) - PP(c1)) * Aval(SidAij(1,s)) (interior) or
// * Abou(c2) (BUFFER)
__global__ __launch_bounds__(kBlock) void k_cor_faces(CalcPSMesh M, const double* __restrict__ pp,
                                                       const double* __restrict__ aval,
                                                       const double* __restrict__ abou, double* __restrict__ u,
                                                       double* __restrict__ v, double* __restrict__ w,
                                                       double* __restrict__ flux) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= M.ns) return;
    const unsigned char k = M.kind[s];
    const int nb = M.nbc;
    const int c1 = nb + M.c1[s], c2 = nb + M.c2[s];
    if (k == kPressure) {
        u[c2] = u[c1];
        v[c2] = v[c1];
        w[c2] = w[c1];
    } else if (k == kInterior) {
        flux[s] = flux[s] + (pp[c2] - pp[c1]) * aval[M.sid1[s]];
    } else if (k == kBuffer) {
        flux[s] = flux[s] + (pp[c2] - pp[c1]) * abou[c2];
    }
}

// :97-115 the mass error per unit volume of every cell, summed over its
// faces in ascending face order; per-block max |b| (:117-120)
__global__ __launch_bounds__(kBlock) void k_cor_cells(CalcPSMesh M, CalcPSFields F, double dt,
                                                       const double* __restrict__ flux, double* __restrict__ b,
                                                       double* __restrict__ part) {
    double vmax = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < M.nc; i += gridDim.x * kBlock) {
        double bb = 0.0;
        for (int e = M.cptr[i]; e < M.cptr[i + 1]; ++e) {
            const int code = M.cface[e];
            if (code & 1) bb = bb + flux[code >> 1];
            else bb = bb - flux[code >> 1];
        }
        const int c = M.nbc + i + 1;
        const double V = M.volume[c];
        bb = bb - (V / dt) * ((1.5 * F.rho_n[c] - 2.0 * F.rho_o[c]) + 0.5 * F.rho_oo[c]);
        bb = bb / V;
        b[i] = bb;
        vmax = fmax(vmax, fabs(bb));
    }
    for (int off = kWave / 2; off > 0; off >>= 1) vmax = fmax(vmax, __shfl_down(vmax, off, kWave));
    __shared__ double sm[kWavesPerBlock];
    if (threadIdx.x % kWave == 0) sm[threadIdx.x / kWave] = vmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sm[0];
        for (int q = 1; q < kWavesPerBlock; ++q) t = fmax(t, sm[q]);
        part[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(kBlock) void k_cor_total(int parts, const double* __restrict__ part,
                                                       double* __restrict__ out) {
    __shared__ double sm[kBlock];
    double t = 0.0;
    for (int q = threadIdx.x; q < parts; q += kBlock) t = fmax(t, part[q]);
    sm[threadIdx.x] = t;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sm[0];
}

template <typename T>
T* upload(std::vector<void*>& owned, const T* src, size_t count, bool device_src, hipStream_t s) {
    T* d = nullptr;
    TF_CHECK(hipMalloc(&d, sizeof(T) * (count ? count : 1)));
    owned.push_back(d);
    if (count)
        TF_CHECK(hipMemcpyAsync(d, src, sizeof(T) * count, device_src ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                s));
    return d;
}

}  // namespace

CalcPS::CalcPS(int nc, int nbc, int ns, int nonz, const int* sidec, const int* sidaij, const int* adia,
               const int* typebc, const double* const geom[7], const double* volume, bool device_ptrs, hipStream_t s)
    : nc_(nc), nbc_(nbc), ns_(ns), nonz_(nonz) {
    if (nc <= 0 || nbc < 0 || ns < 0) throw std::invalid_argument("NC must be positive, NbC and NS not negative");
    if (nonz < nc + 1) throw std::invalid_argument("NONZ must be at least NC + 1");
    // the integer arrays are checked and turned into kernel form on the host
    std::vector<int> h_sidec(3 * (size_t)ns), h_sid(2 * (size_t)ns), h_adia(nc), h_type(nbc + nc + 1);
    auto fetch = [&](void* dst, const void* src, size_t bytes) {
        if (!bytes) return;
        if (device_ptrs) TF_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
        else std::memcpy(dst, src, bytes);
    };
    fetch(h_sidec.data(), sidec, sizeof(int) * h_sidec.size());
    fetch(h_sid.data(), sidaij, sizeof(int) * h_sid.size());
    fetch(h_adia.data(), adia, sizeof(int) * h_adia.size());
    fetch(h_type.data(), typebc, sizeof(int) * h_type.size());
    std::vector<int> c1(ns), c2(ns), s1(ns, 0), s2(ns, 0), ad0(nc);
    std::vector<unsigned char> kind(ns, kNone);
    std::vector<char> covered(nonz, 0);
    for (int i = 0; i < nc; ++i) {
        if (h_adia[i] < 1 || h_adia[i] > nonz - 1) throw std::invalid_argument("Adia out of range");
        ad0[i] = h_adia[i] - 1;
        covered[ad0[i]] = 1;
    }
    std::vector<int> count(nc + 1, 0);
    for (int q = 0; q < ns; ++q) {
        c1[q] = h_sidec[3 * (size_t)q + 1];
        c2[q] = h_sidec[3 * (size_t)q + 2];
        if (c1[q] < 1 || c1[q] > nc) throw std::invalid_argument("SideC(1,s) out of range");
        if (c2[q] > nc || c2[q] < -nbc || c2[q] == 0) throw std::invalid_argument("SideC(2,s) out of range");
        if (c2[q] > 0) {
            kind[q] = kInterior;
            for (int r = 0; r < 2; ++r) {
                const int pos = h_sid[2 * (size_t)q + r];
                if (pos < 1 || pos > nonz - 1) throw std::invalid_argument("SidAij out of range");
                (r ? s2 : s1)[q] = pos - 1;
                covered[pos - 1] = 1;
            }
        } else {
            const int t = h_type[nbc + c2[q]];
            kind[q] = t == kBcBuffer ? kBuffer
                    : t == kBcInflow ? kInflow
                    : (t == kBcOutflow || t == kBcConvect) ? kDiagOut
                    : t == kBcPressure ? kPressure
                    : kNone;
        }
        count[c1[q]]++;
        if (kind[q] == kInterior) count[c2[q]]++;
    }
    // cell -> faces (every face of the cell), ascending face index;
    // entry = 2 s + (cell is c2)
    std::vector<int> cptr(nc + 1, 0), cface;
    for (int i = 0; i < nc; ++i) cptr[i + 1] = cptr[i] + count[i + 1];
    cface.resize(cptr[nc]);
    std::vector<int> fill(cptr.begin(), cptr.end() - 1);
    for (int q = 0; q < ns; ++q) {
        cface[fill[c1[q] - 1]++] = 2 * q;
        if (kind[q] == kInterior) cface[fill[c2[q] - 1]++] = 2 * q + 1;
    }
    std::vector<int> zeros;
    for (int j = 0; j < nonz; ++j)
        if (!covered[j]) zeros.push_back(j);
    n_zero_ = (int)zeros.size();

    M_.nc = nc;
    M_.nbc = nbc;
    M_.ns = ns;
    M_.c1 = upload(owned_, c1.data(), c1.size(), false, s);
    M_.c2 = upload(owned_, c2.data(), c2.size(), false, s);
    M_.sid1 = upload(owned_, s1.data(), s1.size(), false, s);
    M_.sid2 = upload(owned_, s2.data(), s2.size(), false, s);
    M_.kind = upload(owned_, kind.data(), kind.size(), false, s);
    M_.adia = upload(owned_, ad0.data(), ad0.size(), false, s);
    M_.cptr = upload(owned_, cptr.data(), cptr.size(), false, s);
    M_.cface = upload(owned_, cface.data(), cface.size(), false, s);
    zero_idx_ = upload(owned_, zeros.data(), zeros.size(), false, s);
    const double** g[7] = {&M_.sx, &M_.sy, &M_.sz, &M_.dx, &M_.dy, &M_.dz, &M_.f};
    for (int q = 0; q < 7; ++q) *g[q] = upload(owned_, geom[q], (size_t)ns, device_ptrs, s);
    M_.volume = upload(owned_, volume, (size_t)(nbc + nc + 1), device_ptrs, s);
    grid_ = std::max(1, std::min(2048, (nc + kBlock - 1) / kBlock));
    TF_CHECK(hipMalloc(&fa12_, sizeof(double) * (ns ? ns : 1)));
    TF_CHECK(hipMalloc(&part_, sizeof(double) * 3 * grid_));
    TF_CHECK(hipMalloc(&out2_, sizeof(double) * 2));
    TF_CHECK(hipStreamSynchronize(s));
}

double CalcPS::correct(double dt, const CalcPSFields& F, const double* pp, const double* aval, const double* abou,
                       double* u, double* v, double* w, double* flux, double* b, hipStream_t s) {
    if (!(dt != 0.0)) throw std::invalid_argument("dt must be non-zero");
    hipLaunchKernelGGL(k_cor_velocity, dim3((nc_ + kBlock - 1) / kBlock), dim3(kBlock), 0, s, M_, F, u, v, w);
    if (ns_ > 0)
        hipLaunchKernelGGL(k_cor_faces, dim3((ns_ + kBlock - 1) / kBlock), dim3(kBlock), 0, s, M_, pp, aval, abou, u,
                           v, w, flux);
    hipLaunchKernelGGL(k_cor_cells, dim3(grid_), dim3(kBlock), 0, s, M_, F, dt, (const double*)flux, b, part_);
    hipLaunchKernelGGL(k_cor_total, dim3(1), dim3(kBlock), 0, s, grid_, (const double*)part_, out2_);
    TF_CHECK(hipGetLastError());
    double em = 0.0;
    TF_CHECK(hipMemcpyAsync(&em, out2_, sizeof(double), hipMemcpyDeviceToHost, s));
    TF_CHECK(hipStreamSynchronize(s));
    return em;
}

double* CalcPS::staging(size_t doubles) {
    if (doubles > staging_size_) {
        for (void* p : staging_) (void)hipFree(p);
        staging_.clear();
        void* p = nullptr;
        TF_CHECK(hipMalloc(&p, sizeof(double) * doubles));
        staging_.push_back(p);
        staging_size_ = doubles;
    }
    return staging_.empty() ? nullptr : (double*)staging_[0];
}

CalcPS::~CalcPS() {
    for (void* p : owned_) (void)hipFree(p);
    (void)hipFree(fa12_);
    (void)hipFree(part_);
    (void)hipFree(out2_);
    for (void* p : staging_) (void)hipFree(p);
}

void CalcPS::assemble(int nmat, double dt, const CalcPSFields& F, double* aval, double* b, double* flux,
                      double* abou, double* errmax, double* mean, hipStream_t s) {
    if (nmat < 0) throw std::invalid_argument("Nmat must not be negative");
    if (!(dt != 0.0)) throw std::invalid_argument("dt must be non-zero");
    if (nmat == 0) {
        // no face pass (Flux and Abou keep their values): Aval is all zero
        hipLaunchKernelGGL(k_calcps_zero_all, dim3((nonz_ + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nonz_, aval);
    } else {
        if (n_zero_ > 0)
            hipLaunchKernelGGL(k_calcps_zero, dim3((n_zero_ + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n_zero_,
                               (const int*)zero_idx_, aval);
        if (ns_ > 0)
            hipLaunchKernelGGL(k_calcps_faces, dim3((ns_ + kBlock - 1) / kBlock), dim3(kBlock), 0, s, M_, F, aval,
                               flux, abou, fa12_);
    }
    hipLaunchKernelGGL(k_calcps_cells, dim3(grid_), dim3(kBlock), 0, s, M_, F, nmat, dt, (const double*)flux,
                       (const double*)fa12_, aval, b, part_);
    hipLaunchKernelGGL(k_calcps_total, dim3(1), dim3(kBlock), 0, s, grid_, (const double*)part_, out2_);
    hipLaunchKernelGGL(k_calcps_mean, dim3(grid_), dim3(kBlock), 0, s, nc_, (const double*)out2_, b);
    TF_CHECK(hipGetLastError());
    double h[2];
    TF_CHECK(hipMemcpyAsync(h, out2_, sizeof(h), hipMemcpyDeviceToHost, s));
    TF_CHECK(hipStreamSynchronize(s));
    if (errmax) *errmax = h[0];
    if (mean) *mean = h[1];
}

}  // namespace tflows
