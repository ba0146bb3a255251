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
#include <map>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "calcps.h"
#include "device_common.h"

namespace tflows {

namespace {

constexpr int kBcInflow = 1, kBcWall = 2, kBcOutflow = 3, kBcConvect = 5, kBcWallfl = 6, kBcBuffer = 11,
              kBcPressure = 12;  // allp_mod.f90:16-24

// what a face does to its cells (fixed by the face list and TypeBC)
enum : unsigned char {
    kNone = 0,      // SYMMETRY, walls: Flux = 0 (:160-161)
    kInterior = 1,  // diagonal of c1 and c2, b(c1) -= F, b(c2) += F (:86-94, :114-115)
    kBuffer = 2,    // diagonal of c1, Abou(c2), b(c1) -= F (:96-101)
    kInflow = 3,    // b(c1) -= F only (:120-133)
    kDiagOut = 4,   // OUTFLOW / CONVECT / PRESSURE: diagonal of c1, b(c1) -= F (:134-159)
    kPressure = 5,  // as kDiagOut, with the cell's own mass velocity (:150-152)
};

// what a face does in the diffusion loop of the momentum assembly
// (NewUVW.f90:362-410)
enum : unsigned char {
    kUvwNone = 0,      // SYMMETRY, PRESSURE, ...: convection only
    kUvwInterior = 1,  // the "m1" fill of both rows (:364-367)
    kUvwWallLike = 2,  // WALL, CONVECT, INFLOW, WALLFL: diagonal of c1, b(c1) (:385-392)
    kUvwOutflow = 3,   // OUTFLOW: A21 to the diagonal, A12 rhoUi(c2) to b (:396-399)
    kUvwBuffer = 4,    // BUFFER: diagonal of c1, Abou(c2) (:403-406)
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

// ---------------------------------------------------------------------------
// The momentum assembly of one velocity component (Process/NewUVW.f90:137-499).
// ---------------------------------------------------------------------------
constexpr int kFlagCN = 2, kFlagFI = 3, kInertLin = 4, kInertPar = 5, kBlendYes = 10,
              kBlendNo = 11;   // IniPar.f90:22-32
// ConvScheme's face schemes (IniPar.f90:42-49)
constexpr int kCds = 20, kQuick = 21, kLuds = 22, kMinmod = 23, kSmart = 24, kAvlSmart = 25, kSuperbee = 26,
              kGamma = 27;

struct UvwGeom {        // Si, Sj, Sk, Di, Dj, Dk of the component (Processor.f90:314-340)
    const double *si, *sj, *sk, *di, *dj, *dk;
};

// What the face loops read of one cell.
struct UvwCell {
    double rho, ru;     // rho%n and rhoUi = rho%n Ui%n (:138)
    double g[7];        // the seven gradients
    int mat;            // material - 1
    int idx;            // NbC + cell
};

template <int NG>
__device__ __forceinline__ UvwCell uvw_cell(const UvwMesh& U, const UvwFields& F, int idx) {
    UvwCell v;
    v.rho = F.rho_n[idx];
    v.ru = v.rho * F.ui_n[idx];
#pragma unroll
    for (int k = 0; k < 7; ++k) v.g[k] = k < NG ? F.g[k][idx] : 0.0;
    v.mat = U.material[idx] - 1;
    v.idx = idx;
    return v;
}

// The quantities of face s that the reference's two face loops compute
// (:211-238, :280-360), in its operation order; a and b are the cells
// SideC(1,s) and SideC(2,s).
struct UvwFace {
    double conv;        // Flux_u rhoUiS (:220-223)
    double up;          // Flux_u rhoUi(upwind cell) (:228-236)
    bool upwind;        // BLEND_TEM /= NO on either side
    double fex, fim;    // the cross-diffusion parts (:312-330)
    double a12, a21, d12, d21;
};

// MAX / MIN as plain comparisons (the first argument is kept unless the
// other is strictly beyond it), the same in the C restatement.
__device__ __forceinline__ double cs_max(double a, double b) { return b > a ? b : a; }
__device__ __forceinline__ double cs_min(double a, double b) { return b < a ? b : a; }

// Process/ConvScheme.f90:55-135: the face value of Phi = rhoUi at face s for
// scheme `blenda` (YES: the central value phis as it is).  The upwind cell C
// and the downwind cell D follow the sign of the mass flux (:55-65); the
// gradient of C is dotted with Di, Dj, Dk in the order of the call
// (NewUVW.f90:214: dUidi, dUidj, dUidk; CalcZmix.f90:164: dZdx, dZdy, dZdz).
__device__ __forceinline__ double conv_scheme(double phis, int s, const UvwCell& a, const UvwCell& b,
                                              const UvwGeom& G, const CalcPSMesh& M, const UvwFields& F,
                                              int blenda) {
    if (blenda == kBlendYes) return phis;
    const bool pos = (F.flux ? F.flux[s] : F.flux_u[s]) > 0.0;
    const double fj = pos ? 1.0 - M.f[s] : M.f[s];
    const UvwCell& Cc = pos ? a : b;
    const UvwCell& Dc = pos ? b : a;
    const double sign = pos ? 1.0 : -1.0;
    const double gd = Cc.g[0] * G.di[s] + Cc.g[1] * G.dj[s] + Cc.g[2] * G.dk[s];
    const double PhiUstar = pos ? Dc.ru - 2.0 * gd : Dc.ru + 2.0 * gd;                 // :67-75
    const double PhiU = cs_max(F.phimin[Cc.idx], cs_min(PhiUstar, F.phimax[Cc.idx]));   // :77
    double rj = (Cc.ru - PhiU) / (Dc.ru - Cc.ru + 1.0e-16);                             // :79
    const double gD = 0.5 * fj * (1.0 + fj), gU = 0.5 * fj * (1.0 - fj);                // :81-82
    double Phij = 0.0, alfa, beta1, beta2;
    switch (blenda) {                                                                    // :84-114
    case kCds: Phij = fj; break;
    case kQuick:
        rj = (Cc.ru - PhiU) / (Dc.ru - Cc.ru + 1.0e-12);
        alfa = 0.0;
        Phij = (gD - alfa) + (gU + alfa) * rj;
        break;
    case kLuds:
        alfa = 0.5 * fj * (1 + fj);
        Phij = (gD - alfa) + (gU + alfa) * rj;
        break;
    case kMinmod: Phij = fj * cs_max(0.0, cs_min(rj, 1.0)); break;
    case kSmart:
        beta1 = 3.0;
        beta2 = 1.0;
        Phij = cs_max(0.0, cs_min(cs_min((beta1 - 1.0) * rj, gD + gU * rj), beta2));
        break;
    case kAvlSmart:
        beta1 = 1.0 + fj * (2.0 + fj);
        beta2 = fj * (2.0 - fj);
        Phij = cs_max(0.0, cs_min(cs_min((beta1 - 1.0) * rj, gD + gU * rj), beta2));
        break;
    case kSuperbee: Phij = 0.5 * cs_max(cs_max(0.0, cs_min(2.0 * rj, 1.0)), cs_min(rj, 2.0)); break;
    default: break;                     // GAMMA: the value is set below
    }
    if (blenda != kGamma) return Cc.ru + Phij * sign * (b.ru - a.ru);                    // :118
    const double Beta = 0.1;                                                             // :120-134
    const double PhiUg = pos ? 1.0 - (Dc.ru - Cc.ru) / (2.0 * gd) : 1.0 + (Dc.ru - Cc.ru) / (2.0 * gd);
    const double GammaC = PhiUg / Beta;
    if (PhiUg < Beta && PhiUg > 0.0)
        return (1.0 - GammaC * (1.0 - M.f[s])) * Cc.ru + GammaC * (1.0 - M.f[s]) * Dc.ru;
    if (PhiUg < 1.0 && PhiUg >= Beta) return M.f[s] * Cc.ru + (1.0 - M.f[s]) * Dc.ru;
    return Cc.ru;
}

// TVD: some face takes a ConvScheme scheme other than YES (without one,
// every face keeps the central value and the kernel needs no more registers
// than that).
template <bool SCALAR, bool TVD>
__device__ __forceinline__ UvwFace uvw_face(const CalcPSMesh& M, const UvwMesh& U, const UvwGeom& G,
                                            const UvwParams& P, const UvwFields& F, int s, const UvwCell& a,
                                            const UvwCell& b) {
    UvwFace q;
    const double Fu = F.flux_u[s];
    const double fs = M.f[s];
    double rhoUiS = fs * a.ru + (1.0 - fs) * b.ru;                                   // :211
    if (TVD && (P.blend[a.mat] != kBlendNo || P.blend[b.mat] != kBlendNo))            // :213-216
        rhoUiS = conv_scheme(rhoUiS, s, a, b, G, M, F, max(P.blend[a.mat], P.blend[b.mat]));
    q.conv = Fu * rhoUiS;
    q.upwind = P.blend_tem[a.mat] != kBlendNo || P.blend_tem[b.mat] != kBlendNo;
    q.up = Fu * (Fu < 0 ? b.ru : a.ru);
    const double ff = U.ff[s], gg = 1.0 - U.ff[s];
    constexpr int NG = SCALAR ? 3 : 7;
    double gS[7];
#pragma unroll
    for (int k = 0; k < NG; ++k) gS[k] = ff * a.g[k] + gg * b.g[k];                  // :300-308
    const double Si = G.si[s], Sj = G.sj[s], Sk = G.sk[s];
    const double visc = P.visc_rho ? P.visc * (ff * a.rho + gg * b.rho) : P.visc;  // VISeff (:285-295)
    if (SCALAR)                                                                      // CalcZmix.f90:250
        q.fex = visc * ((gS[0] * Si + gS[1] * Sj) + gS[2] * Sk);
    else                                                                             // :312-315
        q.fex = visc * ((((4.0 / 3.0) * gS[0] * Si + (gS[1] + gS[3]) * Sj) + (gS[2] + gS[4]) * Sk) -
                        (2.0 / 3.0) * (gS[5] + gS[6]) * Si);
    const double A0 = visc * U.scoef[s];                                             // :323
    q.fim = A0 * ((gS[0] * G.di[s] + gS[1] * G.dj[s]) + gS[2] * G.dk[s]);            // :325-327
    q.a12 = q.a21 = q.d12 = q.d21 = 0.0;
    if (P.diffus == kFlagFI) {                                                       // :351-354
        q.d12 = 0.0 + A0 / b.rho;
        q.d21 = 0.0 + A0 / a.rho;
    }
    if (P.blend[a.mat] != kBlendNo || P.blend[b.mat] != kBlendNo) {                 // :356-359
        q.a12 = 0.0 - (Fu < 0.0 ? Fu : 0.0);
        q.a21 = 0.0 + (Fu > 0.0 ? Fu : 0.0);
    }
    return q;
}

template <int NG>
__device__ __forceinline__ UvwCell pick(bool first, const UvwCell& x, const UvwCell& y) {
    UvwCell v;
    v.rho = first ? x.rho : y.rho;
    v.ru = first ? x.ru : y.ru;
#pragma unroll
    for (int k = 0; k < 7; ++k) v.g[k] = k < NG ? (first ? x.g[k] : y.g[k]) : 0.0;
    v.mat = first ? x.mat : y.mat;
    v.idx = first ? x.idx : y.idx;
    return v;
}

// CalMinMax (Process/CalMinMax.f90:21-37) of rhoUi = rho%n Ui%n: an interior
// cell over itself and the cells across its interior and BUFFER faces, in
// ascending face order (the order in which the reference's face loop
// reaches it); a boundary cell over itself and, on a BUFFER face, the face's
// cell c1 (a boundary cell has one face, checked on the host).
__global__ __launch_bounds__(kBlock) void k_uvw_minmax(CalcPSMesh M, UvwMesh U, UvwFields F,
                                                        double* __restrict__ pmax, double* __restrict__ pmin) {
    const int nb = M.nbc;
    const int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < M.nc; i += stride) {
        const int ic = nb + i + 1;
        const double v = F.rho_n[ic] * F.ui_n[ic];
        double mx = v, mn = v;
        for (int e = M.cptr[i]; e < M.cptr[i + 1]; ++e) {
            const int2 ie = U.inc[e];
            const int o = ie.y;
            if (o > 0 || U.ukind[ie.x >> 1] == kUvwBuffer) {
                const double w = F.rho_n[nb + o] * F.ui_n[nb + o];
                mx = cs_max(mx, w);
                mn = cs_min(mn, w);
            }
        }
        pmax[ic] = mx;
        pmin[ic] = mn;
    }
    for (int s = blockIdx.x * kBlock + threadIdx.x; s < M.ns; s += stride) {
        const int c2 = M.c2[s];
        if (c2 >= 0) continue;
        const int i2 = nb + c2, i1 = nb + M.c1[s];
        const double v = F.rho_n[i2] * F.ui_n[i2];
        double mx = v, mn = v;
        if (U.ukind[s] == kUvwBuffer) {
            const double w = F.rho_n[i1] * F.ui_n[i1];
            mx = cs_max(mx, w);
            mn = cs_min(mn, w);
        }
        pmax[i2] = mx;
        pmin[i2] = mn;
    }
}

// The CFL and Peclet monitor of CorUVW.f90:124-147: for every material m,
// the largest |dt Flux / (RhoS Scoef |D|^2)| and |Flux / Scoef / (VISc +
// TINY)| over the interior and BUFFER faces with a cell of material m on
// either side, RhoS = f rho%n(c1) + (1 - f) rho%n(c2) (the geometric f, as
// the reference writes it).  A maximum is exact, so the block partials and
// their total give the reference's values whatever the order.  Partials:
// part[(2 m) G + b] the CFL, part[(2 m + 1) G + b] the Peclet number.
__global__ __launch_bounds__(kBlock) void k_cfl_peclet(CalcPSMesh M, UvwMesh U, int nmat, double dt, double visc,
                                                        const double* __restrict__ rho_n,
                                                        const double* __restrict__ flux, double* __restrict__ part) {
    double cfl[kMaxMaterials], pe[kMaxMaterials];
#pragma unroll
    for (int m = 0; m < kMaxMaterials; ++m) cfl[m] = pe[m] = 0.0;
    const int nb = M.nbc;
    for (int s = blockIdx.x * kBlock + threadIdx.x; s < M.ns; s += gridDim.x * kBlock) {
        const int c2 = M.c2[s];
        if (!(c2 > 0 || U.ukind[s] == kUvwBuffer)) continue;
        const int i1 = nb + M.c1[s], i2 = nb + c2;
        const double fs = M.f[s];
        const double rhos = fs * rho_n[i1] + (1.0 - fs) * rho_n[i2];
        const double dx = M.dx[s], dy = M.dy[s], dz = M.dz[s];
        const double sc = U.scoef[s], fl = flux[s];
        const double cfls = fabs(dt * fl / (rhos * sc * (dx * dx + dy * dy + dz * dz)));
        const double pes = fabs(fl / sc / (visc + kTiny));
        const int m1 = U.material[i1] - 1, m2 = U.material[i2] - 1;
#pragma unroll
        for (int m = 0; m < kMaxMaterials; ++m) {
            if (m < nmat && (m == m1 || m == m2)) {
                cfl[m] = cs_max(cfl[m], cfls);
                pe[m] = cs_max(pe[m], pes);
            }
        }
    }
    __shared__ double sm[2 * kMaxMaterials][kWavesPerBlock];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int m = 0; m < kMaxMaterials; ++m) {
        double a = cfl[m], b = pe[m];
        for (int off = kWave / 2; off > 0; off >>= 1) {
            a = cs_max(a, __shfl_xor(a, off, kWave));
            b = cs_max(b, __shfl_xor(b, off, kWave));
        }
        if (lane == 0) {
            sm[2 * m][w] = a;
            sm[2 * m + 1][w] = b;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < 2 * nmat) {
        double v = sm[threadIdx.x][0];
#pragma unroll
        for (int k = 1; k < kWavesPerBlock; ++k) v = cs_max(v, sm[threadIdx.x][k]);
        part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = v;
    }
}

// out[j] = the largest of part[j G .. j G + G - 1], j < nvals (one block)
__global__ __launch_bounds__(kBlock) void k_cfl_total(int nvals, int parts, const double* __restrict__ part,
                                                       double* __restrict__ out) {
    __shared__ double sm[kBlock];
    for (int j = 0; j < nvals; ++j) {
        double t = 0.0;
        for (int q = threadIdx.x; q < parts; q += kBlock) t = cs_max(t, part[(size_t)j * parts + q]);
        sm[threadIdx.x] = t;
        __syncthreads();
        for (int h = kBlock / 2; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) sm[threadIdx.x] = cs_max(sm[threadIdx.x], sm[threadIdx.x + h]);
            __syncthreads();
        }
        if (threadIdx.x == 0) out[j] = sm[0];
        __syncthreads();
    }
}

// One thread per cell: the two face loops walked over the cell's faces in
// ascending face order (the order in which the reference's loops reach the
// cell), then the cell loops (:252-257, :434-499).  SCALAR is the
// mixture-fraction equation (Process/CalcZmix.f90:100-387): three gradients,
// the flux VISc grad(Z) . S, the two-level inertia product in its order, no
// pressure term and no Asave.  The cell's own values are read once and
// kept in LDS for the block; per face the face's values are read, and the
// other cell's from LDS when it belongs to the same block (a tile is four
// runs of 64 cells, listed in U.wbase or consecutive), else from memory:
// the same bits either way.  The diagonal and b stay in registers.  The thread of cell c also writes the
// off-diagonals of row c, (0 - A12) - D12 at SidAij(1,s) and (0 - A21) - D21
// at SidAij(2,s) (:364-367), from the same A12, A21, D12, D21 it adds to the
// diagonal, and Abou of its BUFFER faces (:406); where several faces write
// one position (SHARED) each off-diagonal is accumulated, in face order, on
// a zeroed Aval.
constexpr int kUvwStaged = 6;   // off-diagonals per row staged in LDS (a hexahedron's faces)
template <bool SHARED, bool SCALAR, bool TVD>
__global__ __launch_bounds__(kBlock) void k_uvw_cells(CalcPSMesh M, UvwMesh U, UvwGeom G, UvwParams P,
                                                       UvwFields F, double* __restrict__ aval,
                                                       double* __restrict__ b, double* __restrict__ abou,
                                                       double* __restrict__ asave, double* __restrict__ rhoui) {
    const int nb = M.nbc;
    const int stride = gridDim.x * kBlock;
    const int t0 = blockIdx.x * kBlock + threadIdx.x;
    for (int q = t0; q <= nb; q += stride) rhoui[q] = F.rho_n[q] * F.ui_n[q];       // boundary cells and 0
    const bool diff = P.diffus == kFlagCN || P.diffus == kFlagFI;
    __shared__ double st_val[SHARED ? 1 : kUvwStaged * kBlock];
    __shared__ int st_pos[SHARED ? 1 : kUvwStaged * kBlock];
    // 256-cell tiles; each XCD walks one contiguous eighth of them (in U.wbase or
    // U.tiles order when the mesh has one)
    // the values of the block's own cells, for the faces whose other cell is
    // one of them
    constexpr int NG = SCALAR ? 3 : 7;
    __shared__ double sh_rho[kBlock], sh_ru[kBlock], sh_g[NG][kBlock];
    __shared__ int sh_mat[kBlock];
    const TileSpan sp = block_span_strided((M.nc + kBlock - 1) / kBlock);
    for (int t = sp.first; t < sp.end; t += sp.step) {
        // the first cell of each wavefront of the tile; wavefront w holds 64 consecutive cells
        int wb[kWavesPerBlock];
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w)
            wb[w] = U.wbase ? U.wbase[t * kWavesPerBlock + w] : (U.tiles ? U.tiles[t] : t) * kBlock + w * kWave;
        const int i = wb[threadIdx.x / kWave] + (int)(threadIdx.x & (kWave - 1));
        const bool live = i < M.nc;
        const int c = i + 1, ic = nb + c;
        UvwCell own{};
        if (live) {
            own = uvw_cell<NG>(U, F, ic);
            sh_rho[threadIdx.x] = own.rho;
            sh_ru[threadIdx.x] = own.ru;
#pragma unroll
            for (int k = 0; k < NG; ++k) sh_g[k][threadIdx.x] = own.g[k];
            sh_mat[threadIdx.x] = own.mat;
        }
        __syncthreads();
        if (live) {
            const int e0 = M.cptr[i], e1 = M.cptr[i + 1];
            double Cs = 0.0, Xc = 0.0, Xd = 0.0, diag = 0.0;
            // the row's off-diagonals wait in LDS until the face loop is done: a
            // store among the loop's loads would make every later load wait for it
            int staged = 0;
            auto stage = [&](int pos, double v) {
                if (staged < kUvwStaged) {
                    st_val[staged * kBlock + threadIdx.x] = v;
                    st_pos[staged * kBlock + threadIdx.x] = pos;
                    ++staged;
                } else {
                    aval[pos] = v;
                }
            };
            for (int e = e0; e < e1; ++e) {
                const int2 ie = U.inc[e];
                const int s = ie.x >> 1;
                const bool second = ie.x & 1;                                            // the cell is c2
                const int other = ie.y;
                // a cell of this block: its values from LDS (the same bits)
                int slot = -1;
#pragma unroll
                for (int w = 0; w < kWavesPerBlock; ++w) {
                    const unsigned d = (unsigned)(other - 1 - wb[w]);
                    if (d < (unsigned)kWave) slot = w * kWave + (int)d;
                }
                UvwCell oth;
                if (slot >= 0) {
                    oth.rho = sh_rho[slot];
                    oth.ru = sh_ru[slot];
#pragma unroll
                    for (int k = 0; k < 7; ++k) oth.g[k] = k < NG ? sh_g[k < NG ? k : 0][slot] : 0.0;
                    oth.mat = sh_mat[slot];
                    oth.idx = nb + other;
                } else {
                    oth = uvw_cell<NG>(U, F, nb + other);
                }
                const UvwFace f =
                    uvw_face<SCALAR, TVD>(M, U, G, P, F, s, pick<NG>(!second, own, oth), pick<NG>(!second, oth, own));
                if (!second) {
                    Cs = Cs - f.conv;
                    if (f.upwind) Xc = Xc - f.up;
                    Xd = Xd + f.fex - f.fim;                                             // :338
                } else {
                    Cs = Cs + f.conv;
                    if (f.upwind) Xc = Xc + f.up;
                    Xd = Xd - f.fex + f.fim;                                             // :340
                }
                if (!diff) continue;
                const unsigned char k = U.ukind[s];
                if (second) {                                                            // :366-367
                    diag = diag + f.a12 + f.d12;
                    const int pos = M.sid2[s];
                    if (SHARED) aval[pos] = aval[pos] - f.a21 - f.d21;
                    else stage(pos, 0.0 - f.a21 - f.d21);
                } else if (k == kUvwInterior) {                                          // :364-365, :367
                    diag = diag + f.a21 + f.d21;
                    const int pos = M.sid1[s];
                    if (SHARED) aval[pos] = aval[pos] - f.a12 - f.d12;
                    else stage(pos, 0.0 - f.a12 - f.d12);
                } else if (k == kUvwWallLike || k == kUvwBuffer) {                       // :391, :405
                    diag = diag + f.a21 + f.d21;
                    if (k == kUvwBuffer) abou[nb + other] = -f.a12 - f.d12;             // Abou(-NbC:-1)
                } else if (k == kUvwOutflow) {
                    diag = diag + f.a21;
                }
            }
            double bb = 0.0;
            if (P.convec == kFlagFI) bb = bb + P.urfc[own.mat] * (Cs - Xc);              // :252-257
            if (diff) {
                // b(c1) of WALL-like and OUTFLOW faces (A12 and D12 of :356-359),
                // after the convection term
                for (int e = e0; e < e1; ++e) {
                    const int2 ie = U.inc[e];
                    if (ie.x & 1) continue;
                    const int s = ie.x >> 1;
                    const unsigned char k = U.ukind[s];
                    if (k != kUvwWallLike && k != kUvwOutflow) continue;
                    const int i2 = nb + ie.y;
                    const double rho2 = F.rho_n[i2], ru2 = rho2 * F.ui_n[i2];
                    const double Fu = F.flux_u[s];
                    const int m2 = U.material[i2] - 1;
                    const double a12 = (P.blend[own.mat] != kBlendNo || P.blend[m2] != kBlendNo)
                                           ? 0.0 - (Fu < 0.0 ? Fu : 0.0) : 0.0;
                    if (k == kUvwWallLike) {
                        const double ff = U.ff[s];                                       // VISeff (:285-295)
                        const double visc = P.visc_rho ? P.visc * (ff * own.rho + (1.0 - ff) * rho2) : P.visc;
                        const double d12 = P.diffus == kFlagFI ? 0.0 + (visc * U.scoef[s]) / rho2 : 0.0;
                        bb = bb + (a12 + d12) * ru2;                                     // :392
                    } else {
                        bb = bb + a12 * ru2;                                             // :399
                    }
                }
            }
            if (P.cross == kFlagFI) bb = bb + Xd;                                        // :434-438
            const double V = M.volume[ic];
            if (P.inert == kInertLin) {                                                  // :447-450
                diag = diag + V / P.dt;
                if (SCALAR)                                                              // CalcZmix.f90:371
                    bb = bb + V / P.dt * F.ui_o[ic] * F.rho_o[ic] + F.source[ic] * V;
                else
                    bb = bb + V / P.dt * F.rho_o[ic] * F.ui_o[ic] + F.source[ic] * V;
            }
            if (P.inert == kInertPar) {                                                  // :453-456
                diag = diag + 1.5 * V / P.dt;
                bb = bb + 2.0 * V / P.dt * F.rho_o[ic] * F.ui_o[ic] - 0.5 * V / P.dt * F.rho_oo[ic] * F.ui_oo[ic] +
                     F.source[ic] * V;
            }
            if (!SCALAR) bb = bb - F.pi[ic] * V;                                         // :484
            if (!SCALAR && P.var == 1 && asave) asave[ic] = diag;                        // :497
            bb = bb + diag * (1.0 - P.urf) * own.ru / P.urf;                             // :498
            // the cell's stores, after all its loads
            for (int j = 0; j < staged; ++j) aval[st_pos[j * kBlock + threadIdx.x]] = st_val[j * kBlock + threadIdx.x];
            rhoui[ic] = own.ru;
            aval[M.adia[i]] = diag / P.urf;                                              // :499
            b[i] = bb;
        }
        __syncthreads();   // the next tile's cells overwrite the LDS values
    }
}

// Zmix%n(1:NC) = rhoZ / rho%n, limited to [0, 1] (CalcZmix.f90:401-405)
__global__ __launch_bounds__(kBlock) void k_zmix_update(int nc, int nb, const double* __restrict__ rhoz,
                                                         const double* __restrict__ rho_n,
                                                         double* __restrict__ z_n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nc) return;
    double z = rhoz[nb + i + 1] / rho_n[nb + i + 1];
    z = z < 0.0 ? 0.0 : z;
    z = z > 1.0 ? 1.0 : z;
    z_n[nb + i + 1] = z;
}

// Ui%n(1:NC) = rhoUi(1:NC) / rho%n(1:NC) (:526)
__global__ __launch_bounds__(kBlock) void k_uvw_update(int nc, int nb, const double* __restrict__ rhoui,
                                                        const double* __restrict__ rho_n,
                                                        double* __restrict__ ui_n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < nc) ui_n[nb + i + 1] = rhoui[nb + i + 1] / rho_n[nb + i + 1];
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

// The line and plane offsets of a naturally numbered box: interior faces
// joining cells 1, nx and nx * ny apart (the three most common offsets of a
// sample of the interior faces), planes whole, NC a whole number of planes.
static bool box_offsets(int nc, const std::vector<int>& c1, const std::vector<int>& c2,
                        const std::vector<unsigned char>& kind, int& line, int& plane) {
    std::map<int, int64_t> hist;
    const size_t ns = c1.size(), step = std::max<size_t>(1, ns / 200000);
    for (size_t q = 0; q < ns; q += step)
        if (kind[q] == kInterior) hist[std::abs(c2[q] - c1[q])]++;
    std::vector<std::pair<int64_t, int>> top;
    for (auto& h : hist) top.push_back({h.second, h.first});
    if (top.size() < 3) return false;
    std::sort(top.rbegin(), top.rend());
    int o[3] = {top[0].second, top[1].second, top[2].second};
    std::sort(o, o + 3);
    line = o[1];
    plane = o[2];
    return o[0] == 1 && line >= 2 && plane % line == 0 && nc % plane == 0;
}

// The order in which the cell kernels visit their 256-cell tiles on a box
// whose planes are a whole number of tiles: brick by brick, 8 tile rows (16
// lines of 128 cells) of 32 consecutive planes, bricks in plane-block, then
// row-block order.  The XCD-strided walk gives each XCD 256 consecutive
// entries at a time, one brick, so the cells in flight on an XCD form a
// compact window.  Otherwise empty (tiles in order).
static std::vector<int> brick_tile_order(int nc, int line, int plane) {
    std::vector<int> none;
    if (plane % kBlock != 0) return none;
    const int rows = plane / kBlock, planes = nc / plane;   // tile rows per plane
    const int BY = std::min(rows, 8), BZ = std::min(planes, 32);
    if (rows % BY != 0) return none;
    std::vector<int> order;
    order.reserve((size_t)rows * planes);
    for (int z0 = 0; z0 < planes; z0 += BZ)
        for (int y0 = 0; y0 < rows; y0 += BY)
            for (int z = z0; z < std::min(planes, z0 + BZ); ++z)
                for (int y = y0; y < y0 + BY; ++y) order.push_back(z * rows + y);
    return order;
}

// Cells of each wavefront when a block is a small box: on a box whose lines
// are whole wavefronts, with an even number of lines and of planes, tile t
// is 64 cells of 2 lines of 2 planes, its wavefront w (w = 0..3) the 64
// cells of line (w & 1) of plane (w >> 1), and wb[4 t + w] the first of
// them.  Four of a cell's six neighbours are then in its own block (the
// cell kernel reads those from LDS); tiles go in bricks of 16 lines by 32
// planes (whole lines), one brick per XCD window as above.  Otherwise empty.
static std::vector<int> wave_block_bases(int nc, int line, int plane) {
    std::vector<int> none;
    const int rows = plane / line, planes = nc / plane;
    if (line % kWave != 0 || rows % 2 != 0 || planes % 2 != 0) return none;
    const int xs = line / kWave, BY = std::min(rows, 16), BZ = std::min(planes, 32);
    std::vector<int> wb;
    wb.reserve((size_t)nc / kWave);
    for (int z0 = 0; z0 < planes; z0 += BZ)
        for (int y0 = 0; y0 < rows; y0 += BY)
            for (int z = z0; z < std::min(planes, z0 + BZ); z += 2)
                for (int y = y0; y < std::min(rows, y0 + BY); y += 2)
                    for (int x = 0; x < xs; ++x)
                        for (int w = 0; w < kWavesPerBlock; ++w)
                            wb.push_back((z + (w >> 1)) * plane + (y + (w & 1)) * line + x * kWave);
    // every cell exactly once
    if (wb.size() * kWave != (size_t)nc) return none;
    std::vector<char> seen((size_t)nc, 0);
    for (int base : wb)
        for (int l = 0; l < kWave; ++l) {
            const int i = base + l;
            if (i < 0 || i >= nc || seen[i]) return none;
            seen[i] = 1;
        }
    return wb;
}

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
    std::vector<unsigned char> kind(ns, kNone), ukind(ns, kUvwNone);
    std::vector<int> bfaces(nbc + 1, 0);     // faces of each boundary cell
    std::vector<char> covered(nonz, 0);
    for (int i = 0; i < nc; ++i) {
        if (h_adia[i] < 1 || h_adia[i] > nonz - 1) throw std::invalid_argument("Adia out of range");
        ad0[i] = h_adia[i] - 1;
        if (covered[ad0[i]]) throw std::invalid_argument("two cells share a diagonal position");
        covered[ad0[i]] = 2;
    }
    std::vector<int> count(nc + 1, 0);
    for (int q = 0; q < ns; ++q) {
        c1[q] = h_sidec[3 * (size_t)q + 1];
        c2[q] = h_sidec[3 * (size_t)q + 2];
        if (c1[q] < 1 || c1[q] > nc) throw std::invalid_argument("SideC(1,s) out of range");
        if (c2[q] > nc || c2[q] < -nbc || c2[q] == 0) throw std::invalid_argument("SideC(2,s) out of range");
        if (c2[q] > 0) {
            kind[q] = kInterior;
            ukind[q] = kUvwInterior;
            for (int r = 0; r < 2; ++r) {
                const int pos = h_sid[2 * (size_t)q + r];
                if (pos < 1 || pos > nonz - 1) throw std::invalid_argument("SidAij out of range");
                if (covered[pos - 1] == 2) throw std::invalid_argument("SidAij names a diagonal position");
                (r ? s2 : s1)[q] = pos - 1;
                if (covered[pos - 1]) offdiag_shared_ = true;
                covered[pos - 1] = 1;
            }
        } else {
            const int t = h_type[nbc + c2[q]];
            kind[q] = t == kBcBuffer ? kBuffer
                    : t == kBcInflow ? kInflow
                    : (t == kBcOutflow || t == kBcConvect) ? kDiagOut
                    : t == kBcPressure ? kPressure
                    : kNone;
            ukind[q] = (t == kBcWall || t == kBcConvect || t == kBcInflow || t == kBcWallfl) ? kUvwWallLike
                     : t == kBcOutflow ? kUvwOutflow
                     : t == kBcBuffer ? kUvwBuffer
                     : kUvwNone;
            if (bfaces[-c2[q]]++ > 0) bcell_single_ = false;
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
    U_.ukind = upload(owned_, ukind.data(), ukind.size(), false, s);
    ukind_host_ = ukind;
    c1_host_ = c1;
    c2_host_ = c2;
    M_.adia = upload(owned_, ad0.data(), ad0.size(), false, s);
    M_.cptr = upload(owned_, cptr.data(), cptr.size(), false, s);
    M_.cface = upload(owned_, cface.data(), cface.size(), false, s);
    {
        std::vector<int2> inc(cface.size());
        for (size_t e = 0; e < cface.size(); ++e) {
            const int q = cface[e] >> 1;
            inc[e] = make_int2(cface[e], (cface[e] & 1) ? c1[q] : c2[q]);
        }
        U_.inc = upload(owned_, inc.data(), inc.size(), false, s);
    }
    zero_idx_ = upload(owned_, zeros.data(), zeros.size(), false, s);
    {
        // TFLOWS_UVW_ORDER: 0 tiles in order, 1 at most the brick order, 2 (default) the best that applies
        const char* env = std::getenv("TFLOWS_UVW_ORDER");
        const int want = env ? std::atoi(env) : 2;
        int line = 0, plane = 0;
        if (want > 0 && box_offsets(nc, c1, c2, kind, line, plane)) {
            std::vector<int> wb = want >= 2 ? wave_block_bases(nc, line, plane) : std::vector<int>();
            if (!wb.empty()) {
                U_.wbase = upload(owned_, wb.data(), wb.size(), false, s);
                visit_order_ = 2;
            } else {
                std::vector<int> order = brick_tile_order(nc, line, plane);
                if (!order.empty()) {
                    U_.tiles = upload(owned_, order.data(), order.size(), false, s);
                    visit_order_ = 1;
                }
            }
        }
    }
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
    (void)hipFree(minmax_);
    (void)hipFree(cflpart_);
    (void)hipFree(cflout_);
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

void CalcPS::uvw_setup(const double* scoef, const double* ff, const int* material, int nmaterials, bool device_ptrs,
                       hipStream_t s) {
    if (nmaterials < 1 || nmaterials > kMaxMaterials)
        throw std::invalid_argument("the momentum assembly takes 1 to 8 materials");
    if (ns_ > 0 && (!scoef || !ff)) throw std::invalid_argument("null Scoef or fF");
    if (!material) throw std::invalid_argument("null material");
    const size_t ncell = (size_t)nbc_ + nc_ + 1;
    std::vector<int> mat(ncell);
    if (device_ptrs) TF_CHECK(hipMemcpy(mat.data(), material, sizeof(int) * ncell, hipMemcpyDeviceToHost));
    else std::memcpy(mat.data(), material, sizeof(int) * ncell);
    // cells that no face reaches (cell 0) are never read; the rest must
    // name a material
    std::vector<char> used(ncell, 0);
    for (int q = 0; q < ns_; ++q) used[nbc_ + c1_host_[q]] = used[nbc_ + c2_host_[q]] = 1;
    for (int c = 1; c <= nc_; ++c) used[nbc_ + c] = 1;
    for (size_t c = 0; c < ncell; ++c) {
        if (!used[c]) mat[c] = 1;
        else if (mat[c] < 1 || mat[c] > nmaterials) throw std::invalid_argument("material out of range");
    }
    unsigned pairs[kMaxMaterials] = {};
    for (int q = 0; q < ns_; ++q) pairs[mat[nbc_ + c1_host_[q]] - 1] |= 1u << (mat[nbc_ + c2_host_[q]] - 1);
    if (!U_.ff) {
        double* d = nullptr;
        TF_CHECK(hipMalloc(&d, sizeof(double) * 2 * (ns_ ? ns_ : 1)));
        owned_.push_back(d);
        U_.scoef = d;
        U_.ff = d + ns_;
        int* m = nullptr;
        TF_CHECK(hipMalloc(&m, sizeof(int) * ncell));
        owned_.push_back(m);
        U_.material = m;
    }
    const hipMemcpyKind kind = device_ptrs ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (ns_ > 0) {
        TF_CHECK(hipMemcpyAsync((double*)U_.scoef, scoef, sizeof(double) * ns_, kind, s));
        TF_CHECK(hipMemcpyAsync((double*)U_.ff, ff, sizeof(double) * ns_, kind, s));
    }
    TF_CHECK(hipMemcpyAsync((int*)U_.material, mat.data(), sizeof(int) * ncell, hipMemcpyHostToDevice, s));
    TF_CHECK(hipStreamSynchronize(s));
    nmaterials_ = nmaterials;
    std::memcpy(pairs_, pairs, sizeof(pairs));
}

void CalcPS::uvw_assemble(const UvwParams& P, const UvwFields& F, double* aval, double* b, double* abou,
                          double* asave, double* rhoui, hipStream_t s) {
    if (P.var < 1 || P.var > 3) throw std::invalid_argument("var must be 1, 2 or 3");
    assemble_equation(P, F, aval, b, abou, asave, rhoui, false, s);
}

void CalcPS::assemble_equation(const UvwParams& P, const UvwFields& F, double* aval, double* b, double* abou,
                               double* asave, double* rhoui, bool scalar, hipStream_t s) {
    if (!uvw_ready()) throw std::invalid_argument("the momentum geometry is not set (uvw_setup)");
    if (!(P.dt != 0.0) && (P.inert == kInertLin || P.inert == kInertPar))
        throw std::invalid_argument("dt must be non-zero");
    if (!(P.urf != 0.0)) throw std::invalid_argument("URF must be non-zero");
    // ConvScheme (Process/ConvScheme.f90) runs where max(BLEND) of a face's
    // materials is not NO: YES keeps the central value, CDS .. GAMMA form
    // the face value from PHImax / PHImin of CalMinMax.  YES next to NO (max
    // = NO, which leaves Phij undefined in the reference) and other codes
    // are refused.
    bool tvd = false;
    for (int a = 0; a < nmaterials_; ++a)
        for (int c = 0; c < nmaterials_; ++c)
            if (pairs_[a] >> c & 1u) {
                const int ba = P.blend[a], bc = P.blend[c];
                if (ba == kBlendNo && bc == kBlendNo) continue;
                const int bl = std::max(ba, bc);
                if (bl != kBlendYes && (bl < kCds || bl > kGamma))
                    throw std::invalid_argument(scalar ? "BLEND_TEM of a face is neither YES nor a ConvScheme "
                                                         "scheme (20-27), or YES next to NO"
                                                       : "BLEND of a face is neither YES nor a ConvScheme scheme "
                                                         "(20-27), or YES next to NO");
                tvd = tvd || bl != kBlendYes;
            }
    UvwFields Fx = F;
    if (tvd) {
        if (!bcell_single_)
            throw std::invalid_argument("ConvScheme schemes need every boundary cell on one face");
        const size_t ncell = (size_t)nc_ + nbc_ + 1;
        if (!minmax_) TF_CHECK(hipMalloc(&minmax_, sizeof(double) * 2 * ncell));
        const int work = std::max(nc_, ns_);
        hipLaunchKernelGGL(k_uvw_minmax, dim3(std::max(1, std::min(2048, (work + kBlock - 1) / kBlock))),
                           dim3(kBlock), 0, s, M_, U_, F, minmax_, minmax_ + ncell);
        Fx.phimax = minmax_;
        Fx.phimin = minmax_ + ncell;
    }
    const UvwGeom G = P.var == 1 ? UvwGeom{M_.sx, M_.sy, M_.sz, M_.dx, M_.dy, M_.dz}
                    : P.var == 2 ? UvwGeom{M_.sy, M_.sx, M_.sz, M_.dy, M_.dx, M_.dz}
                                 : UvwGeom{M_.sz, M_.sx, M_.sy, M_.dz, M_.dx, M_.dy};
    const bool diff = P.diffus == kFlagCN || P.diffus == kFlagFI;
    // Aval = 0 (:148): the whole array where the cell kernel accumulates
    // into off-diagonals or writes none, else the positions it never writes
    if (offdiag_shared_ || !diff) {
        hipLaunchKernelGGL(k_calcps_zero_all, dim3((nonz_ + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nonz_, aval);
    } else if (n_zero_ > 0) {
        hipLaunchKernelGGL(k_calcps_zero, dim3((n_zero_ + kBlock - 1) / kBlock), dim3(kBlock), 0, s, n_zero_,
                           (const int*)zero_idx_, aval);
    }
    if (nbc_ > 0)                                                      // :146
        hipLaunchKernelGGL(k_calcps_zero_all, dim3((nbc_ + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nbc_, abou);
    // at most 2048 blocks, a multiple of 8 (the XCD-aware tile walk)
    const int tiles = (nc_ + kBlock - 1) / kBlock;
    const int grid = tiles < kNumXcd ? tiles : std::min(2048, (tiles + kNumXcd - 1) / kNumXcd * kNumXcd);
    auto cells = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, s, M_, U_, G, P, Fx, aval, b, abou, asave, rhoui);
    };
    const int form = (offdiag_shared_ ? 4 : 0) + (scalar ? 2 : 0) + (tvd ? 1 : 0);
    switch (form) {
    case 0: cells(k_uvw_cells<false, false, false>); break;
    case 1: cells(k_uvw_cells<false, false, true>); break;
    case 2: cells(k_uvw_cells<false, true, false>); break;
    case 3: cells(k_uvw_cells<false, true, true>); break;
    case 4: cells(k_uvw_cells<true, false, false>); break;
    case 5: cells(k_uvw_cells<true, false, true>); break;
    case 6: cells(k_uvw_cells<true, true, false>); break;
    default: cells(k_uvw_cells<true, true, true>); break;
    }
    TF_CHECK(hipGetLastError());
}

void CalcPS::zmix_assemble(const UvwParams& Pin, const UvwFields& F, double* aval, double* b, double* abou,
                           double* rhoz, hipStream_t s) {
    // BLEND_TEM drives the face scheme, the upwind matrix and the upwind
    // part (CalcZmix.f90:161-164, :176, :278); the gradients and S, D are
    // taken in x, y, z order; no pressure term, no Asave
    UvwParams P = Pin;
    P.var = 1;
    for (int m = 0; m < kMaxMaterials; ++m) P.blend[m] = P.blend_tem[m];
    assemble_equation(P, F, aval, b, abou, nullptr, rhoz, true, s);
}

void CalcPS::zmix_update(const double* rhoz, const double* rho_n, double* z_n, hipStream_t s) {
    hipLaunchKernelGGL(k_zmix_update, dim3((nc_ + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nc_, nbc_, rhoz, rho_n,
                       z_n);
    TF_CHECK(hipGetLastError());
}

void CalcPS::uvw_update(const double* rhoui, const double* rho_n, double* ui_n, hipStream_t s) {
    hipLaunchKernelGGL(k_uvw_update, dim3((nc_ + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nc_, nbc_, rhoui, rho_n,
                       ui_n);
    TF_CHECK(hipGetLastError());
}

void CalcPS::cfl_peclet(double dt, double visc, const double* rho_n, const double* flux, double* cfl, double* pe,
                        hipStream_t s) {
    if (!uvw_ready()) throw std::invalid_argument("the CFL and Peclet numbers need tflows_calcps_uvw_setup first");
    const int nv = 2 * nmaterials_;
    if (!cflpart_) {
        TF_CHECK(hipMalloc(&cflpart_, sizeof(double) * (size_t)2 * kMaxMaterials * grid_));
        TF_CHECK(hipMalloc(&cflout_, sizeof(double) * 2 * kMaxMaterials));
    }
    hipLaunchKernelGGL(k_cfl_peclet, dim3(grid_), dim3(kBlock), 0, s, M_, U_, nmaterials_, dt, visc, rho_n, flux,
                       cflpart_);
    hipLaunchKernelGGL(k_cfl_total, dim3(1), dim3(kBlock), 0, s, nv, grid_, (const double*)cflpart_, cflout_);
    TF_CHECK(hipGetLastError());
    double h[2 * kMaxMaterials];
    TF_CHECK(hipMemcpyAsync(h, cflout_, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
    TF_CHECK(hipStreamSynchronize(s));
    for (int m = 0; m < nmaterials_; ++m) {
        cfl[m] = h[2 * m];
        pe[m] = h[2 * m + 1];
    }
}

}  // namespace tflows
