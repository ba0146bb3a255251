// calcps.h -- GPU pressure-correction assembly (Process/CalcPS.f90:51-202).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "system.h"

namespace tflows {

// The face list and geometry, resident on the device (kernel form).
struct CalcPSMesh {
    int nc = 0, nbc = 0, ns = 0;
    const int* c1 = nullptr;            // SideC(1,s), 1-based cell
    const int* c2 = nullptr;            // SideC(2,s): cell > 0, boundary cell < 0
    const int* sid1 = nullptr;          // SidAij(1,s) - 1 (interior faces)
    const int* sid2 = nullptr;          // SidAij(2,s) - 1
    const unsigned char* kind = nullptr;  // what the face does to its cells
    const int* adia = nullptr;          // Adia(c) - 1
    const int* cptr = nullptr;          // cell -> faces (NC + 1)
    const int* cface = nullptr;         // 2 s + (the cell is c2), ascending s, every face
    const double *sx = nullptr, *sy = nullptr, *sz = nullptr;
    const double *dx = nullptr, *dy = nullptr, *dz = nullptr, *f = nullptr;
    const double* volume = nullptr;     // (-NbC:NC): volume(c) at [NbC + c]
};

// The per-call cell fields, (-NbC:NC), device pointers.
struct CalcPSFields {
    const double *asave, *rho_n, *rho_o, *rho_oo, *u, *v, *w, *p, *px, *py, *pz;
};

class CalcPS {
public:
    // sidec: SideC(0:2,NS), sidaij: SidAij(1:2,NS), adia: Adia(NC),
    // typebc: TypeBC(-NbC:NC), geom: Sx, Sy, Sz, Dx, Dy, Dz, f (NS),
    // volume(-NbC:NC); host or device pointers.
    CalcPS(int nc, int nbc, int ns, int nonz, const int* sidec, const int* sidaij, const int* adia,
           const int* typebc, const double* const geom[7], const double* volume, bool device_ptrs, hipStream_t s);
    ~CalcPS();
    CalcPS(const CalcPS&) = delete;
    CalcPS& operator=(const CalcPS&) = delete;

    // Aval(NONZ), b(NC), Flux(NS), Abou(-NbC:-1) (NbC entries): device
    // pointers.  errmax, mean: host (may be null).
    void assemble(int nmat, double dt, const CalcPSFields& F, double* aval, double* b, double* flux, double* abou,
                  double* errmax, double* mean, hipStream_t s);
    // The correction after the pressure solve (Process/CorUVW.f90:36-121):
    // U, V, W ((-NbC:NC)) and Flux(NS) in place, the mass error per unit
    // volume b(NC); returns max |b|.  F supplies asave, rho_n/o/oo, px, py,
    // pz (the other members are not read).  Device pointers.
    double correct(double dt, const CalcPSFields& F, const double* pp, const double* aval, const double* abou,
                   double* u, double* v, double* w, double* flux, double* b, hipStream_t s);

    int nc() const { return nc_; }
    int nbc() const { return nbc_; }
    int ns() const { return ns_; }
    int nonz() const { return nonz_; }
    // device staging for host-pointer calls (grown on demand)
    double* staging(size_t doubles);

private:
    int nc_, nbc_, ns_, nonz_;
    CalcPSMesh M_;
    std::vector<void*> owned_, staging_;
    size_t staging_size_ = 0;
    int* zero_idx_ = nullptr;
    int n_zero_ = 0;
    int grid_ = 1;
    double* fa12_ = nullptr;
    double* part_ = nullptr;
    double* out2_ = nullptr;
};

}  // namespace tflows
