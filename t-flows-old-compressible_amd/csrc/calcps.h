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

// Momentum assembly (Process/NewUVW.f90:137-499): the flags and constants
// of one call.  Codes as Process/IniPar.f90:22-32 (CN 2, FI 3, LIN 4, PAR 5,
// YES 10, NO 11); per-material BLEND, BLEND_TEM and URFC for materials
// 1..nmaterials at [m - 1].
constexpr int kMaxMaterials = 8;
struct UvwParams {
    int var = 1;                        // 1 U, 2 V, 3 W: picks Si..Dk as Process/Processor.f90:314-340
    int convec = 3, diffus = 3, cross = 3, inert = 4;
    double visc = 0.0, dt = 1.0, urf = 1.0;
    int blend[kMaxMaterials] = {}, blend_tem[kMaxMaterials] = {};
    double urfc[kMaxMaterials] = {};
    // 1: VISeff = visc * (fF rho%n(c1) + (1 - fF) rho%n(c2)) at each face, the
    // Moin_Problem() == 4 branch of NewUVW.f90:287-295 and CalcZmix.f90:232-242;
    // 0: VISeff = visc
    int visc_rho = 0;
};

// The per-call momentum fields, device pointers; cell arrays (-NbC:NC).
struct UvwFields {
    const double *rho_n, *rho_o, *rho_oo;
    const double *ui_n, *ui_o, *ui_oo, *source, *pi;
    const double* g[7];                 // dUidi, dUidj, dUidk, dUjdi, dUkdi, dUjdj, dUkdk
    const double* flux_u;               // Flux_u(NS)
    const double* flux = nullptr;       // the mass flux Flux(NS) whose sign picks ConvScheme's upwind
                                        // cell; null: Flux_u
    // PHImax / PHImin (-NbC:NC) of CalMinMax, set by the assembly itself
    const double* phimax = nullptr;
    const double* phimin = nullptr;
};

// Geometry the momentum assembly adds to the mesh (device).
struct UvwMesh {
    const int2* inc = nullptr;          // per cface entry: (2 s + role, the face's other cell)
    const int* tiles = nullptr;         // visiting order of the 256-cell tiles (nullptr: in order)
    const int* wbase = nullptr;         // first cell of each wavefront of each visited tile (else: tiles)
    const unsigned char* ukind = nullptr;  // what a face does in the diffusion loop
    const double *scoef = nullptr, *ff = nullptr;
    const int* material = nullptr;      // (-NbC:NC), 1-based
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

    // The momentum mesh data: Scoef(NS), fF(NS) (Process/Calc3.f90:23-70)
    // and material(-NbC:NC) with values 1..nmaterials; host or device.
    void uvw_setup(const double* scoef, const double* ff, const int* material, int nmaterials, bool device_ptrs,
                   hipStream_t s);
    // One velocity component's system (NewUVW.f90:137-499): Aval(NONZ),
    // b(NC), Abou(-NbC:-1), Asave(-NbC:NC) (cells 1..NC, var = 1 only; may
    // be null) and rhoUi(-NbC:NC), the solve's starting point.  Device.
    void uvw_assemble(const UvwParams& P, const UvwFields& F, double* aval, double* b, double* abou, double* asave,
                      double* rhoui, hipStream_t s);
    // Ui%n(1:NC) = rhoUi(1:NC) / rho%n(1:NC) after the solve (:526).
    void uvw_update(const double* rhoui, const double* rho_n, double* ui_n, hipStream_t s);
    // The mixture-fraction system (Process/CalcZmix.f90:100-387): F.ui_* are
    // Z n/o/oo, F.g[0..2] dZ/dx, dZ/dy, dZ/dz, F.pi unused; P.blend_tem is
    // BLEND_TEM, P.urfc URFC_Tem, P.visc VIScZmix, P.urf Zmix%URF (P.var and
    // P.blend are not read).  Aval, b, Abou, rhoZ as for uvw_assemble.
    void zmix_assemble(const UvwParams& P, const UvwFields& F, double* aval, double* b, double* abou, double* rhoz,
                       hipStream_t s);
    // Zmix%n(1:NC) = rhoZ / rho%n limited to [0, 1] (:401-405).
    void zmix_update(const double* rhoz, const double* rho_n, double* z_n, hipStream_t s);
    // The CFL and Peclet numbers of CorUVW.f90:124-147, per material, after
    // uvw_setup (which brings Scoef and the materials): rho_n (-NbC:NC) and
    // the corrected Flux (NS), device pointers; cfl and pe receive
    // nmaterials() values each on the host.
    void cfl_peclet(double dt, double visc, const double* rho_n, const double* flux, double* cfl, double* pe,
                    hipStream_t s);
    bool uvw_ready() const { return U_.ff != nullptr; }
    int nmaterials() const { return nmaterials_; }
    // how the momentum and scalar cell kernels visit their cells: 0 tiles in
    // order, 1 tiles brick by brick, 2 blocks of 64 cells x 2 lines x 2 planes
    int visit_order() const { return visit_order_; }

    int nc() const { return nc_; }
    int nbc() const { return nbc_; }
    int ns() const { return ns_; }
    int nonz() const { return nonz_; }
    // device staging for host-pointer calls (grown on demand)
    double* staging(size_t doubles);

private:
    void assemble_equation(const UvwParams& P, const UvwFields& F, double* aval, double* b, double* abou,
                           double* asave, double* rhoui, bool scalar, hipStream_t s);
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
    // momentum assembly
    UvwMesh U_;
    std::vector<unsigned char> ukind_host_;
    std::vector<int> c1_host_, c2_host_;
    bool offdiag_shared_ = false;       // two faces write one off-diagonal
    bool bcell_single_ = true;          // every boundary cell has at most one face
    double* minmax_ = nullptr;          // PHImax, PHImin of ConvScheme ((-NbC:NC) each)
    double* cflpart_ = nullptr;         // block maxima of the CFL and Peclet numbers
    double* cflout_ = nullptr;          // their totals, 2 per material
    int nmaterials_ = 0;
    int visit_order_ = 0;
    unsigned pairs_[kMaxMaterials] = {};  // material pairs that meet at a face
};

}  // namespace tflows
