/*
 * calcps_oracle.c -- CPU restatement of the pressure-correction assembly of
 * T-FlowS (Process/CalcPS.f90:51-202): the matrix, the mass source and the
 * face fluxes that the pressure solve (CalcPS.f90:214-219, cg_) consumes.
 *
 * TEST INFRASTRUCTURE ONLY, like krylov_oracle.c: only tests/ and
 * bench.py's CPU-baseline leg load it, to compare against or to time.
 *
 * Statement by statement, in the reference's floating-point order (Fortran
 * evaluates a+b+c as (a+b)+c and a*b*c as (a*b)*c; every real is REAL*8
 * under -fdefault-real-8, literals included):
 *
 *   :51-57    Aval(:) = 0, b(:) = 0
 *   :68-168   do m = 1, Nmat; do s = 1, NS -- the face loop: SMDPN, the
 *             interpolated mass velocities, A12, the off-diagonal and
 *             diagonal contributions, the face flux and its effect on b, for
 *             interior / BUFFER faces (:74-115) and for INFLOW (:120-133),
 *             OUTFLOW and CONVECT (:134-148), PRESSURE (:149-159) and every
 *             other boundary type (:160-161, Flux = 0)
 *   :172-175  errmax = max |b|
 *   :188-199  do m = 1, Nmat; do c = 1, NC -- the time term
 *             b -= V/dt (1.5 rho^n - 2 rho^o + 0.5 rho^oo), with the sums of
 *             b V and of V
 *   :201-204  b -= bc_sum / vol (GloSum is a no-op in one domain)
 *
 * Array bases are the reference's: SideC(0:2,NS) is sidec[3*(s-1) + k],
 * SidAij(1:2,NS) is sidaij[2*(s-1) + k-1], Adia(1:NC), Sx..f(1:NS),
 * Aval(1:NONZ), b(1:NC), Flux(1:NS); the cell arrays are declared
 * (-NbC:NC) (GeoAloc.f90:31, 54; UnkAloc.f90:91; CnsLoa.f90:58) and
 * x(c) is x[nbc + c]; Abou(-NbC:-1) (GeoAloc.f90:55) is abou[nbc + c].
 * Boundary-condition codes are those of Library/allp_mod.f90:16-24.
 *
 * The reference routine itself cannot be built here (it uses
 * Moin_Problem_mod, whose module chain needs gfortran's `irand`
 * extension), so this restatement is pinned by closed forms and invariants
 * in tests/test_calcps_oracle.py.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>

enum { BC_INFLOW = 1, BC_WALL = 2, BC_OUTFLOW = 3, BC_SYMMETRY = 4, BC_CONVECT = 5, BC_WALLFL = 6,
       BC_BUFFER = 11, BC_PRESSURE = 12 };

void oracle_calcps(int nc, int nbc, int ns, int nonz, int nmat, const int *sidec, const int *sidaij,
                   const int *adia, const int *typebc, const double *sx, const double *sy, const double *sz,
                   const double *dx, const double *dy, const double *dz, const double *f, const double *volume,
                   const double *asave, const double *rho_n, const double *rho_o, const double *rho_oo,
                   const double *u, const double *v, const double *w, const double *p, const double *px,
                   const double *py, const double *pz, double dt, double *aval, double *b, double *flux,
                   double *abou, double *errmax, double *mean, double *b_pre)
{
#define C(x, c) (x)[nbc + (c)]
    for (int j = 0; j < nonz; ++j) aval[j] = 0.0;                       /* :51 */
    for (int c = 0; c < nc; ++c) b[c] = 0.0;                            /* :56 */

    for (int m = 1; m <= nmat; ++m) {                                   /* :68 */
        for (int s = 1; s <= ns; ++s) {                                 /* :69 */
            const int c1 = sidec[3 * (s - 1) + 1];
            const int c2 = sidec[3 * (s - 1) + 2];
            const double Sx = sx[s - 1], Sy = sy[s - 1], Sz = sz[s - 1];
            const double Dx = dx[s - 1], Dy = dy[s - 1], Dz = dz[s - 1];
            const double fs = f[s - 1];
            if (c2 > 0 || (c2 < 0 && C(typebc, c2) == BC_BUFFER)) {    /* :74 */
                const double SMDPN = ((Sx * Sx + Sy * Sy) + Sz * Sz) / ((Sx * Dx + Sy * Dy) + Sz * Dz);
                const double Urs = (fs * C(rho_n, c1)) * C(u, c1) + ((1.0 - fs) * C(rho_n, c2)) * C(u, c2);
                const double Vrs = (fs * C(rho_n, c1)) * C(v, c1) + ((1.0 - fs) * C(rho_n, c2)) * C(v, c2);
                const double Wrs = (fs * C(rho_n, c1)) * C(w, c1) + ((1.0 - fs) * C(rho_n, c2)) * C(w, c2);
                const double A12 = (0.5 * SMDPN) * (C(volume, c1) / C(asave, c1) + C(volume, c2) / C(asave, c2));
                if (c2 > 0) {                                           /* :86-94 */
                    aval[sidaij[2 * (s - 1)] - 1] = -A12;
                    aval[sidaij[2 * (s - 1) + 1] - 1] = -A12;
                    aval[adia[c1 - 1] - 1] = aval[adia[c1 - 1] - 1] + A12;
                    aval[adia[c2 - 1] - 1] = aval[adia[c2 - 1] - 1] + A12;
                } else {                                                /* :96-101 */
                    aval[adia[c1 - 1] - 1] = aval[adia[c1 - 1] - 1] + A12;
                    C(abou, c2) = -A12;
                }
                const double dPxi = (0.5 * (C(px, c1) + C(px, c2))) * Dx;   /* :104-106 */
                const double dPyi = (0.5 * (C(py, c1) + C(py, c2))) * Dy;
                const double dPzi = (0.5 * (C(pz, c1) + C(pz, c2))) * Dz;
                const double F = (((Urs * Sx + Vrs * Sy) + Wrs * Sz) + A12 * (C(p, c1) - C(p, c2))) +
                                 A12 * ((dPxi + dPyi) + dPzi);          /* :110-112 */
                flux[s - 1] = F;
                b[c1 - 1] = b[c1 - 1] - F;                              /* :114-115 */
                if (c2 > 0) b[c2 - 1] = b[c2 - 1] + F;
            } else {
                const int t = C(typebc, c2);
                if (t == BC_INFLOW) {                                   /* :120-133 */
                    const double Urs = (fs * C(rho_n, c1)) * C(u, c1) + ((1.0 - fs) * C(rho_n, c2)) * C(u, c2);
                    const double Vrs = (fs * C(rho_n, c1)) * C(v, c1) + ((1.0 - fs) * C(rho_n, c2)) * C(v, c2);
                    const double Wrs = (fs * C(rho_n, c1)) * C(w, c1) + ((1.0 - fs) * C(rho_n, c2)) * C(w, c2);
                    const double F = (Urs * Sx + Vrs * Sy) + Wrs * Sz;  /* A12 of :128-130 is unused */
                    flux[s - 1] = F;
                    b[c1 - 1] = b[c1 - 1] - F;
                } else if (t == BC_OUTFLOW || t == BC_CONVECT) {        /* :134-148 */
                    const double Urs = (fs * C(rho_n, c1)) * C(u, c1) + ((1.0 - fs) * C(rho_n, c2)) * C(u, c2);
                    const double Vrs = (fs * C(rho_n, c1)) * C(v, c1) + ((1.0 - fs) * C(rho_n, c2)) * C(v, c2);
                    const double Wrs = (fs * C(rho_n, c1)) * C(w, c1) + ((1.0 - fs) * C(rho_n, c2)) * C(w, c2);
                    const double SMDPN = ((Sx * Sx + Sy * Sy) + Sz * Sz) / ((Sx * Dx + Sy * Dy) + Sz * Dz);
                    const double A12 = (SMDPN * C(volume, c1)) / C(asave, c1);
                    aval[adia[c1 - 1] - 1] = aval[adia[c1 - 1] - 1] + A12;
                    const double F = (Urs * Sx + Vrs * Sy) + Wrs * Sz;
                    flux[s - 1] = F;
                    b[c1 - 1] = b[c1 - 1] - F;
                } else if (t == BC_PRESSURE) {                          /* :149-159 */
                    const double Urs = C(rho_n, c1) * C(u, c1);
                    const double Vrs = C(rho_n, c1) * C(v, c1);
                    const double Wrs = C(rho_n, c1) * C(w, c1);
                    const double F = (Urs * Sx + Vrs * Sy) + Wrs * Sz;
                    flux[s - 1] = F;
                    b[c1 - 1] = b[c1 - 1] - F;
                    const double SMDPN = ((Sx * Sx + Sy * Sy) + Sz * Sz) / ((Sx * Dx + Sy * Dy) + Sz * Dz);
                    const double A12 = (SMDPN * C(volume, c1)) / C(asave, c1);
                    aval[adia[c1 - 1] - 1] = aval[adia[c1 - 1] - 1] + A12;
                } else {                                                /* :160-161 SYMMETRY, walls */
                    flux[s - 1] = 0.0;
                }
            }
        }
    }

    double em = 0.0;                                                    /* :172-175 */
    for (int c = 1; c <= nc; ++c) em = fmax(em, fabs(b[c - 1]));
    *errmax = em;

    double bc_sum = 0.0, vol = 0.0;                                     /* :184-185 */
    for (int m = 1; m <= nmat; ++m) {
        for (int c = 1; c <= nc; ++c) {
            b[c - 1] = b[c - 1] - (C(volume, c) / dt) *
                                      ((1.5 * C(rho_n, c) - 2.0 * C(rho_o, c)) + 0.5 * C(rho_oo, c));
            bc_sum = bc_sum + b[c - 1] * C(volume, c);
            vol = vol + C(volume, c);
        }
    }
    const double mn = bc_sum / vol;
    if (b_pre)                          /* b before the mean (a checker's view) */
        for (int c = 0; c < nc; ++c) b_pre[c] = b[c];
    for (int c = 1; c <= nc; ++c) b[c - 1] = b[c - 1] - mn;             /* :201-204 */
    *mean = mn;
#undef C
}

/*
 * oracle_coruvw -- Process/CorUVW.f90:36-121 and :183 (the correction
 * after the pressure solve; the CFL / Peclet monitor of :124-147 is
 * oracle_cfl_peclet below):
 *
 *   :36-46    U = U - Px V / Asave / rho (both algorithm branches compute
 *             the same values), V, W likewise, cells 1..NC
 *   :48-59    PRESSURE boundary cells take the corrected velocity of c1
 *   :76-87    Flux(s) += (PP(c2) - PP(c1)) * Aval(SidAij(1,s)) for interior
 *             faces, * Abou(c2) for BUFFER faces
 *   :97-115   b = 0; b(c1) -= Flux, b(c2) += Flux (interior) or b(c1) -= Flux
 *             (every boundary face); b -= V/dt (1.5 rho^n - 2 rho^o +
 *             0.5 rho^oo); b /= V
 *   :117-121  errmax = max |b| (glomax is a no-op in one domain), which is
 *             the function's value (:183)
 *
 * Same array bases as oracle_calcps; u, v, w and flux are updated in place.
 */
double oracle_coruvw(int nc, int nbc, int ns, const int *sidec, const int *sidaij, const int *typebc,
                     const double *volume, const double *asave, const double *rho_n, const double *rho_o,
                     const double *rho_oo, const double *px, const double *py, const double *pz, const double *pp,
                     const double *aval, const double *abou, double dt, double *u, double *v, double *w,
                     double *flux, double *b)
{
#define C(x, c) (x)[nbc + (c)]
    for (int c = 1; c <= nc; ++c) {                                     /* :36-46 */
        C(u, c) = C(u, c) - ((C(px, c) * C(volume, c)) / C(asave, c)) / C(rho_n, c);
        C(v, c) = C(v, c) - ((C(py, c) * C(volume, c)) / C(asave, c)) / C(rho_n, c);
        C(w, c) = C(w, c) - ((C(pz, c) * C(volume, c)) / C(asave, c)) / C(rho_n, c);
    }
    for (int s = 1; s <= ns; ++s) {                                     /* :48-59 */
        const int c1 = sidec[3 * (s - 1) + 1], c2 = sidec[3 * (s - 1) + 2];
        if (c2 < 0 && C(typebc, c2) == BC_PRESSURE) {
            C(u, c2) = C(u, c1);
            C(v, c2) = C(v, c1);
            C(w, c2) = C(w, c1);
        }
    }
    for (int s = 1; s <= ns; ++s) {                                     /* :76-87 */
        const int c1 = sidec[3 * (s - 1) + 1], c2 = sidec[3 * (s - 1) + 2];
        if (c2 > 0 || (c2 < 0 && C(typebc, c2) == BC_BUFFER)) {
            if (c2 > 0)
                flux[s - 1] = flux[s - 1] + (C(pp, c2) - C(pp, c1)) * aval[sidaij[2 * (s - 1)] - 1];
            else
                flux[s - 1] = flux[s - 1] + (C(pp, c2) - C(pp, c1)) * C(abou, c2);
        }
    }
    for (int c = 0; c < nc; ++c) b[c] = 0.0;                            /* :97-99 */
    for (int s = 1; s <= ns; ++s) {                                     /* :101-110 */
        const int c1 = sidec[3 * (s - 1) + 1], c2 = sidec[3 * (s - 1) + 2];
        if (c2 > 0 || (c2 < 0 && C(typebc, c2) == BC_BUFFER)) {
            b[c1 - 1] = b[c1 - 1] - flux[s - 1];
            if (c2 > 0) b[c2 - 1] = b[c2 - 1] + flux[s - 1];
        } else {
            b[c1 - 1] = b[c1 - 1] - flux[s - 1];
        }
    }
    for (int c = 1; c <= nc; ++c) {                                     /* :112-115 */
        b[c - 1] = b[c - 1] - (C(volume, c) / dt) *
                                  ((1.5 * C(rho_n, c) - 2.0 * C(rho_o, c)) + 0.5 * C(rho_oo, c));
        b[c - 1] = b[c - 1] / C(volume, c);
    }
    double em = 0.0;                                                    /* :117-121 */
    for (int c = 1; c <= nc; ++c) em = fmax(em, fabs(b[c - 1]));
    return em;
#undef C
}

/*
 * oracle_cfl_peclet -- the CFL and Peclet monitor of Process/CorUVW.f90:
 * 124-147, after the correction:
 *
 *   for m = 1..Nmat: CFLmax(m) = PeMax(m) = 0, then over s = 1..NS with
 *   material(c1) = m or material(c2) = m, and c2 > 0 or TypeBC(c2) = BUFFER:
 *     RhoS = f(s) rho^n(c1) + (1 - f(s)) rho^n(c2)   (the geometric f)
 *     CFLs = |dt Flux(s) / (RhoS Scoef(s) (Dx^2 + Dy^2 + Dz^2))|
 *     PeS  = |Flux(s) / Scoef(s) / (VISc + TINY)|
 *     CFLmax(m) = max(CFLmax(m), CFLs), PeMax(m) = max(PeMax(m), PeS)
 *   (glomax is a no-op in one domain).
 *
 * material is (-NbC:NC), 1-based values; cflmax, pemax receive Nmat values.
 */
void oracle_cfl_peclet(int nc, int nbc, int ns, int nmat, const int *sidec, const int *typebc, const int *material,
                       const double *f, const double *scoef, const double *dx, const double *dy, const double *dz,
                       const double *rho_n, const double *flux, double dt, double visc, double *cflmax,
                       double *pemax)
{
#define C(x, c) (x)[nbc + (c)]
    (void)nc;
    const double TINY = 1.0e-64;                                        /* Library/allp_mod.f90:31 */
    for (int m = 1; m <= nmat; ++m) {
        double cmax = 0.0, pmax = 0.0;
        for (int s = 1; s <= ns; ++s) {
            const int c1 = sidec[3 * (s - 1) + 1], c2 = sidec[3 * (s - 1) + 2];
            if (C(material, c1) != m && C(material, c2) != m) continue;
            if (!(c2 > 0 || (c2 < 0 && C(typebc, c2) == BC_BUFFER))) continue;
            const double RhoS = f[s - 1] * C(rho_n, c1) + (1.0 - f[s - 1]) * C(rho_n, c2);
            const double dd = dx[s - 1] * dx[s - 1] + dy[s - 1] * dy[s - 1] + dz[s - 1] * dz[s - 1];
            const double CFLs = fabs(dt * flux[s - 1] / (RhoS * scoef[s - 1] * dd));
            const double PeS = fabs(flux[s - 1] / scoef[s - 1] / (visc + TINY));
            cmax = CFLs > cmax ? CFLs : cmax;
            pmax = PeS > pmax ? PeS : pmax;
        }
        cflmax[m - 1] = cmax;
        pemax[m - 1] = pmax;
    }
#undef C
}

/*
 * oracle_newuvw -- the momentum assembly of one velocity component,
 * Process/NewUVW.f90:137-499 (everything before the BiCG call at :507; the
 * shift of the old levels at :154-157 and the source terms at :180-199 are
 * the caller's, the source arrives as an array):
 *
 *   :137-138  rhoUi = rho^n Ui^n on (-NbC:NC)
 *   :146-149  Abou = 0, Aval = 0, b = 0
 *   :175-176  C = 0, X = 0
 *   :167-172  PHImax / PHImin of rhoUi (CalMinMax) when a face takes a
 *             ConvScheme scheme
 *   :205-240  convection face loop: rhoUiS = f rhoUi(c1) + (1-f) rhoUi(c2),
 *             then ConvScheme's face value when max(BLEND(m1), BLEND(m2))
 *             /= NO (kept as it is for YES); C(c1) -= Flux_u rhoUiS,
 *             C(c2) += ...; the upwind part X when BLEND_TEM /= NO
 *   :252-257  CONVEC = FI: b += URFC(material) (C - X)
 *   :264-266  X = 0 on the cells
 *   :277-416  diffusion face loop: the cross-diffusion flux Fex - Fim into
 *             X; D12 = A0/rho(c2), D21 = A0/rho(c1) for DIFFUS = FI; the
 *             upwind A12, A21 when BLEND /= NO; the "m1" matrix fill for
 *             interior faces and the boundary branches (WALL, CONVECT,
 *             INFLOW, WALLFL; OUTFLOW; BUFFER into Abou)
 *   :434-438  CROSS = FI: b += X
 *   :447-456  inertia, INERT = LIN (two levels) or PAR (three levels)
 *   :484      b -= Pi V
 *   :497-499  Asave = diagonal (var = 1); under-relaxation with U%URF
 *
 * The flag and boundary codes are those of Process/IniPar.f90:24-32 and
 * Library/allp_mod.f90:16-24.  Per-material arrays BLEND, BLEND_TEM, URFC
 * are 1-based (material m at [m-1]); cell arrays (-NbC:NC) as elsewhere.
 * VISeff (:285-295): VISc, or with visc_rho = 1 the Moin_Problem() == 4
 * branch VISc RhoS, RhoS = fF rho^n(c1) + (1-fF) rho^n(c2).  flux is the module's
 * mass flux Flux, whose sign picks ConvScheme's upwind cell.  A face whose
 * materials' BLEND values have max(...) = NO next to YES (for which the
 * reference leaves Phij undefined), or any other uncovered code, returns
 * -1.  abou is (-NbC:-1); asave (NC) may be NULL; cbuf and xbuf hold
 * NbC+NC+1 doubles.  Returns 0.
 */
enum { FL_FI = 3, FL_CN = 2, IN_LIN = 4, IN_PAR = 5, BL_YES = 10, BL_NO = 11 };
/* ConvScheme's face schemes, Process/IniPar.f90:42-49 */
enum { CS_CDS = 20, CS_QUICK = 21, CS_LUDS = 22, CS_MINMOD = 23, CS_SMART = 24, CS_AVL_SMART = 25,
       CS_SUPERBEE = 26, CS_GAMMA = 27 };

/* MAX / MIN of two or three values as plain comparisons (the first
 * argument is kept unless another is strictly beyond it). */
static double mx2(double a, double b) { return b > a ? b : a; }
static double mn2(double a, double b) { return b < a ? b : a; }
static double mn3(double a, double b, double c) { return mn2(mn2(a, b), c); }
static double mx3(double a, double b, double c) { return mx2(mx2(a, b), c); }

/* A BLEND value max(BLEND(m1), BLEND(m2)) that this restatement covers:
 * YES, or a ConvScheme scheme. */
static int blend_covered(int bl) { return bl == BL_YES || (bl >= CS_CDS && bl <= CS_GAMMA); }

/*
 * Process/CalMinMax.f90:21-37: PHImax / PHImin of every cell (-NbC:NC) over
 * the cell and the neighbours across its interior and BUFFER faces.
 */
void oracle_calminmax(int nc, int nbc, int ns, const int *sidec, const int *typebc, const double *phi,
                      double *phimax, double *phimin)
{
    for (int c = 0; c < nbc + nc + 1; ++c) {                             /* :21-22 */
        phimax[c] = phi[c];
        phimin[c] = phi[c];
    }
    for (int s = 1; s <= ns; ++s) {                                     /* :24-35 */
        const int c1 = sidec[3 * (s - 1) + 1], c2 = sidec[3 * (s - 1) + 2];
        if (c2 > 0 || (c2 < 0 && typebc[nbc + c2] == BC_BUFFER)) {
            const int i1 = nbc + c1, i2 = nbc + c2;
            phimax[i1] = mx2(phimax[i1], phi[i2]);
            phimin[i1] = mn2(phimin[i1], phi[i2]);
            phimax[i2] = mx2(phimax[i2], phi[i1]);
            phimin[i2] = mn2(phimin[i2], phi[i1]);
        }
    }
}

/*
 * Process/ConvScheme.f90:55-135: the face value of Phi at face s for
 * scheme `blenda`, given the central value phis.  The upwind cell C and the
 * downwind cell D follow the sign of Flux(s) (the mass flux of the module,
 * :55-65); gi, gj, gk are the gradient components and di, dj, dk the centre
 * distances in the order the caller passes them.  YES keeps phis (:116).
 * Indices are 0-based into (-NbC:NC) arrays (nbc added by the caller).
 */
static double conv_scheme(double phis, int s, int c1, int c2, int nbc, const double *phi, const double *gi,
                          const double *gj, const double *gk, const double *di, const double *dj,
                          const double *dk, const double *f, const double *flux, const double *phimax,
                          const double *phimin, int blenda)
{
    if (blenda == BL_YES) return phis;                                  /* :116 (PhiU etc. unused) */
    const int pos = flux[s - 1] > 0.0;
    const double fj = pos ? 1.0 - f[s - 1] : f[s - 1];                   /* :55-65 */
    const int Ci = nbc + (pos ? c1 : c2), Di = nbc + (pos ? c2 : c1);
    const double sign = pos ? 1.0 : -1.0;
    const double gd = gi[Ci] * di[s - 1] + gj[Ci] * dj[s - 1] + gk[Ci] * dk[s - 1];
    const double PhiUstar = pos ? phi[Di] - 2.0 * gd : phi[Di] + 2.0 * gd;       /* :67-75 */
    const double PhiU = mx2(phimin[Ci], mn2(PhiUstar, phimax[Ci]));             /* :77 */
    double rj = (phi[Ci] - PhiU) / (phi[Di] - phi[Ci] + 1.0e-16);               /* :79 */
    const double gD = 0.5 * fj * (1.0 + fj), gU = 0.5 * fj * (1.0 - fj);        /* :81-82 */
    double Phij = 0.0, alfa, beta1, beta2;
    switch (blenda) {                                                   /* :84-116 */
    case CS_CDS: Phij = fj; break;
    case CS_QUICK:
        rj = (phi[Ci] - PhiU) / (phi[Di] - phi[Ci] + 1.0e-12);
        alfa = 0.0;
        Phij = (gD - alfa) + (gU + alfa) * rj;
        break;
    case CS_LUDS:
        alfa = 0.5 * fj * (1 + fj);
        Phij = (gD - alfa) + (gU + alfa) * rj;
        break;
    case CS_MINMOD: Phij = fj * mx2(0.0, mn2(rj, 1.0)); break;
    case CS_SMART:
        beta1 = 3.0;
        beta2 = 1.0;
        Phij = mx2(0.0, mn3((beta1 - 1.0) * rj, gD + gU * rj, beta2));
        break;
    case CS_AVL_SMART:
        beta1 = 1.0 + fj * (2.0 + fj);
        beta2 = fj * (2.0 - fj);
        Phij = mx2(0.0, mn3((beta1 - 1.0) * rj, gD + gU * rj, beta2));
        break;
    case CS_SUPERBEE: Phij = 0.5 * mx3(0.0, mn2(2.0 * rj, 1.0), mn2(rj, 2.0)); break;
    default: break;                      /* GAMMA: the value is set below */
    }
    if (blenda != CS_GAMMA) return phi[Ci] + Phij * sign * (phi[nbc + c2] - phi[nbc + c1]);   /* :118 */
    const double Beta = 0.1;                                            /* :120-134 */
    const double PhiUg = pos ? 1.0 - (phi[Di] - phi[Ci]) / (2.0 * gd) : 1.0 + (phi[Di] - phi[Ci]) / (2.0 * gd);
    const double GammaC = PhiUg / Beta;
    if (PhiUg < Beta && PhiUg > 0.0)
        return (1.0 - GammaC * (1.0 - f[s - 1])) * phi[Ci] + GammaC * (1.0 - f[s - 1]) * phi[Di];
    if (PhiUg < 1.0 && PhiUg >= Beta) return f[s - 1] * phi[Ci] + (1.0 - f[s - 1]) * phi[Di];
    return phi[Ci];
}

int oracle_newuvw(int nc, int nbc, int ns, int nonz, int var, int convec, int diffus, int cross, int inert,
                  const int *sidec, const int *sidaij, const int *adia, const int *typebc, const int *material,
                  const int *blend, const int *blend_tem, const double *urfc, const double *f, const double *fF,
                  const double *scoef, const double *flux_u, const double *flux, const double *si,
                  const double *sj, const double *sk, const double *di, const double *dj, const double *dk,
                  const double *volume, const double *rho_n,
                  const double *rho_o, const double *rho_oo, const double *ui_n, const double *ui_o,
                  const double *ui_oo, const double *source, const double *pi, const double *dUidi,
                  const double *dUidj, const double *dUidk, const double *dUjdi, const double *dUkdi,
                  const double *dUjdj, const double *dUkdk, double visc, int visc_rho, double dt, double urf,
                  double *aval,
                  double *b, double *abou, double *asave, double *rhoui, double *cbuf, double *xbuf)
{
#define C(x, c) (x)[nbc + (c)]
#define MAT(arr, c) (arr)[C(material, c) - 1]
    const int nall = nbc + nc + 1;
    for (int c = -nbc; c <= nc; ++c) C(rhoui, c) = C(rho_n, c) * C(ui_n, c);   /* :137-138 */
    for (int c = 0; c < nbc; ++c) abou[c] = 0.0;                        /* :146 */
    for (int j = 0; j < nonz; ++j) aval[j] = 0.0;                       /* :148 */
    for (int c = 0; c < nc; ++c) b[c] = 0.0;                            /* :149 */
    for (int c = 0; c < nall; ++c) { cbuf[c] = 0.0; xbuf[c] = 0.0; }   /* :175-176 */

    /* PHImax / PHImin of rhoUi when any face takes a ConvScheme scheme
     * (:167-172; CalMinMax is called when any material's BLEND /= NO) */
    double *phimax = NULL, *phimin = NULL;
    for (int s = 1; s <= ns; ++s) {
        const int c1 = sidec[3 * (s - 1) + 1], c2 = sidec[3 * (s - 1) + 2];
        if (MAT(blend, c1) != BL_NO || MAT(blend, c2) != BL_NO) {
            const int bl = MAT(blend, c1) > MAT(blend, c2) ? MAT(blend, c1) : MAT(blend, c2);
            if (!blend_covered(bl)) {                                   /* NO next to YES: Phij undefined */
                free(phimax);
                free(phimin);
                return -1;
            }
            if (bl != BL_YES && !phimax) {
                phimax = (double *)malloc(sizeof(double) * nall);
                phimin = (double *)malloc(sizeof(double) * nall);
                oracle_calminmax(nc, nbc, ns, sidec, typebc, rhoui, phimax, phimin);
            }
        }
    }

    for (int s = 1; s <= ns; ++s) {                                     /* :205-240 */
        const int c1 = sidec[3 * (s - 1) + 1], c2 = sidec[3 * (s - 1) + 2];
        double rhoUiS = f[s - 1] * C(rhoui, c1) + (1.0 - f[s - 1]) * C(rhoui, c2);
        if (MAT(blend, c1) != BL_NO || MAT(blend, c2) != BL_NO) {       /* :213-216 */
            const int bl = MAT(blend, c1) > MAT(blend, c2) ? MAT(blend, c1) : MAT(blend, c2);
            rhoUiS = conv_scheme(rhoUiS, s, c1, c2, nbc, rhoui, dUidi, dUidj, dUidk, di, dj, dk, f, flux, phimax,
                                 phimin, bl);
        }
        const double F = flux_u[s - 1];
        if (c2 > 0) {
            C(cbuf, c1) = C(cbuf, c1) - F * rhoUiS;
            C(cbuf, c2) = C(cbuf, c2) + F * rhoUiS;
        } else {
            C(cbuf, c1) = C(cbuf, c1) - F * rhoUiS;
        }
        if (MAT(blend_tem, c1) != BL_NO || MAT(blend_tem, c2) != BL_NO) {
            if (F < 0) {
                C(xbuf, c1) = C(xbuf, c1) - F * C(rhoui, c2);
                if (c2 > 0) C(xbuf, c2) = C(xbuf, c2) + F * C(rhoui, c2);
            } else {
                C(xbuf, c1) = C(xbuf, c1) - F * C(rhoui, c1);
                if (c2 > 0) C(xbuf, c2) = C(xbuf, c2) + F * C(rhoui, c1);
            }
        }
    }
    if (convec == FL_FI)                                                /* :252-257 */
        for (int c = 1; c <= nc; ++c)
            b[c - 1] = b[c - 1] + MAT(urfc, c) * (C(cbuf, c) - C(xbuf, c));
    for (int c = 1; c <= nc; ++c) C(xbuf, c) = 0.0;                     /* :264-266 */

    for (int s = 1; s <= ns; ++s) {                                     /* :277-416 */
        double A12 = 0.0, A21 = 0.0, D12 = 0.0, D21 = 0.0;
        const int c1 = sidec[3 * (s - 1) + 1], c2 = sidec[3 * (s - 1) + 2];
        const double ff = fF[s - 1], gg = 1.0 - fF[s - 1];
        const double RhoS = ff * C(rho_n, c1) + gg * C(rho_n, c2);       /* NewUVW :285, CalcZmix :232 */
        const double VISeff = visc_rho ? visc * RhoS : visc;            /* Moin_Problem() == 4 */
        const double dUidiS = ff * C(dUidi, c1) + gg * C(dUidi, c2);
        const double dUidjS = ff * C(dUidj, c1) + gg * C(dUidj, c2);
        const double dUidkS = ff * C(dUidk, c1) + gg * C(dUidk, c2);
        const double dUjdiS = ff * C(dUjdi, c1) + gg * C(dUjdi, c2);
        const double dUkdiS = ff * C(dUkdi, c1) + gg * C(dUkdi, c2);
        const double dUjdjS = ff * C(dUjdj, c1) + gg * C(dUjdj, c2);
        const double dUkdkS = ff * C(dUkdk, c1) + gg * C(dUkdk, c2);
        const double Fex = VISeff * ((((4.0 / 3.0) * dUidiS * si[s - 1] + (dUidjS + dUjdiS) * sj[s - 1]) +
                                      (dUidkS + dUkdiS) * sk[s - 1]) -
                                     (2.0 / 3.0) * (dUjdjS + dUkdkS) * si[s - 1]);
        const double A0 = VISeff * scoef[s - 1];
        const double Fim = A0 * ((dUidiS * di[s - 1] + dUidjS * dj[s - 1]) + dUidkS * dk[s - 1]);
        C(xbuf, c1) = C(xbuf, c1) + Fex - Fim;                          /* :338-341 */
        if (c2 > 0) C(xbuf, c2) = C(xbuf, c2) - Fex + Fim;
        if (diffus == FL_CN || diffus == FL_FI) {
            if (diffus == FL_FI) {
                D12 = D12 + A0 / C(rho_n, c2);
                D21 = D21 + A0 / C(rho_n, c1);
            }
            if (MAT(blend, c1) != BL_NO || MAT(blend, c2) != BL_NO) {
                const double F = flux_u[s - 1];
                A12 = A12 - (F < 0.0 ? F : 0.0);
                A21 = A21 + (F > 0.0 ? F : 0.0);
            }
            if (c2 > 0) {                                               /* "m1" */
                aval[adia[c1 - 1] - 1] = aval[adia[c1 - 1] - 1] + A21 + D21;
                aval[sidaij[2 * (s - 1) + 1] - 1] = aval[sidaij[2 * (s - 1) + 1] - 1] - A21 - D21;
                aval[adia[c2 - 1] - 1] = aval[adia[c2 - 1] - 1] + A12 + D12;
                aval[sidaij[2 * (s - 1)] - 1] = aval[sidaij[2 * (s - 1)] - 1] - A12 - D12;
            } else if (c2 < 0) {
                const int t = C(typebc, c2);
                if (t == BC_WALL || t == BC_CONVECT || t == BC_INFLOW || t == BC_WALLFL) {
                    aval[adia[c1 - 1] - 1] = aval[adia[c1 - 1] - 1] + A21 + D21;
                    b[c1 - 1] = b[c1 - 1] + (A12 + D12) * C(rhoui, c2);
                } else if (t == BC_OUTFLOW) {
                    aval[adia[c1 - 1] - 1] = aval[adia[c1 - 1] - 1] + A21;
                    b[c1 - 1] = b[c1 - 1] + A12 * C(rhoui, c2);
                } else if (t == BC_BUFFER) {
                    aval[adia[c1 - 1] - 1] = aval[adia[c1 - 1] - 1] + A21 + D21;
                    C(abou, c2) = -A12 - D12;
                }
            }
        }
    }
    if (cross == FL_FI)                                                 /* :434-438 */
        for (int c = 1; c <= nc; ++c) b[c - 1] = b[c - 1] + C(xbuf, c);
    if (inert == IN_LIN)                                                /* :447-450 */
        for (int c = 1; c <= nc; ++c) {
            aval[adia[c - 1] - 1] = aval[adia[c - 1] - 1] + C(volume, c) / dt;
            b[c - 1] = b[c - 1] + C(volume, c) / dt * C(rho_o, c) * C(ui_o, c) + C(source, c) * C(volume, c);
        }
    if (inert == IN_PAR)                                                /* :453-456 */
        for (int c = 1; c <= nc; ++c) {
            aval[adia[c - 1] - 1] = aval[adia[c - 1] - 1] + 1.5 * C(volume, c) / dt;
            b[c - 1] = b[c - 1] + 2.0 * C(volume, c) / dt * C(rho_o, c) * C(ui_o, c) -
                       0.5 * C(volume, c) / dt * C(rho_oo, c) * C(ui_oo, c) + C(source, c) * C(volume, c);
        }
    for (int c = 1; c <= nc; ++c) b[c - 1] = b[c - 1] - C(pi, c) * C(volume, c);   /* :484 */
    if (var == 1 && asave)                                              /* :497 */
        for (int c = 1; c <= nc; ++c) asave[c - 1] = aval[adia[c - 1] - 1];
    for (int c = 1; c <= nc; ++c) {                                     /* :498-499 */
        b[c - 1] = b[c - 1] + aval[adia[c - 1] - 1] * (1.0 - urf) * C(rhoui, c) / urf;
        aval[adia[c - 1] - 1] = aval[adia[c - 1] - 1] / urf;
    }
    free(phimax);
    free(phimin);
    return 0;
#undef MAT
#undef C
}

/*
 * oracle_calczmix -- the assembly of the mixture-fraction equation,
 * Process/CalcZmix.f90:100-387 (everything before the CG call at :394; the
 * old-level shift :110-113 and the source :141-151 are the caller's):
 *
 *   :100      rhoZ = rho^n Z^n on (-NbC:NC)
 *   :104-107  Abou = 0, Aval = 0, b = 0
 *   :134-135  C = 0, X = 0
 *   :126-131  PHImax / PHImin of rhoZ (CalMinMax) when a face takes a
 *             ConvScheme scheme
 *   :155-189  convection face loop: rhoZS = f rhoZ(c1) + (1-f) rhoZ(c2),
 *             then ConvScheme's face value (kept as it is for YES); C,
 *             and the upwind part X when BLEND_TEM /= NO
 *   :201-207  CONVEC = FI: b += URFC_Tem(material) (C - X)
 *   :217-219  X = 0 on the cells
 *   :224-341  diffusion face loop: Fex = VISeff grad(Z)_f . S,
 *             Fim = A0 grad(Z)_f . D, X += Fex - Fim; D12 = A0/rho(c2),
 *             D21 = A0/rho(c1) for DIFFUS = FI; the upwind A12, A21 when
 *             BLEND_TEM /= NO; the "m1" fill and the boundary branches
 *   :359-361  CROSS = FI: b += X
 *   :369-378  inertia, LIN: b += V/dt Z_o rho_o + S V (in that order);
 *             PAR: b += 2 V/dt rho_o Z_o - V/(2 dt) rho_oo Z_oo + S V
 *   :386-387  under-relaxation with Zmix%URF
 *
 * VISeff (:232-242): VIScZmix, or with visc_rho = 1 the Moin_Problem() == 4
 * branch VIScZmix RhoS, RhoS as for oracle_newuvw.  Per-material BLEND_TEM and
 * URFC_Tem at [m-1]; flux as for oracle_newuvw.  Returns -1 where a face's
 * max(BLEND_TEM) is NO next to YES or an uncovered code; abou is
 * (-NbC:-1); cbuf, xbuf hold NbC+NC+1 doubles.  Returns 0.
 */
int oracle_calczmix(int nc, int nbc, int ns, int nonz, int convec, int diffus, int cross, int inert,
                    const int *sidec, const int *sidaij, const int *adia, const int *typebc, const int *material,
                    const int *blend_tem, const double *urfc_tem, const double *f, const double *fF,
                    const double *scoef, const double *flux_u, const double *flux, const double *sx,
                    const double *sy, const double *sz,
                    const double *dx, const double *dy, const double *dz, const double *volume,
                    const double *rho_n, const double *rho_o, const double *rho_oo, const double *z_n,
                    const double *z_o, const double *z_oo, const double *source, const double *dZdx,
                    const double *dZdy, const double *dZdz, double visc, int visc_rho, double dt, double urf,
                    double *aval,
                    double *b, double *abou, double *rhoz, double *cbuf, double *xbuf)
{
#define C(x, c) (x)[nbc + (c)]
#define MAT(arr, c) (arr)[C(material, c) - 1]
    const int nall = nbc + nc + 1;
    for (int c = -nbc; c <= nc; ++c) C(rhoz, c) = C(rho_n, c) * C(z_n, c);    /* :100 */
    for (int c = 0; c < nbc; ++c) abou[c] = 0.0;                        /* :104 */
    for (int j = 0; j < nonz; ++j) aval[j] = 0.0;                       /* :106 */
    for (int c = 0; c < nc; ++c) b[c] = 0.0;                            /* :107 */
    for (int c = 0; c < nall; ++c) { cbuf[c] = 0.0; xbuf[c] = 0.0; }   /* :134-135 */

    /* PHImax / PHImin of rhoZ when any face takes a ConvScheme scheme
     * (:126-131) */
    double *phimax = NULL, *phimin = NULL;
    for (int s = 1; s <= ns; ++s) {
        const int c1 = sidec[3 * (s - 1) + 1], c2 = sidec[3 * (s - 1) + 2];
        if (MAT(blend_tem, c1) != BL_NO || MAT(blend_tem, c2) != BL_NO) {
            const int bl = MAT(blend_tem, c1) > MAT(blend_tem, c2) ? MAT(blend_tem, c1) : MAT(blend_tem, c2);
            if (!blend_covered(bl)) {                                   /* NO next to YES: Phij undefined */
                free(phimax);
                free(phimin);
                return -1;
            }
            if (bl != BL_YES && !phimax) {
                phimax = (double *)malloc(sizeof(double) * nall);
                phimin = (double *)malloc(sizeof(double) * nall);
                oracle_calminmax(nc, nbc, ns, sidec, typebc, rhoz, phimax, phimin);
            }
        }
    }

    for (int s = 1; s <= ns; ++s) {                                     /* :155-189 */
        const int c1 = sidec[3 * (s - 1) + 1], c2 = sidec[3 * (s - 1) + 2];
        double rhoZS = f[s - 1] * C(rhoz, c1) + (1.0 - f[s - 1]) * C(rhoz, c2);
        const int up = MAT(blend_tem, c1) != BL_NO || MAT(blend_tem, c2) != BL_NO;
        if (up) {                                                       /* :163-166 */
            const int bl = MAT(blend_tem, c1) > MAT(blend_tem, c2) ? MAT(blend_tem, c1) : MAT(blend_tem, c2);
            rhoZS = conv_scheme(rhoZS, s, c1, c2, nbc, rhoz, dZdx, dZdy, dZdz, dx, dy, dz, f, flux, phimax, phimin,
                                bl);
        }
        const double F = flux_u[s - 1];
        if (c2 > 0) {
            C(cbuf, c1) = C(cbuf, c1) - F * rhoZS;
            C(cbuf, c2) = C(cbuf, c2) + F * rhoZS;
        } else {
            C(cbuf, c1) = C(cbuf, c1) - F * rhoZS;
        }
        if (up) {
            if (F < 0) {
                C(xbuf, c1) = C(xbuf, c1) - F * C(rhoz, c2);
                if (c2 > 0) C(xbuf, c2) = C(xbuf, c2) + F * C(rhoz, c2);
            } else {
                C(xbuf, c1) = C(xbuf, c1) - F * C(rhoz, c1);
                if (c2 > 0) C(xbuf, c2) = C(xbuf, c2) + F * C(rhoz, c1);
            }
        }
    }
    if (convec == FL_FI)                                                /* :201-207 */
        for (int c = 1; c <= nc; ++c)
            b[c - 1] = b[c - 1] + MAT(urfc_tem, c) * (C(cbuf, c) - C(xbuf, c));
    for (int c = 1; c <= nc; ++c) C(xbuf, c) = 0.0;                     /* :217-219 */

    for (int s = 1; s <= ns; ++s) {                                     /* :224-341 */
        double A12 = 0.0, A21 = 0.0, D12 = 0.0, D21 = 0.0;
        const int c1 = sidec[3 * (s - 1) + 1], c2 = sidec[3 * (s - 1) + 2];
        const double ff = fF[s - 1], gg = 1.0 - fF[s - 1];
        const double RhoS = ff * C(rho_n, c1) + gg * C(rho_n, c2);       /* NewUVW :285, CalcZmix :232 */
        const double VISeff = visc_rho ? visc * RhoS : visc;            /* Moin_Problem() == 4 */
        const double dZdxS = ff * C(dZdx, c1) + gg * C(dZdx, c2);
        const double dZdyS = ff * C(dZdy, c1) + gg * C(dZdy, c2);
        const double dZdzS = ff * C(dZdz, c1) + gg * C(dZdz, c2);
        const double Fex = VISeff * ((dZdxS * sx[s - 1] + dZdyS * sy[s - 1]) + dZdzS * sz[s - 1]);   /* :250 */
        const double A0 = VISeff * scoef[s - 1];
        const double Fim = A0 * ((dZdxS * dx[s - 1] + dZdyS * dy[s - 1]) + dZdzS * dz[s - 1]);     /* :255-258 */
        C(xbuf, c1) = C(xbuf, c1) + Fex - Fim;                          /* :261-264 */
        if (c2 > 0) C(xbuf, c2) = C(xbuf, c2) - Fex + Fim;
        if (diffus == FL_CN || diffus == FL_FI) {
            if (diffus == FL_FI) {
                D12 = D12 + A0 / C(rho_n, c2);
                D21 = D21 + A0 / C(rho_n, c1);
            }
            if (MAT(blend_tem, c1) != BL_NO || MAT(blend_tem, c2) != BL_NO) {
                const double F = flux_u[s - 1];
                A12 = A12 - (F < 0.0 ? F : 0.0);
                A21 = A21 + (F > 0.0 ? F : 0.0);
            }
            if (c2 > 0) {                                               /* "m1" */
                aval[adia[c1 - 1] - 1] = aval[adia[c1 - 1] - 1] + A21 + D21;
                aval[sidaij[2 * (s - 1) + 1] - 1] = aval[sidaij[2 * (s - 1) + 1] - 1] - A21 - D21;
                aval[adia[c2 - 1] - 1] = aval[adia[c2 - 1] - 1] + A12 + D12;
                aval[sidaij[2 * (s - 1)] - 1] = aval[sidaij[2 * (s - 1)] - 1] - A12 - D12;
            } else if (c2 < 0) {
                const int t = C(typebc, c2);
                if (t == BC_INFLOW || t == BC_WALL || t == BC_CONVECT || t == BC_WALLFL) {
                    aval[adia[c1 - 1] - 1] = aval[adia[c1 - 1] - 1] + A21 + D21;
                    b[c1 - 1] = b[c1 - 1] + (A12 + D12) * C(rhoz, c2);
                } else if (t == BC_OUTFLOW) {
                    aval[adia[c1 - 1] - 1] = aval[adia[c1 - 1] - 1] + A21;
                    b[c1 - 1] = b[c1 - 1] + A12 * C(rhoz, c2);
                } else if (t == BC_BUFFER) {
                    aval[adia[c1 - 1] - 1] = aval[adia[c1 - 1] - 1] + A21 + D21;
                    C(abou, c2) = -A12 - D12;
                }
            }
        }
    }
    if (cross == FL_FI)                                                 /* :359-361 */
        for (int c = 1; c <= nc; ++c) b[c - 1] = b[c - 1] + C(xbuf, c);
    if (inert == IN_LIN)                                                /* :369-372 */
        for (int c = 1; c <= nc; ++c) {
            aval[adia[c - 1] - 1] = aval[adia[c - 1] - 1] + C(volume, c) / dt;
            b[c - 1] = b[c - 1] + C(volume, c) / dt * C(z_o, c) * C(rho_o, c) + C(source, c) * C(volume, c);
        }
    if (inert == IN_PAR)                                                /* :375-378 */
        for (int c = 1; c <= nc; ++c) {
            aval[adia[c - 1] - 1] = aval[adia[c - 1] - 1] + 1.5 * C(volume, c) / dt;
            b[c - 1] = b[c - 1] + 2.0 * C(volume, c) / dt * C(rho_o, c) * C(z_o, c) -
                       0.5 * C(volume, c) / dt * C(rho_oo, c) * C(z_oo, c) + C(source, c) * C(volume, c);
        }
    for (int c = 1; c <= nc; ++c) {                                     /* :386-387 */
        b[c - 1] = b[c - 1] + aval[adia[c - 1] - 1] * (1.0 - urf) * C(rhoz, c) / urf;
        aval[adia[c - 1] - 1] = aval[adia[c - 1] - 1] / urf;
    }
    free(phimax);
    free(phimin);
    return 0;
#undef MAT
#undef C
}

/* Zmix%n(1:NC) = rhoZ / rho%n, then limited to [0, 1] (CalcZmix.f90:401-405):
 * z = min(max(rhoZ / rho, 0), 1), with max and min taken as (a < b ? b : a)
 * and (a > b ? b : a). */
void oracle_zmix_update(int nc, int nbc, const double *rhoz, const double *rho_n, double *z_n)
{
    for (int c = 1; c <= nc; ++c) {
        double z = rhoz[nbc + c] / rho_n[nbc + c];
        z = z < 0.0 ? 0.0 : z;
        z = z > 1.0 ? 1.0 : z;
        z_n[nbc + c] = z;
    }
}
