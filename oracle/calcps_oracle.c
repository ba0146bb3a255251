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
 * after the pressure solve; the CFL / Peclet monitor of :126-147 only feeds
 * the printed monitor lines and is left out):
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
