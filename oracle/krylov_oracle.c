/*
 * krylov_oracle.c -- CPU restatement of the T-FlowS Krylov solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load it, and only to compare against or to time as a baseline.  The
 * product library never links or calls it.
 *
 * It restates, statement by statement and in the same floating-point
 * operation order, the single-domain (Sequential build, NPro = 0) solvers:
 *
 *   oracle_prec1_  <- Process/Prec1.f90:45-66
 *   oracle_resid_  <- Process/Resid.f90:29-34 (buffer loop :36-43 is dead
 *                     with NPro = 0)
 *   oracle_cg_     <- Process/CG.f90:55-262
 *   oracle_bicg_   <- Process/BiCG.f90:43-254
 *   oracle_cgs_    <- Process/CGS.f90:55-337
 *
 * with GloSum / Exchng collapsed to no-ops (Sequential/GloSum.f90,
 * Sequential/Exchng.f90).  The argument lists are the Fortran ones, passed
 * by reference, with the reference's array bases: x and Ab are declared
 * x(-NB:N), Ab(-NB:-1), so Fortran x(i) is x[NB + i]; every other array is
 * 1-based, A(j) is A[j - 1]; the stored index values are 1-based.
 *
 * Semantics kept on purpose (they are the reference's, not bugs of this
 * file): the tolerance is absolute on ||r||_2; r1 enters as b and leaves as
 * the final residual; niter leaves as the iteration count, or cap + 1 when
 * the loop runs out (Fortran DO semantics); on the early exit when
 * ||b|| < tol IniRes is not written and FinRes receives the value of the
 * module variable `error` left by the previous call (Library/sol_mod.f90:11);
 * BiCG multiplies its shadow by A, not A^T (BiCG.f90:196-200); the DIC
 * pivots use a_ik^2 (Prec1.f90:56); the Jacobi apply is a true division.
 *
 * Build with -ffp-contract=off and no fast-math so that no multiply-add is
 * fused: the reference is compiled for baseline x86-64, which has no FMA.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define TINY 1.0e-64 /* Library/allp_mod.f90:31 */

/* sol_mod's `error` persists between calls (Library/sol_mod.f90:11). */
static double g_error = 0.0;

void oracle_reset_state(void) { g_error = 0.0; }
double oracle_get_error(void) { return g_error; }
void oracle_set_error(double e) { g_error = e; }

/* 1-based accessors */
#define AV(j) A[(j) - 1]
#define ACOL(i) Acol[(i) - 1]
#define AROW(j) Arow[(j) - 1]
#define ADIA(i) Adia[(i) - 1]
#define X(i) x[nb + (i)]
#define V1(v, i) (v)[(i) - 1]

void oracle_prec1_(const int *N, const int *NONZ, const double *A, const int *Acol,
                   const int *Arow, const int *Adia, double *D, const int *prec)
{
    const int n = *N;
    (void)NONZ;
    if (*prec == 1) { /* Prec1.f90:45-48 */
        for (int i = 1; i <= n; ++i) V1(D, i) = AV(ADIA(i));
    } else if (*prec == 2) { /* Prec1.f90:51-59 */
        for (int i = 1; i <= n; ++i) {
            double sum1 = AV(ADIA(i));
            for (int j = ACOL(i); j <= ADIA(i) - 1; ++j) {
                int k = AROW(j);
                sum1 = sum1 - V1(D, k) * AV(j) * AV(j);
            }
            V1(D, i) = 1.0 / sum1;
        }
    } else { /* Prec1.f90:63-66 */
        for (int i = 1; i <= n; ++i) V1(D, i) = 1.0;
    }
}

void oracle_resid_(const int *N, const int *NB, const int *NONZ, const double *A,
                   const int *Acol, const int *Arow, const double *Ab, const double *x,
                   double *r1)
{
    const int n = *N, nb = *NB;
    (void)NONZ; (void)Ab;
    for (int i = 1; i <= n; ++i)          /* Resid.f90:29-34 */
        for (int j = ACOL(i); j <= ACOL(i + 1) - 1; ++j) {
            int k = AROW(j);
            V1(r1, i) = V1(r1, i) - AV(j) * X(k);
        }
}

/* Preconditioner apply z = M^-1 u, in the reference's three forms
 * (CG.f90:124-165).  `u` and `z` are 1-based views of length n. */
static void prec_apply(int n, const double *A, const int *Acol, const int *Arow,
                       const int *Adia, const double *D, int prec, const double *u, double *z)
{
    if (prec == 1) {
        for (int i = 1; i <= n; ++i) V1(z, i) = V1(u, i) / V1(D, i);
    } else if (prec == 2) {
        for (int i = 1; i <= n; ++i) {          /* forward, CG.f90:135-142 */
            double sum1 = V1(u, i);
            for (int j = ACOL(i); j <= ADIA(i) - 1; ++j) {
                int k = AROW(j);
                sum1 = sum1 - AV(j) * V1(z, k);
            }
            V1(z, i) = sum1 * V1(D, i);
        }
        for (int i = 1; i <= n; ++i)            /* CG.f90:144-146 */
            V1(z, i) = V1(z, i) / (V1(D, i) + TINY);
        for (int i = n; i >= 1; --i) {          /* backward, CG.f90:149-156 */
            double sum1 = V1(z, i);
            for (int j = ADIA(i) + 1; j <= ACOL(i + 1) - 1; ++j) {
                int k = AROW(j);
                sum1 = sum1 - AV(j) * V1(z, k);
            }
            V1(z, i) = sum1 * V1(D, i);
        }
    } else {
        for (int i = 1; i <= n; ++i) V1(z, i) = V1(u, i);
    }
}

/* q = A p over full rows, accumulating from 0.0 in row order (CG.f90:193-199).
 * p is a 1-based view. */
static void spmv(int n, const double *A, const int *Acol, const int *Arow, const double *p,
                 double *q)
{
    for (int i = 1; i <= n; ++i) {
        double s = 0.0;
        for (int j = ACOL(i); j <= ACOL(i + 1) - 1; ++j) {
            int k = AROW(j);
            s = s + AV(j) * V1(p, k);
        }
        V1(q, i) = s;
    }
}

static double norm2(int n, const double *v)
{
    double s = 0.0;
    for (int i = 1; i <= n; ++i) s = s + V1(v, i) * V1(v, i);
    return sqrt(s);
}

static double *alloc0(int n) { return (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

void oracle_cg_(const int *N, const int *NB, const int *NONZ, const double *A, const int *Acol,
                const int *Arow, const int *Adia, const double *Ab, double *x, double *r1,
                const int *prec, int *niter, const double *tol, double *IniRes,
                double *FinRes)
{
    const int n = *N, nb = *NB, cap = *niter, pr = *prec;
    const double tl = *tol;
    double *D = alloc0(n), *p1 = alloc0(n), *q1 = alloc0(n);
    double rho = 0.0, rhoold = 0.0, alfa, beta, bnrm2;
    int iter;

    oracle_prec1_(N, NONZ, A, Acol, Arow, Adia, D, prec);        /* CG.f90:55 */
    bnrm2 = 0.0;                                                  /* CG.f90:62-67 */
    for (int i = 1; i <= n; ++i) bnrm2 = bnrm2 + V1(r1, i) * V1(r1, i);
    bnrm2 = sqrt(bnrm2);
    if (bnrm2 < tl) { iter = 0; goto done; }                      /* CG.f90:69-72 */

    oracle_resid_(N, NB, NONZ, A, Acol, Arow, Ab, x, r1);         /* CG.f90:78 */
    for (int i = 1; i <= n; ++i) V1(p1, i) = V1(r1, i);           /* CG.f90:85-87 */
    g_error = norm2(n, r1);                                       /* CG.f90:93-98 */
    *IniRes = g_error;
    if (g_error < tl) { iter = 0; goto done; }

    for (iter = 1; iter <= cap; ++iter) {                         /* CG.f90:113 */
        prec_apply(n, A, Acol, Arow, Adia, D, pr, r1, q1);        /* z in q1 */
        rho = 0.0;                                                /* CG.f90:172-176 */
        for (int i = 1; i <= n; ++i) rho = rho + V1(r1, i) * V1(q1, i);
        if (iter == 1) {
            for (int i = 1; i <= n; ++i) V1(p1, i) = V1(q1, i);
        } else {
            beta = rho / rhoold;                                  /* CG.f90:183-186 */
            for (int i = 1; i <= n; ++i) V1(p1, i) = V1(q1, i) + beta * V1(p1, i);
        }
        spmv(n, A, Acol, Arow, p1, q1);                           /* CG.f90:193-199 */
        alfa = 0.0;                                               /* CG.f90:214-219 */
        for (int i = 1; i <= n; ++i) alfa = alfa + V1(p1, i) * V1(q1, i);
        alfa = rho / alfa;
        for (int i = 1; i <= n; ++i) {                            /* CG.f90:226-229 */
            X(i) = X(i) + alfa * V1(p1, i);
            V1(r1, i) = V1(r1, i) - alfa * V1(q1, i);
        }
        g_error = norm2(n, r1);                                   /* CG.f90:248-253 */
        if (g_error < tl) goto done;
        rhoold = rho;
    }
done:
    *FinRes = g_error;                                            /* CG.f90:261-262 */
    *niter = iter;
    free(D); free(p1); free(q1);
}

void oracle_bicg_(const int *N, const int *NB, const int *NONZ, const double *A,
                  const int *Acol, const int *Arow, const int *Adia, const double *Ab,
                  double *x, double *r1, const int *prec, int *niter, const double *tol,
                  double *IniRes, double *FinRes)
{
    const int n = *N, nb = *NB, cap = *niter, pr = *prec;
    const double tl = *tol;
    double *D = alloc0(n), *p1 = alloc0(n), *p2 = alloc0(n), *q1 = alloc0(n),
           *q2 = alloc0(n), *r2 = alloc0(n);
    double rho = 0.0, rhoold = 0.0, alfa, beta, bnrm2;
    int iter;

    oracle_prec1_(N, NONZ, A, Acol, Arow, Adia, D, prec);        /* BiCG.f90:43 */
    bnrm2 = 0.0;
    for (int i = 1; i <= n; ++i) bnrm2 = bnrm2 + V1(r1, i) * V1(r1, i);
    bnrm2 = sqrt(bnrm2);
    if (bnrm2 < tl) { iter = 0; goto done; }                      /* BiCG.f90:57-60 */

    oracle_resid_(N, NB, NONZ, A, Acol, Arow, Ab, x, r1);         /* BiCG.f90:66 */
    g_error = norm2(n, r1);                                       /* BiCG.f90:72-77 */
    *IniRes = g_error;
    if (g_error < tl) { iter = 0; goto done; }
    for (int i = 1; i <= n; ++i) V1(r2, i) = V1(r1, i);           /* BiCG.f90:93-95 */

    for (iter = 1; iter <= cap; ++iter) {
        /* BiCG.f90:112-163: both systems share D and the same sweeps; when
         * prec == 2 the reference interleaves the two sweeps row by row,
         * which is arithmetically the same as two separate sweeps. */
        prec_apply(n, A, Acol, Arow, Adia, D, pr, r1, q1);
        prec_apply(n, A, Acol, Arow, Adia, D, pr, r2, q2);
        rho = 0.0;                                                /* BiCG.f90:169-173 */
        for (int i = 1; i <= n; ++i) rho = rho + V1(q1, i) * V1(r2, i);
        if (iter == 1) {
            for (int i = 1; i <= n; ++i) { V1(p1, i) = V1(q1, i); V1(p2, i) = V1(q2, i); }
        } else {
            beta = rho / rhoold;                                  /* BiCG.f90:181-185 */
            for (int i = 1; i <= n; ++i) {
                V1(p1, i) = V1(q1, i) + beta * V1(p1, i);
                V1(p2, i) = V1(q2, i) + beta * V1(p2, i);
            }
        }
        spmv(n, A, Acol, Arow, p1, q1);                           /* BiCG.f90:193-201: */
        spmv(n, A, Acol, Arow, p2, q2);                           /* shadow uses A, not A^T */
        alfa = 0.0;                                               /* BiCG.f90:218-223 */
        for (int i = 1; i <= n; ++i) alfa = alfa + V1(p2, i) * V1(q1, i);
        alfa = rho / alfa;
        for (int i = 1; i <= n; ++i) {                            /* BiCG.f90:230-234 */
            X(i) = X(i) + alfa * V1(p1, i);
            V1(r1, i) = V1(r1, i) - alfa * V1(q1, i);
            V1(r2, i) = V1(r2, i) - alfa * V1(q2, i);
        }
        g_error = norm2(n, r1);                                   /* BiCG.f90:240-245 */
        if (g_error < tl) goto done;
        rhoold = rho;
    }
done:
    *FinRes = g_error;
    *niter = iter;
    free(D); free(p1); free(p2); free(q1); free(q2); free(r2);
}

void oracle_cgs_(const int *N, const int *NB, const int *NONZ, const double *A,
                 const int *Acol, const int *Arow, const int *Adia, const double *Ab,
                 double *x, double *r1, const int *prec, int *niter, const double *tol,
                 double *IniRes, double *FinRes)
{
    const int n = *N, nb = *NB, cap = *niter, pr = *prec;
    const double tl = *tol;
    double *D = alloc0(n), *p1 = alloc0(n), *p2 = alloc0(n), *q1 = alloc0(n),
           *q2 = alloc0(n), *r2 = alloc0(n), *u1 = alloc0(n), *u2 = alloc0(n),
           *v2 = alloc0(n), *w = alloc0(n);
    double rho = 0.0, rhoold = 0.0, alfa, beta, bnrm2;
    int iter;

    oracle_prec1_(N, NONZ, A, Acol, Arow, Adia, D, prec);        /* CGS.f90:55 */
    bnrm2 = 0.0;
    for (int i = 1; i <= n; ++i) bnrm2 = bnrm2 + V1(r1, i) * V1(r1, i);
    bnrm2 = sqrt(bnrm2);
    if (bnrm2 < tl) { iter = 0; goto done; }                      /* CGS.f90:69-72 */

    oracle_resid_(N, NB, NONZ, A, Acol, Arow, Ab, x, r1);         /* CGS.f90:78 */
    for (int i = 1; i <= n; ++i) V1(r2, i) = V1(r1, i);           /* CGS.f90:84-86 */
    g_error = norm2(n, r1);                                       /* CGS.f90:92-97 */
    *IniRes = g_error;
    if (g_error < tl) { iter = 0; goto done; }

    for (iter = 1; iter <= cap; ++iter) {                         /* CGS.f90:112 */
        rho = 0.0;                                                /* CGS.f90:118-122 */
        for (int i = 1; i <= n; ++i) rho = rho + V1(r1, i) * V1(r2, i);
        if (iter == 1) {                                          /* CGS.f90:124-128 */
            for (int i = 1; i <= n; ++i) { V1(u1, i) = V1(r1, i); V1(u2, i) = V1(u1, i); }
        } else {                                                  /* CGS.f90:130-134 */
            beta = rho / rhoold;
            for (int i = 1; i <= n; ++i) {
                V1(u1, i) = V1(r1, i) + beta * V1(q1, i);
                V1(u2, i) = V1(u1, i) + beta * (V1(q1, i) + beta * V1(u2, i));
            }
        }
        prec_apply(n, A, Acol, Arow, Adia, D, pr, u2, p2);        /* CGS.f90:145-187 */
        spmv(n, A, Acol, Arow, p2, v2);                           /* CGS.f90:193-198 */
        /* CGS.f90:199 accumulates into alfa inside the SpMV loop; the value is
         * overwritten at :215 before any use, so it is not restated. */
        alfa = 0.0;                                               /* CGS.f90:215-220 */
        for (int i = 1; i <= n; ++i) alfa = alfa + V1(r2, i) * V1(v2, i);
        alfa = rho / alfa;
        for (int i = 1; i <= n; ++i) V1(q1, i) = V1(u1, i) - alfa * V1(v2, i); /* :226-228 */
        for (int i = 1; i <= n; ++i) V1(w, i) = V1(u1, i) + V1(q1, i);          /* :240/250 */
        prec_apply(n, A, Acol, Arow, Adia, D, pr, w, p1);        /* CGS.f90:238-280 */
        for (int i = 1; i <= n; ++i) X(i) = X(i) + alfa * V1(p1, i);            /* :286-288 */
        spmv(n, A, Acol, Arow, p1, q2);                           /* CGS.f90:294-300 */
        for (int i = 1; i <= n; ++i) V1(r1, i) = V1(r1, i) - alfa * V1(q2, i);  /* :315-317 */
        g_error = norm2(n, r1);                                   /* CGS.f90:323-328 */
        if (g_error < tl) goto done;
        rhoold = rho;
    }
done:
    *FinRes = g_error;
    *niter = iter;
    free(D); free(p1); free(p2); free(q1); free(q2); free(r2);
    free(u1); free(u2); free(v2); free(w);
}

/* --------------------------------------------------------------------------
 * Checker helpers beyond the reference (used by the tests):
 * SpMV and the DIC apply as stand-alone functions, with the same order of
 * operations as inside the solvers, on 1-based T-FlowS arrays.
 * ------------------------------------------------------------------------ */
void oracle_spmv(int n, const double *A, const int *Acol, const int *Arow, const double *p,
                 double *q)
{
    spmv(n, A, Acol, Arow, p, q);
}

void oracle_prec_apply(int n, const double *A, const int *Acol, const int *Arow,
                       const int *Adia, const double *D, int prec, const double *u, double *z)
{
    prec_apply(n, A, Acol, Arow, Adia, D, prec, u, z);
}
