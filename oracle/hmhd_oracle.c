/*
 * hmhd_oracle.c -- CPU restatement of the hosono-2DMHD per-step update.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product library never links
 * or calls it.
 *
 * It restates, in plain C with the same loop orders, the same left-to-right
 * operation order and the same MOOD bookkeeping, the reference routines
 * (paths relative to the reference repository):
 *
 *   src/eos.f90:9-26                eos_pres            (pressure floor)
 *   src/convert.f90:28-92           cons2flux
 *   src/convert.f90:159-196         prim2cons
 *   src/convert.f90:198-234         cons2prim
 *   src/convert.f90:236-254         cons2pres           (unfloored)
 *   src/eigen.f90:257-321           eigen_cf, fast speed only (c_s and c_a
 *                                   feed eigenvalues HLLC never reads)
 *   src/riemannSolver.f90:23-221    face flux + HLLC-MHD (Li 2005)
 *   src/reconstruct.f90:18-175      GP face reconstruction + admissibility
 *   src/gridBlock.f90:194-235       lower_order
 *   src/getFluxes.f90:16-90         get_fluxes
 *   src/FLICDdivB.f90:13-70         flicd
 *   src/detect.f90:18-265           detect_all, CAD, PAD
 *   src/boundaryConditions.f90      bc_fill (outflow/periodic), periodic_mood
 *   src/timeStep.f90:22-171         forward Euler MOOD loop, SSP-RK3, advance
 *   src/cfl.f90:14-52               cfl_dt
 *   src/GP.f90:37-381 +
 *   src/linAlgQuadPrecision.f90:107-365   GP prediction vectors in binary128
 *
 * MAX/MIN follow the semantics of the reference as compiled by amdflang:
 * MAX(a,b) = a>b ? a : b, MIN(a,b) = a<b ? a : b (left fold for more
 * arguments), MINVAL/MAXVAL skip NaN.  Build with -ffp-contract=off so no
 * multiply-add is fused.
 *
 * Arrays use the reference's Fortran layout with 0-based padded indices:
 *   U(v,i,j)  -> U[(j*mx + i)*8 + v],   V(v,i,j) -> V[(j*mx + i)*10 + v]
 *   faceVals(v,q,f,i,j), flux(v,d,i,j), scheme(i,j), ...
 */
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "hmhd_oracle.h"

/* ---------------------------------------------------------------- */
/* indices (src/definitions.f90:11-35), shifted to 0-based            */
enum { NORTH = 0, SOUTH = 1, EAST = 2, WEST = 3 };
enum { DENS = 0, MOMX = 1, MOMY = 2, MOMZ = 3, ENER = 4, MAGX = 5, MAGY = 6, MAGZ = 7 };
enum { VELX = 1, VELY = 2, VELZ = 3, PRES = 4, EINT = 8, GAMM = 9 };
enum { NC = 8, NP = 10, NSMAX = 25, NQMAX = 4 };

static inline double fmaxF(double a, double b) { return a > b ? a : b; }
static inline double fminF(double a, double b) { return a < b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

struct orc {
    orc_cfg c;
    int ngc, mx, my, s_x, e_x, s_y, e_y;
    double dx, dy, mindxdy, maxdxdy;
    /* GP tables, per radius 1..3 */
    int ns[4], nq[4];
    int st[4][NSMAX][2];
    double pv[4][4][NQMAX][NSMAX]; /* [rr][face][q][k] */
    double qw[4][NQMAX];           /* [rr][q] */
    /* state */
    double *U, *V;
    /* block scratch (src/gridBlock.f90:33-90) */
    double *faceVals; /* (8, NQMAX, 4, i, j) */
    double *flux, *fluxRaw; /* (8, 2, i, j) */
    int *scheme;
    unsigned char *rc, *rcF;
    /* work arrays */
    unsigned char *cellAdm, *recCell, *recNext, *checkCell, *lowerCell;
    double *Us[4];
    /* statistics of the last advance */
    int passes[3], inner[3];
    int *stage_scheme; /* 3 * mx*my */
};

#define IDX(o, i, j) ((size_t)(j) * (size_t)(o)->mx + (size_t)(i))
#define UA(o, A, v, i, j) (A)[IDX(o, i, j) * NC + (v)]
#define FV(o, v, q, f, i, j) (o)->faceVals[((IDX(o, i, j) * 4 + (f)) * NQMAX + (q)) * NC + (v)]
#define FL(o, A, v, d, i, j) (A)[(IDX(o, i, j) * 2 + (d)) * NC + (v)]

/* ---------------------------------------------------------------- */
/* eos / convert                                                      */

/* src/eos.f90:9-26 */
static double eos_pres(const orc_cfg *c, double dens, double eint, double gama) {
    double pres = (gama - 1.0) * dens * eint;
    if (pres < c->small_p) pres = c->small_p;
    return pres;
}

/* src/convert.f90:198-234 */
static void cons2prim(const orc_cfg *c, const double *U, double *V) {
    double Bp = 0.0;
    if (!c->force_hydro) Bp = 0.5 * (U[MAGX] * U[MAGX] + U[MAGY] * U[MAGY] + U[MAGZ] * U[MAGZ]);
    double dens = U[DENS];
    double vx = U[MOMX] / dens, vy = U[MOMY] / dens, vz = U[MOMZ] / dens;
    double eint = (U[ENER] - Bp) / dens - 0.5 * (vx * vx + vy * vy + vz * vz);
    double pres = eos_pres(c, dens, eint, c->gamma);
    V[DENS] = dens;
    V[VELX] = vx;
    V[VELY] = vy;
    V[VELZ] = vz;
    if (c->force_hydro) {
        V[MAGX] = V[MAGY] = V[MAGZ] = 0.0;
    } else {
        V[MAGX] = U[MAGX];
        V[MAGY] = U[MAGY];
        V[MAGZ] = U[MAGZ];
    }
    V[PRES] = pres;
    V[EINT] = eint;
    V[GAMM] = c->gamma;
}

/* src/convert.f90:236-254: pressure without the floor */
static double cons2pres(const orc_cfg *c, const double *U) {
    double Bp = 0.0;
    if (!c->force_hydro) Bp = 0.5 * (U[MAGX] * U[MAGX] + U[MAGY] * U[MAGY] + U[MAGZ] * U[MAGZ]);
    return (c->gamma - 1.0) *
           (U[ENER] - Bp - 0.5 * (U[MOMX] * U[MOMX] + U[MOMY] * U[MOMY] + U[MOMZ] * U[MOMZ]) / U[DENS]);
}

/* src/convert.f90:159-196 */
static void prim2cons(const orc_cfg *c, const double *V, double *U) {
    double Bp;
    if (c->force_hydro)
        Bp = 0.0;
    else
        Bp = 0.5 * (V[MAGX] * V[MAGX] + V[MAGY] * V[MAGY] + V[MAGZ] * V[MAGZ]);
    double rhoEint = V[PRES] / (c->gamma - 1.0);
    double E = rhoEint + V[DENS] * (0.5 * (V[VELX] * V[VELX] + V[VELY] * V[VELY] + V[VELZ] * V[VELZ])) + Bp;
    U[DENS] = V[DENS];
    U[MOMX] = V[DENS] * V[VELX];
    U[MOMY] = V[DENS] * V[VELY];
    U[MOMZ] = V[DENS] * V[VELZ];
    if (c->force_hydro) {
        U[MAGX] = U[MAGY] = U[MAGZ] = 0.0;
    } else {
        U[MAGX] = V[MAGX];
        U[MAGY] = V[MAGY];
        U[MAGZ] = V[MAGZ];
    }
    U[ENER] = E;
}

/* src/convert.f90:28-92; dir 0 = x, 1 = y */
static void cons2flux(const orc_cfg *c, const double *U, int dir, double *F) {
    double Bp = 0.0;
    if (!c->force_hydro) Bp = 0.5 * (U[MAGX] * U[MAGX] + U[MAGY] * U[MAGY] + U[MAGZ] * U[MAGZ]);
    double E = U[ENER];
    double V[NP];
    cons2prim(c, U, V);
    double Ptot = V[PRES] + Bp;
    int mN, mT1, mT2 = MOMZ, bN, bT1, bT2 = MAGZ;
    if (dir == 0) {
        mN = MOMX; mT1 = MOMY; bN = MAGX; bT1 = MAGY;
    } else {
        mN = MOMY; mT1 = MOMX; bN = MAGY; bT1 = MAGX;
    }
    /* prim velocity/field indices coincide with the momentum/field ones */
    for (int v = 0; v < NC; v++) F[v] = 0.0;
    F[DENS] = U[mN];
    F[mN] = U[mN] * U[mN] / U[DENS] + Ptot;
    F[mT1] = U[mN] * V[mT1];
    F[mT2] = U[mN] * V[mT2];
    F[ENER] = (E + Ptot) * V[mN];
    if (!c->force_hydro) {
        F[mN] = F[mN] - V[bN] * V[bN];
        F[mT1] = F[mT1] - V[bT1] * V[bN];
        F[mT2] = F[mT2] - V[bT2] * V[bN];
        F[bT1] = V[mN] * V[bT1] - V[mT1] * V[bN];
        F[bT2] = V[mN] * V[bT2] - V[mT2] * V[bN];
        F[ENER] = F[ENER] - V[bN] * (V[VELX] * V[MAGX] + V[VELY] * V[MAGY] + V[VELZ] * V[MAGZ]);
    }
}

/* src/eigen.f90:305-321 */
static double eigen_cfs(double gam, double rho, double p, double BB, double BN, double pm) {
    double cfs = (gam * p + BB) / rho;
    cfs = 0.5 * (cfs + pm * sqrt(cfs * cfs - 4.0 * gam * p * (BN * BN) / (rho * rho)));
    cfs = sqrt(fmaxF(1.0e-14, cfs));
    return cfs;
}

/* src/eigen.f90:257-303: only eigenvalues 1 (v_N - c_f) and 7 (v_N + c_f) */
static void eigen_fast(const orc_cfg *c, const double *V, int dir, double *l1, double *l7) {
    int dN = dir;
    double rho = fmaxF(V[DENS], c->small_rho);
    double velN = V[VELX + dN];
    double BN = V[MAGX + dN];
    double p = fmaxF(V[PRES], c->small_p);
    double gam = V[GAMM];
    double BB = V[MAGX] * V[MAGX] + V[MAGY] * V[MAGY] + V[MAGZ] * V[MAGZ];
    double cf = eigen_cfs(gam, rho, p, BB, BN, +1.0);
    *l1 = velN - cf;
    *l7 = velN + cf;
}

/* src/riemannSolver.f90:68-221 */
static void hllc(const orc_cfg *c, const double *uL, const double *uR, int dir, double *flux) {
    double vL[NP], vR[NP], vA[NP];
    cons2prim(c, uL, vL);
    cons2prim(c, uR, vR);
    int mN, mT1, mT2 = MOMZ, bN, bT1, bT2 = MAGZ;
    if (dir == 0) {
        mN = MOMX; mT1 = MOMY; bN = MAGX; bT1 = MAGY;
    } else {
        mN = MOMY; mT1 = MOMX; bN = MAGY; bT1 = MAGX;
    }
    double eL1, eL7, eR1, eR7, eA1, eA7;
    eigen_fast(c, vL, dir, &eL1, &eL7);
    eigen_fast(c, vR, dir, &eR1, &eR7);
    for (int k = 0; k < NP; k++) vA[k] = (vL[k] + vR[k]) / 2.0;
    eigen_fast(c, vA, dir, &eA1, &eA7);
    double sL = fminF(eL1, eA1);
    double sR = fmaxF(eA7, eR7);

    double fL[NC], fR[NC], uHLL[NC], uStr[NC];
    cons2flux(c, uL, dir, fL);
    cons2flux(c, uR, dir, fR);
    for (int v = 0; v < NC; v++) uHLL[v] = (sR * uR[v] - sL * uL[v] + fL[v] - fR[v]) / (sR - sL);
    double velHLL[3], magHLL[3];
    for (int k = 0; k < 3; k++) {
        velHLL[k] = uHLL[MOMX + k] / uHLL[DENS];
        magHLL[k] = uHLL[MAGX + k];
    }
    double mdvHLL = magHLL[0] * velHLL[0] + magHLL[1] * velHLL[1] + magHLL[2] * velHLL[2];
    double mdvL = vL[MAGX] * vL[VELX] + vL[MAGY] * vL[VELY] + vL[MAGZ] * vL[VELZ];
    double mdvR = vR[MAGX] * vR[VELX] + vR[MAGY] * vR[VELY] + vR[MAGZ] * vR[VELZ];

    double velNL = vL[mN], velNR = vR[mN];
    double velT1L = vL[mT1], velT1R = vR[mT1];
    double velT2L = vL[mT2], velT2R = vR[mT2];
    double momNL = uL[mN], momNR = uR[mN];
    double magNL = vL[bN], magNR = vR[bN], magNHLL = uHLL[bN];
    double magT1L = vL[bT1], magT1R = vR[bT1], magT1HLL = uHLL[bT1];
    double magT2L = vL[bT2], magT2R = vR[bT2], magT2HLL = uHLL[bT2];
    double rhoL = vL[DENS], rhoR = vR[DENS];
    double pL = vL[PRES], pR = vR[PRES];
    double pTotL = pL + 0.5 * (vL[MAGX] * vL[MAGX] + vL[MAGY] * vL[MAGY] + vL[MAGZ] * vL[MAGZ]);
    double pTotR = pR + 0.5 * (vR[MAGX] * vR[MAGX] + vR[MAGY] * vR[MAGY] + vR[MAGZ] * vR[MAGZ]);
    double dL = sL - velNL, dR = sR - velNR;
    double sStr = (momNR * dR - momNL * dL + pTotL - pTotR - magNL * magNL + magNR * magNR) /
                  (rhoR * dR - rhoL * dL);
    double pStr = rhoL * dL * (sStr - velNL) + pTotL - magNL * magNL + magNHLL * magNHLL;

    if (0.0 < sL) {
        for (int v = 0; v < NC; v++) flux[v] = fL[v];
    } else if ((sL <= 0.0) && (0.0 < sStr)) {
        uStr[DENS] = rhoL * dL / (sL - sStr);
        uStr[mN] = uStr[DENS] * sStr;
        uStr[mT1] = uStr[DENS] * velT1L - (magNHLL * magT1HLL - magNL * magT1L) / (sL - sStr);
        uStr[mT2] = uStr[DENS] * velT2L - (magNHLL * magT2HLL - magNL * magT2L) / (sL - sStr);
        uStr[bN] = magNHLL;
        uStr[bT1] = magT1HLL;
        uStr[bT2] = magT2HLL;
        uStr[ENER] = uL[ENER] * dL + pStr * sStr - pTotL * velNL - magNHLL * mdvHLL + magNL * mdvL;
        uStr[ENER] = uStr[ENER] / (sL - sStr);
        for (int v = 0; v < NC; v++) flux[v] = fL[v] + sL * (uStr[v] - uL[v]);
    } else if ((sStr <= 0.0) && (0.0 <= sR)) {
        uStr[DENS] = rhoR * dR / (sR - sStr);
        uStr[mN] = uStr[DENS] * sStr;
        uStr[mT1] = uStr[DENS] * velT1R - (magNHLL * magT1HLL - magNR * magT1R) / (sR - sStr);
        uStr[mT2] = uStr[DENS] * velT2R - (magNHLL * magT2HLL - magNR * magT2R) / (sR - sStr);
        uStr[bN] = magNHLL;
        uStr[bT1] = magT1HLL;
        uStr[bT2] = magT2HLL;
        uStr[ENER] = uR[ENER] * dR + pStr * sStr - pTotR * velNR - magNHLL * mdvHLL + magNR * mdvR;
        uStr[ENER] = uStr[ENER] / (sR - sStr);
        for (int v = 0; v < NC; v++) flux[v] = fR[v] + sR * (uStr[v] - uR[v]);
    } else {
        for (int v = 0; v < NC; v++) flux[v] = fR[v];
    }
}

/* src/riemannSolver.f90:23-47: uL, uR are (8, NQMAX) column-major */
static void face_flux(const struct orc *o, const double *uL, const double *uR, int dir, int rr,
                      double *flux) {
    if (rr == 0) {
        hllc(&o->c, uL, uR, dir, flux);
    } else {
        double F[NC];
        for (int v = 0; v < NC; v++) flux[v] = 0.0;
        for (int q = 0; q < o->nq[rr]; q++) {
            hllc(&o->c, uL + q * NC, uR + q * NC, dir, F);
            for (int v = 0; v < NC; v++) flux[v] = flux[v] + o->qw[rr][q] * F[v];
        }
    }
}

/* src/detect.f90:238-265 */
static int detect_PAD(double dens, double pres) { return (dens > 0.0) && (pres > 0.0); }
static int detect_CAD(double dens, double pres) {
    return !(isnan(dens) || isnan(pres) || !isfinite(dens) || !isfinite(pres));
}

/* ---------------------------------------------------------------- */
/* GP prediction vectors in binary128 (src/GP.f90:37-381,             */
/* src/linAlgQuadPrecision.f90:107-365)                               */

typedef __float128 q_t;

/* src/GP.f90:216-280 */
static q_t gp_volvol(const q_t *p1, const q_t *p2, const q_t *dl, q_t ell) {
    const q_t sqrt_two = sqrtq(2.0Q), qpi = M_PIq;
    q_t sqrt2_ell = sqrt_two * ell;
    q_t two_ell_sq = 2.0Q * (ell * ell);
    q_t sqrt_pi_ell_sq = sqrtq(qpi) * (ell * ell);
    q_t kernel = 1.0Q;
    for (int i = 0; i < 2; i++) {
        q_t sep = p1[i] - p2[i];
        q_t r1 = (sep + dl[i]) / sqrt2_ell;
        q_t r2 = (sep - dl[i]) / sqrt2_ell;
        q_t r3 = -((sep + dl[i]) * (sep + dl[i]) / two_ell_sq);
        q_t r4 = -((sep - dl[i]) * (sep - dl[i]) / two_ell_sq);
        q_t r5 = sep / sqrt2_ell;
        q_t r6 = -(sep * sep / two_ell_sq);
        kernel = kernel * sqrt_pi_ell_sq / (dl[i] * dl[i]) *
                 (r1 * erfq(r1) + r2 * erfq(r2) + (expq(r3) + expq(r4)) / sqrtq(qpi) -
                  2.0Q * (r5 * erfq(r5) + expq(r6) / sqrtq(qpi)));
    }
    return kernel;
}

/* src/GP.f90:282-335 */
static q_t gp_volpoint(const q_t *p1, const q_t *p2, const q_t *dl, q_t ell) {
    const q_t sqrt_two = sqrtq(2.0Q), sqrt_half_pi = sqrtq(M_PIq / 2.0Q);
    q_t sqrt2_ell = sqrt_two * ell;
    q_t kernel = 1.0Q;
    for (int i = 0; i < 2; i++) {
        q_t half_delta = dl[i] / 2.0Q;
        q_t sep = p1[i] - p2[i];
        q_t r1 = (sep + half_delta) / sqrt2_ell;
        q_t r2 = (sep - half_delta) / sqrt2_ell;
        kernel = kernel * sqrt_half_pi * (ell / dl[i]) * (erfq(r1) - erfq(r2));
    }
    return kernel;
}

/* src/linAlgQuadPrecision.f90:107-144, :146-175, :206-235, :336-365.
 * A is m x m (row-major a[r*m+c]), B is m x n (row-major); X overwrites B. */
static int spd_solve(int m, int n, const q_t *A, q_t *B) {
    q_t R[NSMAX * NSMAX], Y[NSMAX * 4 * NQMAX];
    for (int r = 0; r < m; r++)
        for (int cc = 0; cc < m; cc++) R[r * m + cc] = (cc < r) ? 0.0Q : A[r * m + cc];
    for (int k = 0; k < m; k++) {
        if (!(R[k * m + k] > 0.0Q)) return -1;
        for (int j = k + 1; j < m; j++)
            for (int l = j; l < m; l++) R[j * m + l] = R[j * m + l] - R[k * m + l] * R[k * m + j] / R[k * m + k];
        q_t sk = sqrtq(R[k * m + k]);
        for (int l = k; l < m; l++) R[k * m + l] = R[k * m + l] / sk;
    }
    /* forward substitution with L = R^T: L(i,j) = R(j,i) */
    for (int cc = 0; cc < n; cc++) Y[0 * n + cc] = B[0 * n + cc] / R[0];
    for (int i = 1; i < m; i++) {
        for (int cc = 0; cc < n; cc++) {
            q_t ss = 0.0Q;
            for (int j = 0; j < i; j++) ss = ss + R[j * m + i] * Y[j * n + cc];
            Y[i * n + cc] = (B[i * n + cc] - ss) / R[i * m + i];
        }
    }
    /* back substitution with U = R */
    for (int cc = 0; cc < n; cc++) B[(m - 1) * n + cc] = Y[(m - 1) * n + cc] / R[(m - 1) * m + (m - 1)];
    for (int i = m - 2; i >= 0; i--) {
        for (int cc = 0; cc < n; cc++) {
            q_t ss = 0.0Q;
            for (int j = i + 1; j < m; j++) ss = ss + R[i * m + j] * B[j * n + cc];
            B[i * n + cc] = (Y[i * n + cc] - ss) / R[i * m + i];
        }
    }
    return 0;
}

/* src/GP.f90:37-200 */
static int gp_init(struct orc *o) {
    const orc_cfg *c = &o->c;
    q_t dl[2];
    dl[0] = ((q_t)c->xend - (q_t)c->xbeg) / (q_t)c->nx;
    dl[1] = ((q_t)c->yend - (q_t)c->ybeg) / (q_t)c->ny;
    q_t mindl = dl[0] < dl[1] ? dl[0] : dl[1];
    q_t ell = (q_t)c->ell_over_delta * mindl;
    memset(o->pv, 0, sizeof o->pv);
    memset(o->qw, 0, sizeof o->qw);
    memset(o->st, 0, sizeof o->st);
    for (int rr = 1; rr <= c->R; rr++) {
        int ns = 1 + 2 * rr * (rr + 1), nq = rr + 1, np = 4 * nq;
        o->ns[rr] = ns;
        o->nq[rr] = nq;
        int cnt = 0;
        for (int i = -rr; i <= rr; i++)
            for (int j = -rr; j <= rr; j++)
                if (abs(i) + abs(j) <= rr) {
                    o->st[rr][cnt][0] = i;
                    o->st[rr][cnt][1] = j;
                    cnt++;
                }
        q_t qpts[NQMAX];
        if (rr == 1) {
            qpts[0] = 1.0Q / 2.0Q / sqrtq(3.0Q);
            qpts[1] = -qpts[0];
            o->qw[rr][0] = 1.0 / 2.0;
            o->qw[rr][1] = o->qw[rr][0];
        } else if (rr == 2) {
            qpts[0] = 1.0Q / 2.0Q * sqrtq(3.0Q / 5.0Q);
            qpts[1] = 0.0Q;
            qpts[2] = -qpts[0];
            o->qw[rr][0] = 5.0 / 18.0;
            o->qw[rr][1] = 8.0 / 18.0;
            o->qw[rr][2] = o->qw[rr][0];
        } else {
            qpts[0] = 1.0Q / 2.0Q * sqrtq(3.0Q / 7.0Q + 2.0Q / 7.0Q * sqrtq(6.0Q / 5.0Q));
            qpts[1] = 1.0Q / 2.0Q * sqrtq(3.0Q / 7.0Q - 2.0Q / 7.0Q * sqrtq(6.0Q / 5.0Q));
            qpts[2] = -qpts[1];
            qpts[3] = -qpts[0];
            o->qw[rr][0] = (18.0 - sqrt(30.0)) / 72.0;
            o->qw[rr][1] = (18.0 + sqrt(30.0)) / 72.0;
            o->qw[rr][2] = o->qw[rr][1];
            o->qw[rr][3] = o->qw[rr][0];
        }
        q_t XX[NSMAX][2], XXs[4 * NQMAX][2];
        for (int k = 0; k < ns; k++) {
            XX[k][0] = (q_t)o->st[rr][k][0] * dl[0];
            XX[k][1] = (q_t)o->st[rr][k][1] * dl[1];
        }
        for (int q = 0; q < nq; q++) {
            XXs[NORTH * nq + q][0] = qpts[q] * dl[0];
            XXs[NORTH * nq + q][1] = dl[1] / 2.0Q;
            XXs[SOUTH * nq + q][0] = qpts[q] * dl[0];
            XXs[SOUTH * nq + q][1] = -dl[1] / 2.0Q;
            XXs[EAST * nq + q][0] = dl[0] / 2.0Q;
            XXs[EAST * nq + q][1] = qpts[q] * dl[1];
            XXs[WEST * nq + q][0] = -dl[0] / 2.0Q;
            XXs[WEST * nq + q][1] = qpts[q] * dl[1];
        }
        q_t Kcov[NSMAX * NSMAX], Kstr[NSMAX * 4 * NQMAX];
        for (int i = 0; i < ns; i++)
            for (int j = 0; j < ns; j++) Kcov[i * ns + j] = gp_volvol(XX[i], XX[j], dl, ell);
        for (int i = 0; i < ns; i++)
            for (int j = 0; j < np; j++) Kstr[i * np + j] = gp_volpoint(XX[i], XXs[j], dl, ell);
        if (spd_solve(ns, np, Kcov, Kstr) != 0) return -1;
        for (int p = 0; p < np; p++) {
            q_t sum = 0.0Q;
            for (int k = 0; k < ns; k++) sum = sum + Kstr[k * np + p];
            for (int k = 0; k < ns; k++) Kstr[k * np + p] = Kstr[k * np + p] / sum;
        }
        for (int f = 0; f < 4; f++)
            for (int q = 0; q < nq; q++)
                for (int k = 0; k < ns; k++) o->pv[rr][f][q][k] = (double)Kstr[k * np + f * nq + q];
    }
    return 0;
}

/* ---------------------------------------------------------------- */
/* block operations                                                   */

/* src/gridBlock.f90:194-235 */
static void lower_order(struct orc *o, int i, int j) {
    o->scheme[IDX(o, i, j)] = imax(o->scheme[IDX(o, i, j)] - 1, 0);
    o->rc[IDX(o, i, j)] = 1;
    o->rc[IDX(o, i + 1, j)] = 1;
    o->rc[IDX(o, i - 1, j)] = 1;
    o->rc[IDX(o, i, j + 1)] = 1;
    o->rc[IDX(o, i, j - 1)] = 1;
    for (int jj = imax(j - 2, 0); jj <= imin(j + 2, o->my - 1); jj++)
        for (int ii = imax(i - 2, 0); ii <= imin(i + 2, o->mx - 1); ii++)
            if (abs(ii - i) + abs(jj - j) <= 2) o->rcF[IDX(o, ii, jj)] = 1;
}

static const int FOFF[4][2] = {{0, 1}, {0, -1}, {1, 0}, {-1, 0}};

/* src/reconstruct.f90:18-175 */
static void reconstruct(struct orc *o, const double *U, int *inner_iters) {
    const orc_cfg *c = &o->c;
    size_t ncell = (size_t)o->mx * o->my;
    memcpy(o->recCell, o->rc, ncell);
    int admissible = 0;
    int it = 0;
    int j0 = o->s_y - 2, j1 = o->e_y + 2, i0 = o->s_x - 2, i1 = o->e_x + 2;
    while (!admissible) {
        it++;
        memset(o->cellAdm, 1, ncell);
#pragma omp parallel for collapse(2) schedule(static)
        for (int j = j0; j <= j1; j++)
            for (int i = i0; i <= i1; i++) {
                if (!o->recCell[IDX(o, i, j)]) continue;
                for (int f = 0; f < 4; f++) {
                    int ia = FOFF[f][0], ja = FOFF[f][1];
                    int rr = imin(o->scheme[IDX(o, i, j)], o->scheme[IDX(o, i + ia, j + ja)]);
                    if (rr == 0) {
                        for (int v = 0; v < NC; v++) FV(o, v, 0, f, i, j) = UA(o, U, v, i, j);
                    } else {
                        int ns = o->ns[rr], nq = o->nq[rr];
                        double sd[NSMAX][NC];
                        for (int k = 0; k < ns; k++)
                            for (int v = 0; v < NC; v++)
                                sd[k][v] = UA(o, U, v, i + o->st[rr][k][0], j + o->st[rr][k][1]);
                        for (int q = 0; q < nq; q++) {
                            for (int v = 0; v < NC; v++) FV(o, v, q, f, i, j) = 0.0;
                            for (int k = 0; k < ns; k++)
                                for (int v = 0; v < NC; v++)
                                    FV(o, v, q, f, i, j) = FV(o, v, q, f, i, j) + o->pv[rr][f][q][k] * sd[k][v];
                        }
                        for (int q = 0; q < nq; q++) {
                            double *fv = &FV(o, 0, q, f, i, j);
                            double pres = cons2pres(c, fv);
                            if (!(detect_CAD(fv[DENS], pres) && detect_PAD(fv[DENS], pres)))
                                o->cellAdm[IDX(o, i, j)] = 0;
                        }
                    }
                }
            }
        admissible = 1;
        memset(o->recNext, 0, ncell);
        for (int j = j0; j <= j1; j++)
            for (int i = i0; i <= i1; i++)
                if (!o->cellAdm[IDX(o, i, j)]) {
                    lower_order(o, i, j);
                    o->recNext[IDX(o, i - 1, j)] = 1;
                    o->recNext[IDX(o, i, j)] = 1;
                    o->recNext[IDX(o, i + 1, j)] = 1;
                    o->recNext[IDX(o, i, j - 1)] = 1;
                    o->recNext[IDX(o, i, j + 1)] = 1;
                    admissible = 0;
                }
        memcpy(o->recCell, o->recNext, ncell);
    }
    if (inner_iters) *inner_iters += it;
}

/* src/getFluxes.f90:16-90 */
static void get_fluxes(struct orc *o) {
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = o->s_y - 1; j <= o->e_y + 2; j++)
        for (int i = o->s_x - 1; i <= o->e_x + 2; i++) {
            if (o->rc[IDX(o, i, j)] || o->rc[IDX(o, i - 1, j)]) {
                int rr = imin(o->scheme[IDX(o, i, j)], o->scheme[IDX(o, i - 1, j)]);
                face_flux(o, &FV(o, 0, 0, EAST, i - 1, j), &FV(o, 0, 0, WEST, i, j), 0, rr,
                          &FL(o, o->fluxRaw, 0, 0, i, j));
                for (int v = 0; v < NC; v++) FL(o, o->flux, v, 0, i, j) = FL(o, o->fluxRaw, v, 0, i, j);
            }
            if (o->rc[IDX(o, i, j)] || o->rc[IDX(o, i, j - 1)]) {
                int rr = imin(o->scheme[IDX(o, i, j)], o->scheme[IDX(o, i, j - 1)]);
                face_flux(o, &FV(o, 0, 0, NORTH, i, j - 1), &FV(o, 0, 0, SOUTH, i, j), 1, rr,
                          &FL(o, o->fluxRaw, 0, 1, i, j));
                for (int v = 0; v < NC; v++) FL(o, o->flux, v, 1, i, j) = FL(o, o->fluxRaw, v, 1, i, j);
            }
        }
}

/* src/FLICDdivB.f90:13-70 */
static void flicd(struct orc *o) {
    double *R = o->fluxRaw;
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = o->s_y; j <= o->e_y + 1; j++)
        for (int i = o->s_x; i <= o->e_x + 1; i++) {
            if (!o->rcF[IDX(o, i, j)]) continue;
            FL(o, o->flux, MAGY, 0, i, j) =
                (2.0 * FL(o, R, MAGY, 0, i, j) + FL(o, R, MAGY, 0, i + 1, j) + FL(o, R, MAGY, 0, i - 1, j) -
                 FL(o, R, MAGX, 1, i, j + 1) - FL(o, R, MAGX, 1, i, j) - FL(o, R, MAGX, 1, i - 1, j) -
                 FL(o, R, MAGX, 1, i - 1, j + 1)) /
                8.0;
            FL(o, o->flux, MAGX, 1, i, j) =
                (2.0 * FL(o, R, MAGX, 1, i, j) + FL(o, R, MAGX, 1, i, j + 1) + FL(o, R, MAGX, 1, i, j - 1) -
                 FL(o, R, MAGY, 0, i + 1, j) - FL(o, R, MAGY, 0, i, j) - FL(o, R, MAGY, 0, i, j - 1) -
                 FL(o, R, MAGY, 0, i + 1, j - 1)) /
                8.0;
        }
    size_t ncell = (size_t)o->mx * o->my;
    for (size_t p = 0; p < ncell; p++) {
        o->flux[(p * 2 + 0) * NC + MAGX] = 0.0;
        o->flux[(p * 2 + 1) * NC + MAGY] = 0.0;
    }
}

/* NaN-skipping MINVAL / MAXVAL over n values */
static double minval(const double *x, int n) {
    double m = x[0];
    for (int k = 1; k < n; k++)
        if (isnan(m) || x[k] < m) m = x[k];
    return m;
}
static double maxval(const double *x, int n) {
    double m = x[0];
    for (int k = 1; k < n; k++)
        if (isnan(m) || x[k] > m) m = x[k];
    return m;
}

/* src/detect.f90:18-236 */
static int detect_all(struct orc *o, const double *U, const double *Unew) {
    const orc_cfg *c = &o->c;
    size_t ncell = (size_t)o->mx * o->my;
    int exit_loop = 1;
    memcpy(o->checkCell, o->rcF, ncell);
    memset(o->rc, 0, ncell);
    memset(o->rcF, 0, ncell);
    memset(o->lowerCell, 0, ncell);
    const double dx = o->dx, dy = o->dy;
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = o->s_y; j <= o->e_y; j++)
        for (int i = o->s_x; i <= o->e_x; i++) {
            if (!o->checkCell[IDX(o, i, j)]) continue;
            if (o->scheme[IDX(o, i, j)] <= 0) continue;
            const double *un = &UA(o, Unew, 0, i, j);
            double presNew = cons2pres(c, un);
            if (!detect_CAD(un[DENS], presNew)) {
                o->lowerCell[IDX(o, i, j)] = 1;
                continue;
            }
            if (!detect_PAD(un[DENS], presNew)) {
                o->lowerCell[IDX(o, i, j)] = 1;
                continue;
            }
            double V[NP], Vip[NP], Vim[NP], Vjp[NP], Vjm[NP];
            cons2prim(c, &UA(o, U, 0, i, j), V);
            cons2prim(c, &UA(o, U, 0, i + 1, j), Vip);
            cons2prim(c, &UA(o, U, 0, i - 1, j), Vim);
            cons2prim(c, &UA(o, U, 0, i, j + 1), Vjp);
            cons2prim(c, &UA(o, U, 0, i, j - 1), Vjm);
            double div_vel = (Vip[VELX] - Vim[VELX]) / (2.0 * dx);
            div_vel = div_vel + (Vjp[VELY] - Vjm[VELY]) / (2.0 * dy);
            double sigma_v = 1.0;
            int shock1 = div_vel < -(sigma_v * (o->maxdxdy * o->maxdxdy));
            double mach = V[VELX] * V[VELX] + V[VELY] * V[VELY];
            mach = V[DENS] * mach / (V[GAMM] * V[PRES]);
            mach = sqrt(mach);
            double sigma_M = 0.2;
            int shock2 = mach > sigma_M;
            double pg = fabs(Vip[PRES] - Vim[PRES]) / fminF(Vip[PRES], Vim[PRES]);
            pg = pg + fabs(Vjp[PRES] - Vjm[PRES]) / fminF(Vjp[PRES], Vjm[PRES]);
            double sigma_p = 10.0;
            int shock3 = pg > sigma_p * o->maxdxdy;
            if (!(shock1 && shock2 && shock3)) continue;

            double densNew = un[DENS];
            double d1 = UA(o, U, DENS, i, j), d2 = UA(o, U, DENS, i + 1, j), d3 = UA(o, U, DENS, i - 1, j);
            double d4 = UA(o, U, DENS, i, j + 1), d5 = UA(o, U, DENS, i, j - 1);
            double mn = fminF(fminF(fminF(fminF(d1, d2), d3), d4), d5);
            double mxv = fmaxF(fmaxF(fmaxF(fmaxF(d1, d2), d3), d4), d5);
            if (mxv - mn < o->mindxdy * o->mindxdy * o->mindxdy) continue;
            if ((mn <= densNew) && (densNew <= mxv)) continue;

            double dd[2][5];
            int cnt = 0;
            for (int ii = -1; ii <= 1; ii++)
                for (int jj = -1; jj <= 1; jj++)
                    if (abs(ii) + abs(jj) <= 1) {
                        dd[0][cnt] = (UA(o, U, DENS, i + ii + 1, j + jj) - 2.0 * UA(o, U, DENS, i + ii, j + jj) +
                                      UA(o, U, DENS, i + ii - 1, j + jj)) /
                                     (dx * dx);
                        dd[1][cnt] = (UA(o, U, DENS, i + ii, j + jj + 1) - 2.0 * UA(o, U, DENS, i + ii, j + jj) +
                                      UA(o, U, DENS, i + ii, j + jj - 1)) /
                                     (dy * dy);
                        cnt++;
                    }
            double mnx = minval(dd[0], 5), mny = minval(dd[1], 5);
            double mxx = maxval(dd[0], 5), mxy = maxval(dd[1], 5);
            int sx1 = mnx * mxx > -o->mindxdy;
            int sy1 = mny * mxy > -o->mindxdy;
            int sx2 = fmaxF(fabs(mnx), fabs(mxx)) < o->mindxdy;
            int sy2 = fmaxF(fabs(mny), fabs(mxy)) < o->mindxdy;
            int sx3 = fminF(fabs(mnx), fabs(mxx)) >= 0.5 * fmaxF(fabs(mnx), fabs(mxx));
            int sy3 = fminF(fabs(mny), fabs(mxy)) >= 0.5 * fmaxF(fabs(mny), fabs(mxy));
            int smooth1 = sx1 && sx2 && sy1 && sy2;
            int smooth2 = sx1 && sx3 && sy1 && sy3;
            if (!(smooth1 || smooth2)) o->lowerCell[IDX(o, i, j)] = 1;
        }
    for (int j = o->s_y; j <= o->e_y; j++)
        for (int i = o->s_x; i <= o->e_x; i++)
            if (o->lowerCell[IDX(o, i, j)]) {
                lower_order(o, i, j);
                exit_loop = 0;
            }
    return exit_loop;
}

/* src/boundaryConditions.f90:189-254 */
static void periodic_mood(struct orc *o) {
    int nx = o->c.nx, ny = o->c.ny;
    for (int j = 0; j < o->my; j++)
        for (int i = 0; i < o->mx; i++) {
            int ip;
            if (i < o->s_x)
                ip = i + nx;
            else if (i > o->e_x)
                ip = i - nx;
            else
                continue;
            size_t a = IDX(o, i, j), b = IDX(o, ip, j);
            unsigned char e = o->rcF[a] || o->rcF[b];
            o->rcF[a] = e;
            o->rcF[b] = e;
            e = o->rc[a] || o->rc[b];
            o->rc[a] = e;
            o->rc[b] = e;
            int ls = imin(o->scheme[a], o->scheme[b]);
            o->scheme[a] = ls;
            o->scheme[b] = ls;
        }
    for (int j = 0; j < o->my; j++) {
        int jp;
        if (j < o->s_y)
            jp = j + ny;
        else if (j > o->e_y)
            jp = j - ny;
        else
            continue;
        for (int i = 0; i < o->mx; i++) {
            size_t a = IDX(o, i, j), b = IDX(o, i, jp);
            unsigned char e = o->rcF[a] || o->rcF[b];
            o->rcF[a] = e;
            o->rcF[b] = e;
            e = o->rc[a] || o->rc[b];
            o->rc[a] = e;
            o->rc[b] = e;
            int ls = imin(o->scheme[a], o->scheme[b]);
            o->scheme[a] = ls;
            o->scheme[b] = ls;
        }
    }
}

/* src/boundaryConditions.f90:118-187, for an array of nv variables per cell */
static void bc_fill(const struct orc *o, double *A, int nv) {
    int ngc = o->ngc;
    if (o->c.bc == ORC_BC_OUTFLOW) {
        for (int j = 0; j < o->my; j++) {
            for (int i = 0; i < ngc; i++)
                memcpy(&A[IDX(o, i, j) * nv], &A[IDX(o, o->s_x, j) * nv], nv * sizeof(double));
            for (int i = o->e_x + 1; i < o->mx; i++)
                memcpy(&A[IDX(o, i, j) * nv], &A[IDX(o, o->e_x, j) * nv], nv * sizeof(double));
        }
        for (int j = 0; j < ngc; j++)
            memcpy(&A[IDX(o, 0, j) * nv], &A[IDX(o, 0, o->s_y) * nv], (size_t)o->mx * nv * sizeof(double));
        for (int j = o->e_y + 1; j < o->my; j++)
            memcpy(&A[IDX(o, 0, j) * nv], &A[IDX(o, 0, o->e_y) * nv], (size_t)o->mx * nv * sizeof(double));
    } else {
        for (int j = 0; j < o->my; j++) {
            for (int g = 1; g <= ngc; g++)
                memcpy(&A[IDX(o, o->s_x - g, j) * nv], &A[IDX(o, o->e_x - g + 1, j) * nv], nv * sizeof(double));
            for (int g = 1; g <= ngc; g++)
                memcpy(&A[IDX(o, o->e_x + g, j) * nv], &A[IDX(o, o->s_x + g - 1, j) * nv], nv * sizeof(double));
        }
        for (int g = 1; g <= ngc; g++)
            memcpy(&A[IDX(o, 0, o->s_y - g) * nv], &A[IDX(o, 0, o->e_y - g + 1) * nv],
                   (size_t)o->mx * nv * sizeof(double));
        for (int g = 1; g <= ngc; g++)
            memcpy(&A[IDX(o, 0, o->e_y + g) * nv], &A[IDX(o, 0, o->s_y + g - 1) * nv],
                   (size_t)o->mx * nv * sizeof(double));
    }
}

/* src/timeStep.f90:45-130 */
static void forward_euler(struct orc *o, double dt, const double *U, double *Unew, int stage) {
    size_t ncell = (size_t)o->mx * o->my;
    for (size_t p = 0; p < ncell; p++) {
        o->scheme[p] = o->c.R;
        o->rc[p] = 1;
        o->rcF[p] = 1;
    }
    const double dx = o->dx, dy = o->dy;
    int exit_loop = 0;
    o->passes[stage] = 0;
    o->inner[stage] = 0;
    while (!exit_loop) {
        o->passes[stage]++;
        reconstruct(o, U, &o->inner[stage]);
        get_fluxes(o);
        flicd(o);
#pragma omp parallel for collapse(2) schedule(static)
        for (int j = o->s_y; j <= o->e_y; j++)
            for (int i = o->s_x; i <= o->e_x; i++) {
                if (!o->rcF[IDX(o, i, j)]) continue;
                for (int v = 0; v < NC; v++)
                    UA(o, Unew, v, i, j) =
                        UA(o, U, v, i, j) -
                        (dt / dx) * (FL(o, o->flux, v, 0, i + 1, j) - FL(o, o->flux, v, 0, i, j)) -
                        (dt / dy) * (FL(o, o->flux, v, 1, i, j + 1) - FL(o, o->flux, v, 1, i, j));
            }
        exit_loop = detect_all(o, U, Unew);
        if (o->c.bc == ORC_BC_PERIODIC) periodic_mood(o);
    }
    memcpy(o->stage_scheme + (size_t)stage * ncell, o->scheme, ncell * sizeof(int));
}

/* src/timeStep.f90:132-171 and :22-43 */
void orc_advance(orc_t *o, double dt) {
    size_t n = (size_t)o->mx * o->my * NC;
    double *U = o->U, *U1 = o->Us[0], *U2 = o->Us[1], *U3 = o->Us[2], *Un = o->Us[3];
    memset(U1, 0, n * sizeof(double));
    memset(U2, 0, n * sizeof(double));
    memset(U3, 0, n * sizeof(double));
    forward_euler(o, dt, U, U1, 0);
    bc_fill(o, U1, NC);
    forward_euler(o, dt, U1, U2, 1);
    bc_fill(o, U2, NC);
    for (size_t p = 0; p < n; p++) U2[p] = (3.0 * U[p] + U2[p]) / 4.0;
    bc_fill(o, U2, NC);
    forward_euler(o, dt, U2, U3, 2);
    bc_fill(o, U3, NC);
    for (size_t p = 0; p < n; p++) Un[p] = (U[p] + 2.0 * U3[p]) / 3.0;
    memcpy(U, Un, n * sizeof(double));
    for (int j = o->s_y; j <= o->e_y; j++)
        for (int i = o->s_x; i <= o->e_x; i++)
            cons2prim(&o->c, &UA(o, U, 0, i, j), &o->V[IDX(o, i, j) * NP]);
    bc_fill(o, U, NC);
    bc_fill(o, o->V, NP);
}

/* src/cfl.f90:14-52 */
double orc_cfl_dt(const orc_t *o) {
    const orc_cfg *c = &o->c;
    double xmax = 0.0, ymax = 0.0;
    for (int j = o->s_y; j <= o->e_y; j++)
        for (int i = o->s_x; i <= o->e_x; i++) {
            const double *V = &o->V[IDX(o, i, j) * NP];
            double cs;
            if (c->force_hydro) {
                cs = sqrt(V[GAMM] * V[PRES] / V[DENS]);
            } else {
                double BB = V[MAGX] * V[MAGX] + V[MAGY] * V[MAGY] + V[MAGZ] * V[MAGZ];
                double ab[3] = {fabs(V[MAGX]), fabs(V[MAGY]), fabs(V[MAGZ])};
                double bn = minval(ab, 3);
                cs = eigen_cfs(V[GAMM], V[DENS], V[PRES], BB, bn, +1.0);
            }
            double xl = fabs(V[VELX]) + cs, yl = fabs(V[VELY]) + cs;
            xmax = fmaxF(xmax, xl);
            ymax = fmaxF(ymax, yl);
        }
    return c->cfl * fminF(o->dx / xmax, o->dy / ymax);
}

/* ---------------------------------------------------------------- */
/* public API                                                         */

orc_t *orc_create(const orc_cfg *cfg) {
    if (cfg->R < 1 || cfg->R > 3 || cfg->nx < 1 || cfg->ny < 1) return NULL;
    struct orc *o = calloc(1, sizeof *o);
    if (!o) return NULL;
    o->c = *cfg;
    o->ngc = cfg->R + 2;
    o->mx = cfg->nx + 2 * o->ngc;
    o->my = cfg->ny + 2 * o->ngc;
    o->s_x = o->ngc;
    o->e_x = o->ngc + cfg->nx - 1;
    o->s_y = o->ngc;
    o->e_y = o->ngc + cfg->ny - 1;
    o->dx = (cfg->xend - cfg->xbeg) / cfg->nx;
    o->dy = (cfg->yend - cfg->ybeg) / cfg->ny;
    /* src/grid.f90:82-83: MINVAL / MAXVAL of (dx, dy) */
    o->mindxdy = (o->dy < o->dx) ? o->dy : o->dx;
    o->maxdxdy = (o->dy > o->dx) ? o->dy : o->dx;
    if (gp_init(o) != 0) {
        free(o);
        return NULL;
    }
    size_t ncell = (size_t)o->mx * o->my;
    o->U = calloc(ncell * NC, sizeof(double));
    o->V = calloc(ncell * NP, sizeof(double));
    o->faceVals = calloc(ncell * 4 * NQMAX * NC, sizeof(double));
    o->flux = calloc(ncell * 2 * NC, sizeof(double));
    o->fluxRaw = calloc(ncell * 2 * NC, sizeof(double));
    o->scheme = calloc(ncell, sizeof(int));
    o->rc = calloc(ncell, 1);
    o->rcF = calloc(ncell, 1);
    o->cellAdm = calloc(ncell, 1);
    o->recCell = calloc(ncell, 1);
    o->recNext = calloc(ncell, 1);
    o->checkCell = calloc(ncell, 1);
    o->lowerCell = calloc(ncell, 1);
    for (int k = 0; k < 4; k++) o->Us[k] = calloc(ncell * NC, sizeof(double));
    o->stage_scheme = calloc(3 * ncell, sizeof(int));
    return o;
}

void orc_destroy(orc_t *o) {
    if (!o) return;
    free(o->U); free(o->V); free(o->faceVals); free(o->flux); free(o->fluxRaw);
    free(o->scheme); free(o->rc); free(o->rcF); free(o->cellAdm); free(o->recCell);
    free(o->recNext); free(o->checkCell); free(o->lowerCell);
    for (int k = 0; k < 4; k++) free(o->Us[k]);
    free(o->stage_scheme);
    free(o);
}

void orc_dims(const orc_t *o, int *ngc, int *mx, int *my) {
    *ngc = o->ngc;
    *mx = o->mx;
    *my = o->my;
}

void orc_gp_tables(const orc_t *o, double *pv, double *qw, int *st) {
    /* pv: (nsmax, nqmax, 4, R) Fortran order with nsmax/nqmax for R;
       qw: (nqmax, R); st: (2, nsmax, R) */
    int R = o->c.R, nsm = o->ns[R], nqm = o->nq[R];
    for (int rr = 1; rr <= R; rr++)
        for (int f = 0; f < 4; f++)
            for (int q = 0; q < nqm; q++)
                for (int k = 0; k < nsm; k++)
                    pv[(((size_t)(rr - 1) * 4 + f) * nqm + q) * nsm + k] = o->pv[rr][f][q][k];
    for (int rr = 1; rr <= R; rr++)
        for (int q = 0; q < nqm; q++) qw[(rr - 1) * nqm + q] = o->qw[rr][q];
    for (int rr = 1; rr <= R; rr++)
        for (int k = 0; k < nsm; k++) {
            st[((rr - 1) * nsm + k) * 2 + 0] = o->st[rr][k][0];
            st[((rr - 1) * nsm + k) * 2 + 1] = o->st[rr][k][1];
        }
}

void orc_set_state(orc_t *o, const double *U, const double *V) {
    size_t ncell = (size_t)o->mx * o->my;
    memcpy(o->U, U, ncell * NC * sizeof(double));
    if (V) {
        memcpy(o->V, V, ncell * NP * sizeof(double));
    } else {
        for (int j = o->s_y; j <= o->e_y; j++)
            for (int i = o->s_x; i <= o->e_x; i++)
                cons2prim(&o->c, &UA(o, o->U, 0, i, j), &o->V[IDX(o, i, j) * NP]);
        bc_fill(o, o->V, NP);
    }
}

void orc_get_state(const orc_t *o, double *U, double *V) {
    size_t ncell = (size_t)o->mx * o->my;
    if (U) memcpy(U, o->U, ncell * NC * sizeof(double));
    if (V) memcpy(V, o->V, ncell * NP * sizeof(double));
}

void orc_get_stage_schemes(const orc_t *o, int *out) {
    memcpy(out, o->stage_scheme, 3 * (size_t)o->mx * o->my * sizeof(int));
}

void orc_get_stats(const orc_t *o, int *passes, int *inner) {
    for (int s = 0; s < 3; s++) {
        passes[s] = o->passes[s];
        inner[s] = o->inner[s];
    }
}

/* One forward-Euler stage from U (padded, guards filled), exposing the
   stage's MOOD state; Unew receives interior values only. */
void orc_forward_euler(orc_t *o, double dt, const double *U, double *Unew, int *scheme_out) {
    size_t n = (size_t)o->mx * o->my;
    forward_euler(o, dt, U, Unew, 0);
    if (scheme_out) memcpy(scheme_out, o->scheme, n * sizeof(int));
}

/* src/reconstruct.f90 driven as in test/test_reconstruct_faceValsWithGP.f90:
   scheme set to `radius` everywhere, every cell flagged, faceVals returned
   as (8, NQMAX, 4, i, j). */
void orc_reconstruct_all(orc_t *o, const double *U, int radius, double *faceVals_out) {
    size_t ncell = (size_t)o->mx * o->my;
    for (size_t p = 0; p < ncell; p++) {
        o->scheme[p] = radius;
        o->rc[p] = 1;
        o->rcF[p] = 1;
    }
    memset(o->faceVals, 0, ncell * 4 * NQMAX * NC * sizeof(double));
    reconstruct(o, U, NULL);
    memcpy(faceVals_out, o->faceVals, ncell * 4 * NQMAX * NC * sizeof(double));
}

void orc_bc_fill(const orc_t *o, double *A, int nv) { bc_fill(o, A, nv); }

/* building blocks, exposed for unit tests */
void orc_hllc(const orc_t *o, const double *uL, const double *uR, int dir, double *flux) {
    hllc(&o->c, uL, uR, dir, flux);
}
void orc_cons2prim(const orc_t *o, const double *U, double *V) { cons2prim(&o->c, U, V); }
void orc_prim2cons(const orc_t *o, const double *V, double *U) { prim2cons(&o->c, V, U); }
void orc_cons2flux(const orc_t *o, const double *U, int dir, double *F) { cons2flux(&o->c, U, dir, F); }
double orc_cons2pres(const orc_t *o, const double *U) { return cons2pres(&o->c, U); }
double orc_eigen_cfs(double gam, double rho, double p, double BB, double BN, double pm) {
    return eigen_cfs(gam, rho, p, BB, BN, pm);
}
