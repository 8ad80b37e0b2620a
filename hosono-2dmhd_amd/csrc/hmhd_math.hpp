// hmhd_math.hpp -- per-cell and per-face FP64 arithmetic of the step, on the
// device.  Every expression keeps the reference's operand order so that the
// library, built without multiply-add contraction, reproduces the reference
// bit for bit:
//   pressure floor          src/eos.f90:9-26
//   cons -> prim            src/convert.f90:198-234
//   unfloored pressure      src/convert.f90:236-254
//   physical flux           src/convert.f90:28-92
//   fast magnetosonic speed src/eigen.f90:257-321 (eigenvalues 1 and 7 only:
//                           HLLC reads no other)
//   HLLC-MHD (Li 2005)      src/riemannSolver.f90:68-221
//   CAD / PAD               src/detect.f90:238-265
// MAX(a,b) is a>b?a:b and MIN(a,b) is a<b?a:b, the rule the reference's
// compiler applies.
// The file also compiles on the host with HMHD_MATH_HOST defined (the
// instrumented operation count, flop_count.cpp, runs these same formulas).
#pragma once

#ifndef HMHD_MATH_HOST
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include "hmhd_internal.hpp"

namespace hmhd {

enum : int { DENS = 0, MOMX = 1, MOMY = 2, MOMZ = 3, ENER = 4, MAGX = 5, MAGY = 6, MAGZ = 7 };


struct Phys {
    double gamma, small_p, small_rho;
    int hydro;
};

__device__ __forceinline__ double fmaxF(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double fminF(double a, double b) { return a < b ? a : b; }

// primitive state without the two constants eint/gamma the step never reads
// from a Riemann state except through gamma itself
struct Prim {
    double d, vx, vy, vz, p, bx, by, bz;
};

__device__ __forceinline__ double eos_pres(const Phys& ph, double dens, double eint) {
    double pres = (ph.gamma - 1.0) * dens * eint;
    if (pres < ph.small_p) pres = ph.small_p;
    return pres;
}

__device__ __forceinline__ double mag_pressure(const Phys& ph, const double* U) {
    return ph.hydro ? 0.0 : 0.5 * (U[MAGX] * U[MAGX] + U[MAGY] * U[MAGY] + U[MAGZ] * U[MAGZ]);
}

__device__ __forceinline__ Prim cons2prim(const Phys& ph, const double* U) {
    double Bp = mag_pressure(ph, U);
    Prim V;
    V.d = U[DENS];
    V.vx = U[MOMX] / V.d;
    V.vy = U[MOMY] / V.d;
    V.vz = U[MOMZ] / V.d;
    double eint = (U[ENER] - Bp) / V.d - 0.5 * (V.vx * V.vx + V.vy * V.vy + V.vz * V.vz);
    V.p = eos_pres(ph, V.d, eint);
    if (ph.hydro) {
        V.bx = V.by = V.bz = 0.0;
    } else {
        V.bx = U[MAGX];
        V.by = U[MAGY];
        V.bz = U[MAGZ];
    }
    return V;
}

__device__ __forceinline__ double cons2pres(const Phys& ph, const double* U) {
    double Bp = mag_pressure(ph, U);
    return (ph.gamma - 1.0) *
           (U[ENER] - Bp - 0.5 * (U[MOMX] * U[MOMX] + U[MOMY] * U[MOMY] + U[MOMZ] * U[MOMZ]) / U[DENS]);
}

__device__ __forceinline__ bool admissible(double dens, double pres) {
    // CAD: no NaN/Inf; PAD: dens > 0 and pres > 0
    bool cad = !(__builtin_isnan(dens) || __builtin_isnan(pres) || !__builtin_isfinite(dens) ||
                 !__builtin_isfinite(pres));
    return cad && (dens > 0.0) && (pres > 0.0);
}

__device__ __forceinline__ double eigen_cfs(double gam, double rho, double p, double BB, double BN) {
    double cfs = (gam * p + BB) / rho;
    cfs = 0.5 * (cfs + 1.0 * __builtin_sqrt(cfs * cfs - 4.0 * gam * p * (BN * BN) / (rho * rho)));
    return __builtin_sqrt(fmaxF(1.0e-14, cfs));
}

// The two speeds cfl_computedt maximises for one cell (src/cfl.f90:28-48):
// |vx| + c and |vy| + c, c the sound speed (hydro) or the fast speed with
// B_N = MIN(|Bx|, |By|, |Bz|), NaN-skipping as MINVAL; returned as the bit
// patterns of the two non-negative values, whose unsigned order is their
// numeric order (NaN above +inf).
__device__ __forceinline__ void cfl_speed_bits(double gamma, int hydro, double d, double vx, double vy, double p,
                                               double b1, double b2, double b3, unsigned long long& bx,
                                               unsigned long long& by) {
    double cs;
    if (hydro) {
        cs = __builtin_sqrt(gamma * p / d);
    } else {
        const double BB = b1 * b1 + b2 * b2 + b3 * b3;
        const double ab[3] = {__builtin_fabs(b1), __builtin_fabs(b2), __builtin_fabs(b3)};
        double bn = ab[0];
        for (int k = 1; k < 3; k++)
            if (__builtin_isnan(bn) || ab[k] < bn) bn = ab[k];
        cs = eigen_cfs(gamma, d, p, BB, bn);
    }
    const double xl = __builtin_fabs(vx) + cs, yl = __builtin_fabs(vy) + cs;
    bx = (unsigned long long)__double_as_longlong(xl);
    by = (unsigned long long)__double_as_longlong(yl);
}

// fast speed of the split system in direction DIR (0 = x, 1 = y)
template <int DIR>
__device__ __forceinline__ double fast_speed(const Phys& ph, const Prim& V) {
    double rho = fmaxF(V.d, ph.small_rho);
    double BN = DIR == 0 ? V.bx : V.by;
    double p = fmaxF(V.p, ph.small_p);
    double BB = V.bx * V.bx + V.by * V.by + V.bz * V.bz;
    return eigen_cfs(ph.gamma, rho, p, BB, BN);
}

// indices of normal / transverse components for direction DIR
template <int DIR>
struct Dir {
    static constexpr int mN = DIR == 0 ? MOMX : MOMY;
    static constexpr int mT1 = DIR == 0 ? MOMY : MOMX;
    static constexpr int mT2 = MOMZ;
    static constexpr int bN = DIR == 0 ? MAGX : MAGY;
    static constexpr int bT1 = DIR == 0 ? MAGY : MAGX;
    static constexpr int bT2 = MAGZ;
};

__device__ __forceinline__ double pcomp(const Prim& V, int k) {
    // velocity / field component by conservative index
    switch (k) {
        case MOMX: return V.vx;
        case MOMY: return V.vy;
        case MOMZ: return V.vz;
        case MAGX: return V.bx;
        case MAGY: return V.by;
        default: return V.bz;
    }
}

template <int DIR>
__device__ __forceinline__ void cons2flux(const Phys& ph, const double* U, const Prim& V, double* F) {
    using D = Dir<DIR>;
    double Bp = mag_pressure(ph, U);
    double E = U[ENER];
    double Ptot = V.p + Bp;
    const double vN = pcomp(V, D::mN), vT1 = pcomp(V, D::mT1), vT2 = pcomp(V, D::mT2);
    const double bN = pcomp(V, D::bN), bT1 = pcomp(V, D::bT1), bT2 = pcomp(V, D::bT2);
#pragma unroll
    for (int v = 0; v < NCONS; v++) F[v] = 0.0;
    F[DENS] = U[D::mN];
    F[D::mN] = U[D::mN] * U[D::mN] / U[DENS] + Ptot;
    F[D::mT1] = U[D::mN] * vT1;
    F[D::mT2] = U[D::mN] * vT2;
    F[ENER] = (E + Ptot) * vN;
    if (!ph.hydro) {
        F[D::mN] = F[D::mN] - bN * bN;
        F[D::mT1] = F[D::mT1] - bT1 * bN;
        F[D::mT2] = F[D::mT2] - bT2 * bN;
        F[D::bT1] = vN * bT1 - vT1 * bN;
        F[D::bT2] = vN * bT2 - vT2 * bN;
        F[ENER] = F[ENER] - bN * (V.vx * V.bx + V.vy * V.by + V.vz * V.bz);
    }
}

// What HLLC computes from one side's state alone: its primitives, fast
// speed and physical flux, B.v and the total pressure.
struct HSide {
    Prim V;
    double c, mdv, pTot;
    double f[NCONS];
};
template <int DIR>
__device__ __forceinline__ HSide hllc_side(const Phys& ph, const double* u) {
    HSide h;
    h.V = cons2prim(ph, u);
    h.c = fast_speed<DIR>(ph, h.V);
    cons2flux<DIR>(ph, u, h.V, h.f);
    h.mdv = h.V.bx * h.V.vx + h.V.by * h.V.vy + h.V.bz * h.V.vz;
    h.pTot = h.V.p + 0.5 * (h.V.bx * h.V.bx + h.V.by * h.V.by + h.V.bz * h.V.bz);
    return h;
}

// HLLC-MHD (src/riemannSolver.f90:68-221) from the two sides' states and
// their hllc_side values; hllc() below is both together.
template <int DIR>
__device__ __forceinline__ void hllc_from_sides(const Phys& ph, const double* uL, const double* uR, const HSide& hL,
                                                const HSide& hR, double* flux) {
    using D = Dir<DIR>;
    const Prim& vL = hL.V;
    const Prim& vR = hR.V;
    Prim vA;
    vA.d = (vL.d + vR.d) / 2.0;
    vA.vx = (vL.vx + vR.vx) / 2.0;
    vA.vy = (vL.vy + vR.vy) / 2.0;
    vA.vz = (vL.vz + vR.vz) / 2.0;
    vA.p = (vL.p + vR.p) / 2.0;
    vA.bx = (vL.bx + vR.bx) / 2.0;
    vA.by = (vL.by + vR.by) / 2.0;
    vA.bz = (vL.bz + vR.bz) / 2.0;
    const double cL = hL.c;
    const double cR = hR.c;
    const double cA = fast_speed<DIR>(ph, vA);
    const double vNL = pcomp(vL, D::mN), vNR = pcomp(vR, D::mN), vNA = pcomp(vA, D::mN);
    const double sL = fminF(vNL - cL, vNA - cA);
    const double sR = fmaxF(vNA + cA, vNR + cR);

    const double* fL = hL.f;
    const double* fR = hR.f;
    double uHLL[NCONS];
#pragma unroll
    for (int v = 0; v < NCONS; v++) uHLL[v] = (sR * uR[v] - sL * uL[v] + fL[v] - fR[v]) / (sR - sL);
    const double velx = uHLL[MOMX] / uHLL[DENS], vely = uHLL[MOMY] / uHLL[DENS], velz = uHLL[MOMZ] / uHLL[DENS];
    const double mdvHLL = uHLL[MAGX] * velx + uHLL[MAGY] * vely + uHLL[MAGZ] * velz;
    const double mdvL = hL.mdv;
    const double mdvR = hR.mdv;

    const double velT1L = pcomp(vL, D::mT1), velT1R = pcomp(vR, D::mT1);
    const double velT2L = pcomp(vL, D::mT2), velT2R = pcomp(vR, D::mT2);
    const double momNL = uL[D::mN], momNR = uR[D::mN];
    const double magNL = pcomp(vL, D::bN), magNR = pcomp(vR, D::bN), magNHLL = uHLL[D::bN];
    const double magT1L = pcomp(vL, D::bT1), magT1R = pcomp(vR, D::bT1), magT1HLL = uHLL[D::bT1];
    const double magT2L = pcomp(vL, D::bT2), magT2R = pcomp(vR, D::bT2), magT2HLL = uHLL[D::bT2];
    const double rhoL = vL.d, rhoR = vR.d;
    const double pTotL = hL.pTot;
    const double pTotR = hR.pTot;
    const double dL = sL - vNL, dR = sR - vNR;
    const double sStr =
        (momNR * dR - momNL * dL + pTotL - pTotR - magNL * magNL + magNR * magNR) / (rhoR * dR - rhoL * dL);
    const double pStr = rhoL * dL * (sStr - vNL) + pTotL - magNL * magNL + magNHLL * magNHLL;

    if (0.0 < sL) {
#pragma unroll
        for (int v = 0; v < NCONS; v++) flux[v] = fL[v];
    } else if ((sL <= 0.0) && (0.0 < sStr)) {
        double uS[NCONS];
        uS[DENS] = rhoL * dL / (sL - sStr);
        uS[D::mN] = uS[DENS] * sStr;
        uS[D::mT1] = uS[DENS] * velT1L - (magNHLL * magT1HLL - magNL * magT1L) / (sL - sStr);
        uS[D::mT2] = uS[DENS] * velT2L - (magNHLL * magT2HLL - magNL * magT2L) / (sL - sStr);
        uS[D::bN] = magNHLL;
        uS[D::bT1] = magT1HLL;
        uS[D::bT2] = magT2HLL;
        uS[ENER] = uL[ENER] * dL + pStr * sStr - pTotL * vNL - magNHLL * mdvHLL + magNL * mdvL;
        uS[ENER] = uS[ENER] / (sL - sStr);
#pragma unroll
        for (int v = 0; v < NCONS; v++) flux[v] = fL[v] + sL * (uS[v] - uL[v]);
    } else if ((sStr <= 0.0) && (0.0 <= sR)) {
        double uS[NCONS];
        uS[DENS] = rhoR * dR / (sR - sStr);
        uS[D::mN] = uS[DENS] * sStr;
        uS[D::mT1] = uS[DENS] * velT1R - (magNHLL * magT1HLL - magNR * magT1R) / (sR - sStr);
        uS[D::mT2] = uS[DENS] * velT2R - (magNHLL * magT2HLL - magNR * magT2R) / (sR - sStr);
        uS[D::bN] = magNHLL;
        uS[D::bT1] = magT1HLL;
        uS[D::bT2] = magT2HLL;
        uS[ENER] = uR[ENER] * dR + pStr * sStr - pTotR * vNR - magNHLL * mdvHLL + magNR * mdvR;
        uS[ENER] = uS[ENER] / (sR - sStr);
#pragma unroll
        for (int v = 0; v < NCONS; v++) flux[v] = fR[v] + sR * (uS[v] - uR[v]);
    } else {
#pragma unroll
        for (int v = 0; v < NCONS; v++) flux[v] = fR[v];
    }
}

template <int DIR>
__device__ __forceinline__ void hllc(const Phys& ph, const double* uL, const double* uR, double* flux) {
    const HSide hL = hllc_side<DIR>(ph, uL);
    const HSide hR = hllc_side<DIR>(ph, uR);
    hllc_from_sides<DIR>(ph, uL, uR, hL, hR, flux);
}


// ------------------------------------------------------------------ //
// Reconstruction, FV update and detection formulas of the per-cell kernels.

// stencil of radius RR in the reference's order: i outer, j inner,
// |i|+|j| <= RR (src/GP.f90:85-95)
template <int RR>
struct Sten {
    static constexpr int n = gp_ns(RR);
    int x[n], y[n];
    constexpr Sten() : x(), y() {
        int c = 0;
        for (int i = -RR; i <= RR; i++)
            for (int j = -RR; j <= RR; j++)
                if ((i < 0 ? -i : i) + (j < 0 ? -j : j) <= RR) {
                    x[c] = i;
                    y[c] = j;
                    c++;
                }
    }
};

// GP point value at quadrature point q of face f of the cell whose stencil
// the accessor A is centred on; radius RR (src/reconstruct.f90:109-124).
// The stencil loop is unrolled eight cells at a time: unrolled in full, the
// compiler hoists every stencil load of both face sides, and the first-pass
// face kernel then needs more than 256 VGPRs; by eight it needs at most 166
// and runs three waves per SIMD.
template <int RR, class A>
__device__ __forceinline__ void gp_point(const A& a, const double* __restrict__ w, double* out) {
    constexpr Sten<RR> st{};
#pragma unroll
    for (int v = 0; v < NCONS; v++) out[v] = 0.0;
#pragma unroll 8
    for (int k = 0; k < st.n; k++) {
        const double wk = w[k];
#pragma unroll
        for (int v = 0; v < NCONS; v++) out[v] = out[v] + wk * a.at(v, st.x[k], st.y[k]);
    }
}

// FLICD correction (src/FLICDdivB.f90:38-54): (2 a + b + c - d - e - f - g) / 8,
// summed left to right
__device__ __forceinline__ double flicd_corr(double a, double b, double c, double d, double e, double f, double g) {
    return (2.0 * a + b + c - d - e - f - g) / 8.0;
}

// the unsplit FV update of one variable (src/timeStep.f90:112-114)
__device__ __forceinline__ double fv_update(double uc, double dtdx, double fxp, double fxm, double dtdy, double fyp,
                                            double fym) {
    return uc - dtdx * (fxp - fxm) - dtdy * (fyp - fym);
}

// SSP-RK3 combination of stage `stage` (src/timeStep.f90:165,169)
__device__ __forceinline__ double rk_combine(int stage, double ub, double un) {
    return stage == 0 ? un : (stage == 1 ? (3.0 * ub + un) / 4.0 : (ub + 2.0 * un) / 3.0);
}

// NaN-skipping MINVAL / MAXVAL of five values
__device__ __forceinline__ double minval5(const double* x) {
    double m = x[0];
#pragma unroll
    for (int k = 1; k < 5; k++)
        if (__builtin_isnan(m) || x[k] < m) m = x[k];
    return m;
}
__device__ __forceinline__ double maxval5(const double* x) {
    double m = x[0];
#pragma unroll
    for (int k = 1; k < 5; k++)
        if (__builtin_isnan(m) || x[k] > m) m = x[k];
    return m;
}

// The values around one cell that detection reads, loaded together
// before any is used: U of the cross neighbours and the density of the
// cells at taxicab distance 2 (guard cells through the index map).
struct Around {
    double e[NCONS], w[NCONS], n[NCONS], s[NCONS];  // U at (i+1,j), (i-1,j), (i,j+1), (i,j-1)
    double d2[8];  // rho at (i+2,j), (i-2,j), (i,j+2), (i,j-2), (i+1,j+1), (i+1,j-1), (i-1,j+1), (i-1,j-1)
};

// density at offset (di, dj), |di| + |dj| <= 2, from the centre cell and `a`
__device__ __forceinline__ double dens_at(const double* uc, const Around& a, int di, int dj) {
    if (di == 0 && dj == 0) return uc[DENS];
    if (di == 1 && dj == 0) return a.e[DENS];
    if (di == -1 && dj == 0) return a.w[DENS];
    if (di == 0 && dj == 1) return a.n[DENS];
    if (di == 0 && dj == -1) return a.s[DENS];
    if (di == 2) return a.d2[0];
    if (di == -2) return a.d2[1];
    if (dj == 2) return a.d2[2];
    if (dj == -2) return a.d2[3];
    if (di == 1) return dj == 1 ? a.d2[4] : a.d2[5];
    return dj == 1 ? a.d2[6] : a.d2[7];
}

// src/detect.f90:83-213 for one checked cell; true = lower the order
__device__ __forceinline__ bool detect_cell(const Inst& I, const Phys& ph, const Around& a, const double* un,
                                            const double* uc) {
    const double presNew = cons2pres(ph, un);
    if (!admissible(un[DENS], presNew)) return true;
    const Prim V = cons2prim(ph, uc);
    const Prim Vip = cons2prim(ph, a.e);
    const Prim Vim = cons2prim(ph, a.w);
    const Prim Vjp = cons2prim(ph, a.n);
    const Prim Vjm = cons2prim(ph, a.s);
    double div_vel = (Vip.vx - Vim.vx) / (2.0 * I.dx);
    div_vel = div_vel + (Vjp.vy - Vjm.vy) / (2.0 * I.dy);
    const bool shock1 = div_vel < -(1.0 * (I.maxdxdy * I.maxdxdy));
    double mach = V.vx * V.vx + V.vy * V.vy;
    mach = V.d * mach / (ph.gamma * V.p);
    mach = __builtin_sqrt(mach);
    const bool shock2 = mach > 0.2;
    double pg = __builtin_fabs(Vip.p - Vim.p) / fminF(Vip.p, Vim.p);
    pg = pg + __builtin_fabs(Vjp.p - Vjm.p) / fminF(Vjp.p, Vjm.p);
    const bool shock3 = pg > 10.0 * I.maxdxdy;
    if (!(shock1 && shock2 && shock3)) return false;

    const double densNew = un[DENS];
    const double d1 = uc[DENS], d2 = a.e[DENS], d3 = a.w[DENS], d4 = a.n[DENS], d5 = a.s[DENS];
    const double mn = fminF(fminF(fminF(fminF(d1, d2), d3), d4), d5);
    const double mxd = fmaxF(fmaxF(fmaxF(fmaxF(d1, d2), d3), d4), d5);
    if (mxd - mn < I.mindxdy * I.mindxdy * I.mindxdy) return false;
    if ((mn <= densNew) && (densNew <= mxd)) return false;

    // u2 smoothness on the five cells of the cross (src/detect.f90:177-213)
    double ddx[5], ddy[5];
    const double dx2 = I.dx * I.dx, dy2 = I.dy * I.dy;
    constexpr int ci[5] = {-1, 0, 0, 0, 1}, cj[5] = {0, -1, 0, 1, 0};
#pragma unroll
    for (int c = 0; c < 5; c++) {
        const double dcc = dens_at(uc, a, ci[c], cj[c]);
        ddx[c] = (dens_at(uc, a, ci[c] + 1, cj[c]) - 2.0 * dcc + dens_at(uc, a, ci[c] - 1, cj[c])) / dx2;
        ddy[c] = (dens_at(uc, a, ci[c], cj[c] + 1) - 2.0 * dcc + dens_at(uc, a, ci[c], cj[c] - 1)) / dy2;
    }
    const double mnx = minval5(ddx), mny = minval5(ddy), mxx = maxval5(ddx), mxy = maxval5(ddy);
    const double m = I.mindxdy;
    const bool sx1 = mnx * mxx > -m, sy1 = mny * mxy > -m;
    const double ax = fmaxF(__builtin_fabs(mnx), __builtin_fabs(mxx)), ay = fmaxF(__builtin_fabs(mny), __builtin_fabs(mxy));
    const bool sx2 = ax < m, sy2 = ay < m;
    const bool sx3 = fminF(__builtin_fabs(mnx), __builtin_fabs(mxx)) >= 0.5 * ax;
    const bool sy3 = fminF(__builtin_fabs(mny), __builtin_fabs(mxy)) >= 0.5 * ay;
    const bool smooth1 = sx1 && sx2 && sy1 && sy2;
    const bool smooth2 = sx1 && sx3 && sy1 && sy3;
    return !(smooth1 || smooth2);
}

}  // namespace hmhd
