// gp_tables.cpp -- GP prediction vectors in binary128, rounded to FP64.
//
// Gaussian-process face predictors of the GP-MOOD reconstruction
// (src/GP.f90:37-200): for every radius rr <= R, the diamond stencil
// |i|+|j| <= rr, Gauss-Legendre points on the four faces, the integrated
// squared-exponential covariances (Reyes et al. 2018 eqs. 22-23; Bourgeois &
// Lee 2022 eqs. 9-12; src/GP.f90:216-335), the SPD solve Kcov P = Kstar by
// Cholesky (src/linAlgQuadPrecision.f90:107-365), each column normalised to
// sum 1, then rounded to FP64.  binary128 is needed: Kcov has condition
// number ~7e10 at R = 2, l = 12 dx.  Host code, built by g++ with
// libquadmath; the arithmetic order follows the reference so the FP64
// vectors come out identical.
#include <quadmath.h>

#include <array>
#include <cmath>
#include <vector>

#include "gp_tables.hpp"

namespace hmhd {
namespace {

using quad = __float128;

struct Pt {
    quad x, y;
};

// covariance of two cell averages (src/GP.f90:216-280)
quad cov_avg_avg(const Pt& a, const Pt& b, const quad dl[2], quad ell) {
    const quad root2 = sqrtq(2.0Q), rpi = sqrtq(M_PIq);
    const quad s2l = root2 * ell, tl2 = 2.0Q * (ell * ell), pil2 = rpi * (ell * ell);
    const quad sep[2] = {a.x - b.x, a.y - b.y};
    quad k = 1.0Q;
    for (int d = 0; d < 2; d++) {
        const quad h = dl[d], r = sep[d];
        const quad r1 = (r + h) / s2l, r2 = (r - h) / s2l, r5 = r / s2l;
        const quad r3 = -((r + h) * (r + h) / tl2), r4 = -((r - h) * (r - h) / tl2), r6 = -(r * r / tl2);
        const quad bracket = r1 * erfq(r1) + r2 * erfq(r2) + (expq(r3) + expq(r4)) / rpi -
                             2.0Q * (r5 * erfq(r5) + expq(r6) / rpi);
        k = k * pil2 / (h * h) * bracket;
    }
    return k;
}

// covariance of a cell average and a point value (src/GP.f90:282-335)
quad cov_avg_point(const Pt& a, const Pt& p, const quad dl[2], quad ell) {
    const quad s2l = sqrtq(2.0Q) * ell, rhp = sqrtq(M_PIq / 2.0Q);
    const quad sep[2] = {a.x - p.x, a.y - p.y};
    quad k = 1.0Q;
    for (int d = 0; d < 2; d++) {
        const quad half = dl[d] / 2.0Q;
        const quad r1 = (sep[d] + half) / s2l, r2 = (sep[d] - half) / s2l;
        k = k * rhp * (ell / dl[d]) * (erfq(r1) - erfq(r2));
    }
    return k;
}

// Solve A X = B (A n x n SPD, B n x m, both row-major) via A = R^T R with
// R upper triangular, forward then back substitution.
bool spd_solve(int n, int m, const std::vector<quad>& A, std::vector<quad>& B) {
    std::vector<quad> Rm(A);
    auto R = [&](int r, int c) -> quad& { return Rm[(size_t)r * n + c]; };
    for (int r = 0; r < n; r++)
        for (int c = 0; c < r; c++) R(r, c) = 0.0Q;
    for (int k = 0; k < n; k++) {
        if (!(R(k, k) > 0.0Q)) return false;
        for (int r = k + 1; r < n; r++)
            for (int c = r; c < n; c++) R(r, c) = R(r, c) - R(k, c) * R(k, r) / R(k, k);
        const quad piv = sqrtq(R(k, k));
        for (int c = k; c < n; c++) R(k, c) = R(k, c) / piv;
    }
    std::vector<quad> Y((size_t)n * m);
    for (int c = 0; c < m; c++) {
        Y[c] = B[c] / R(0, 0);
        for (int r = 1; r < n; r++) {
            quad acc = 0.0Q;
            for (int k = 0; k < r; k++) acc = acc + R(k, r) * Y[(size_t)k * m + c];
            Y[(size_t)r * m + c] = (B[(size_t)r * m + c] - acc) / R(r, r);
        }
        B[(size_t)(n - 1) * m + c] = Y[(size_t)(n - 1) * m + c] / R(n - 1, n - 1);
        for (int r = n - 2; r >= 0; r--) {
            quad acc = 0.0Q;
            for (int k = r + 1; k < n; k++) acc = acc + R(r, k) * B[(size_t)k * m + c];
            B[(size_t)r * m + c] = (Y[(size_t)r * m + c] - acc) / R(r, r);
        }
    }
    return true;
}

}  // namespace

int gp_compute(int R, double ell_over_delta, double xbeg, double xend, double ybeg, double yend, int nx, int ny,
               double* pv, double* qw) {
    if (R < 1 || R > kMaxR || nx < 1 || ny < 1) return -1;
    const quad dl[2] = {((quad)xend - (quad)xbeg) / (quad)nx, ((quad)yend - (quad)ybeg) / (quad)ny};
    const quad ell = (quad)ell_over_delta * (dl[1] < dl[0] ? dl[1] : dl[0]);
    for (int k = 0; k < kPvSet; k++) pv[k] = 0.0;
    for (int k = 0; k < (kMaxR + 1) * 4; k++) qw[k] = 0.0;

    for (int rr = 1; rr <= R; rr++) {
        const int ns = gp_ns(rr), nq = gp_nq(rr), np = 4 * nq;
        std::vector<Pt> cells;
        for (int i = -rr; i <= rr; i++)
            for (int j = -rr; j <= rr; j++)
                if (std::abs(i) + std::abs(j) <= rr) cells.push_back({(quad)i * dl[0], (quad)j * dl[1]});

        // Gauss-Legendre abscissae on [-1/2, 1/2] (most positive first) and
        // FP64 weights (src/GP.f90:119-151)
        std::array<quad, 4> g{};
        double* w = qw + rr * 4;
        if (rr == 1) {
            g[0] = 1.0Q / 2.0Q / sqrtq(3.0Q);
            g[1] = -g[0];
            w[0] = w[1] = 1.0 / 2.0;
        } else if (rr == 2) {
            g[0] = 1.0Q / 2.0Q * sqrtq(3.0Q / 5.0Q);
            g[1] = 0.0Q;
            g[2] = -g[0];
            w[0] = w[2] = 5.0 / 18.0;
            w[1] = 8.0 / 18.0;
        } else {
            const quad t = 2.0Q / 7.0Q * sqrtq(6.0Q / 5.0Q);
            g[0] = 1.0Q / 2.0Q * sqrtq(3.0Q / 7.0Q + t);
            g[1] = 1.0Q / 2.0Q * sqrtq(3.0Q / 7.0Q - t);
            g[2] = -g[1];
            g[3] = -g[0];
            w[0] = w[3] = (18.0 - std::sqrt(30.0)) / 72.0;
            w[1] = w[2] = (18.0 + std::sqrt(30.0)) / 72.0;
        }
        // prediction points: face f, point q at column f*nq + q (N, S, E, W)
        std::vector<Pt> pts(np);
        for (int q = 0; q < nq; q++) {
            pts[0 * nq + q] = {g[q] * dl[0], dl[1] / 2.0Q};
            pts[1 * nq + q] = {g[q] * dl[0], -dl[1] / 2.0Q};
            pts[2 * nq + q] = {dl[0] / 2.0Q, g[q] * dl[1]};
            pts[3 * nq + q] = {-dl[0] / 2.0Q, g[q] * dl[1]};
        }
        std::vector<quad> K((size_t)ns * ns), P((size_t)ns * np);
        for (int a = 0; a < ns; a++)
            for (int b = 0; b < ns; b++) K[(size_t)a * ns + b] = cov_avg_avg(cells[a], cells[b], dl, ell);
        for (int a = 0; a < ns; a++)
            for (int p = 0; p < np; p++) P[(size_t)a * np + p] = cov_avg_point(cells[a], pts[p], dl, ell);
        if (!spd_solve(ns, np, K, P)) return -2;
        double* blk = pv + gp_pv_off(rr);
        for (int p = 0; p < np; p++) {
            quad sum = 0.0Q;
            for (int a = 0; a < ns; a++) sum = sum + P[(size_t)a * np + p];
            // column p = f*nq + q maps to blk[(f*nq + q)*ns + k]
            for (int a = 0; a < ns; a++) blk[(size_t)p * ns + a] = (double)(P[(size_t)a * np + p] / sum);
        }
    }
    return 0;
}

}  // namespace hmhd
