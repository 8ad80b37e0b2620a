// flop_count.cpp -- the instrumented FP64 operation count of one zone-update.
//
// The device's own formulas (hmhd_math.hpp: GP point values, admissibility,
// HLLC-MHD, FLICD, the FV update, detection, the RK combination, the CFL
// speed) are compiled here on the host with a scalar type that counts each
// +, -, *, / and sqrt as one operation; a sign flip, min, max, abs and every
// comparison count zero.  The program runs the first MOOD pass of a stage
// over a real state, in the order the kernels run it:
//
//   face   k_face_full: per padded cell of [s-2, e+3]^2, for its x- and
//          y-face, both sides' GP point values at every Gauss point, their
//          admissibility, and, on the flux faces [s-1, e+2]^2, HLLC and the
//          weighted sum;
//   update k_update at the first pass (every cell flagged): the FLICD
//          corrections of the cell's faces, the FV update of 8 variables and
//          detection (with the branches the state takes);
//   rk     the RK combinations of stages 1 and 2;
//   cfl    cons2prim and the two speeds of k_cfl.
//
// flops_per_zone_update = 3 (face + update) + rk + cfl, per interior cell;
// face_flops_per_launch is one launch of the face kernel over the grid.
// The face and update terms are averaged over `windows` windows of 32 x 32
// interior cells (seeded positions); the cells of the face kernel's domain
// outside the interior are counted in full.
//
// Usage: hmhd_flop_count U.bin nx ny R ell_over_delta gamma small_p small_rho
//                        hydro bc xbeg xend ybeg yend cfl windows seed
// U.bin: the padded state U(8, mx, my) as raw FP64, (my, mx, 8) in C order;
// only its interior is read, guard cells through the boundary condition bc
// (0 periodic, 1 outflow) as the kernels read them.
// One JSON line on stdout.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

namespace fc {
unsigned long long n_ops = 0;
struct F {
    double v;
    F() : v(0.0) {}
    F(double x) : v(x) {}
};
inline F operator+(F a, F b) { ++n_ops; return F(a.v + b.v); }
inline F operator-(F a, F b) { ++n_ops; return F(a.v - b.v); }
inline F operator*(F a, F b) { ++n_ops; return F(a.v * b.v); }
inline F operator/(F a, F b) { ++n_ops; return F(a.v / b.v); }
inline F operator-(F a) { return F(-a.v); }
inline bool operator<(F a, F b) { return a.v < b.v; }
inline bool operator>(F a, F b) { return a.v > b.v; }
inline bool operator<=(F a, F b) { return a.v <= b.v; }
inline bool operator>=(F a, F b) { return a.v >= b.v; }
inline F cnt_sqrt(F x) { ++n_ops; return F(std::sqrt(x.v)); }
inline F cnt_fabs(F x) { return F(std::fabs(x.v)); }
inline bool cnt_isnan(F x) { return std::isnan(x.v); }
inline bool cnt_isfinite(F x) { return std::isfinite(x.v); }
inline long long cnt_bits(F x) {
    long long b;
    std::memcpy(&b, &x.v, 8);
    return b;
}
}  // namespace fc

using CountedDouble = fc::F;
using fc::F;
#define HMHD_MATH_HOST
#define __device__
#define __host__
#define __forceinline__ inline
#define __restrict__
#define __builtin_sqrt fc::cnt_sqrt
#define __builtin_fabs fc::cnt_fabs
#define __builtin_isnan fc::cnt_isnan
#define __builtin_isfinite fc::cnt_isfinite
#define __double_as_longlong fc::cnt_bits
#define double CountedDouble
#include "hmhd_math.hpp"
#undef double

namespace hmhd {
int gp_compute(int R, double ell_over_delta, double xbeg, double xend, double ybeg, double yend, int nx, int ny,
               double* pv, double* qw);
}

using namespace hmhd;

namespace {

struct State {
    int nx, ny, R, ngc, mx, my, bc, hydro;
    std::vector<double> U;  // (my, mx, 8)
    // the padded index p read through the boundary condition, as the
    // kernels' index map does (any p the face domain's stencils reach)
    int map(int p, int n) const {
        int k = p - ngc;
        if (bc == BC_PERIODIC) {
            k += k < 0 ? n : 0;
            k -= k >= n ? n : 0;
        } else {
            k = k < 0 ? 0 : (k > n - 1 ? n - 1 : k);
        }
        return k + ngc;
    }
    double u(int v, int i, int j) const { return U[((size_t)map(j, ny) * mx + map(i, nx)) * 8 + v]; }
};

// a stencil accessor on the padded state
struct PAcc {
    const State* S;
    int i, j;
    F at(int v, int dx, int dy) const { return F(S->u(v, i + dx, j + dy)); }
};

struct Tally {
    unsigned long long recon = 0, admiss = 0, hllc = 0, sums = 0, hllc_calls = 0;
    unsigned long long flicd = 0, fv = 0, detect = 0;
    unsigned long long face() const { return recon + admiss + hllc + sums; }
};

F pvF[kPvSet], qwF[(kMaxR + 1) * 4];
const F* pv_at(int rr, int f, int q) { return &pvF[gp_pv_off(rr) + (f * gp_nq(rr) + q) * gp_ns(rr)]; }

template <int R>
void gp(const PAcc& a, const F* w, F* out) {
    gp_point<R>(a, w, out);
}

void recon(const State& S, const PAcc& a, int f, int q, F* out) {
    if (S.R == 1) gp<1>(a, pv_at(1, f, q), out);
    else if (S.R == 2) gp<2>(a, pv_at(2, f, q), out);
    else gp<3>(a, pv_at(3, f, q), out);
}

enum { FN = 0, FS = 1, FE = 2, FW = 3 };

// The face kernel's work for padded cell (i, j), as face_tile_full does it;
// the raw fluxes of its x- and y-face go to fx / fy when it owns them.
void face_cell(const State& S, const Phys& ph, int i, int j, Tally& t, F* fx, F* fy, bool& hasx, bool& hasy) {
    const int s = S.ngc, ex = S.ngc + S.nx - 1, ey = S.ngc + S.ny - 1, NQ = S.R + 1;
    const bool in_dom = i >= s - 2 && i <= ex + 3 && j >= s - 2 && j <= ey + 3;
    const bool flux_cell = in_dom && i >= s - 1 && i <= ex + 2 && j >= s - 1 && j <= ey + 2;
    const bool cellC = in_dom && i <= ex + 2 && j <= ey + 2;
    const bool cellW = in_dom && i >= s - 1 && j <= ey + 2;
    const bool cellS = in_dom && j >= s - 1 && i <= ex + 2;
    const bool do_x = cellC || cellW || flux_cell, do_y = cellC || cellS || flux_cell;
    hasx = hasy = false;
    const PAcc aC{&S, i, j}, aW{&S, i - 1, j}, aS{&S, i, j - 1};
    for (int d = 0; d < 2; d++) {
        if (d == 0 ? !do_x : !do_y) continue;
        F acc[NCONS];
        for (int q = 0; q < NQ; q++) {
            F uL[NCONS], uR[NCONS];
            unsigned long long c0 = fc::n_ops;
            recon(S, d == 0 ? aW : aS, d == 0 ? FE : FN, q, uL);
            recon(S, aC, d == 0 ? FW : FS, q, uR);
            unsigned long long c1 = fc::n_ops;
            (void)admissible(uL[DENS], cons2pres(ph, uL));
            (void)admissible(uR[DENS], cons2pres(ph, uR));
            unsigned long long c2 = fc::n_ops;
            t.recon += c1 - c0;
            t.admiss += c2 - c1;
            if (flux_cell) {
                F Fl[NCONS];
                if (d == 0) hllc<0>(ph, uL, uR, Fl);
                else hllc<1>(ph, uL, uR, Fl);
                unsigned long long c3 = fc::n_ops;
                t.hllc += c3 - c2;
                t.hllc_calls += 1;
                const F wq = qwF[S.R * 4 + q];
                for (int v = 0; v < NCONS; v++) acc[v] = (q == 0 ? F(0.0) : acc[v]) + wq * Fl[v];
                t.sums += fc::n_ops - c3;
            }
        }
        if (flux_cell) {
            F* dst = d == 0 ? fx : fy;
            for (int v = 0; v < NCONS; v++) dst[v] = acc[v];
            (d == 0 ? hasx : hasy) = true;
        }
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 18) {
        std::fprintf(stderr,
                     "usage: %s U.bin nx ny R ell gamma small_p small_rho hydro bc xbeg xend ybeg yend cfl windows "
                     "seed\n",
                     argv[0]);
        return 2;
    }
    State S;
    S.nx = std::atoi(argv[2]);
    S.ny = std::atoi(argv[3]);
    S.R = std::atoi(argv[4]);
    const double ell = std::atof(argv[5]);
    Phys ph;
    ph.gamma = std::atof(argv[6]);
    ph.small_p = std::atof(argv[7]);
    ph.small_rho = std::atof(argv[8]);
    S.hydro = ph.hydro = std::atoi(argv[9]);
    S.bc = std::atoi(argv[10]);
    const double xbeg = std::atof(argv[11]), xend = std::atof(argv[12]), ybeg = std::atof(argv[13]),
                 yend = std::atof(argv[14]);
    const double cfl = std::atof(argv[15]);
    const int windows = std::atoi(argv[16]);
    const unsigned seed = (unsigned)std::atoi(argv[17]);
    if (S.R < 1 || S.R > kMaxR || S.nx < 32 || S.ny < 32 || windows < 1) {
        std::fprintf(stderr, "need 1 <= R <= %d, nx, ny >= 32 and windows >= 1\n", kMaxR);
        return 2;
    }
    S.ngc = S.R + 2;
    S.mx = S.nx + 2 * S.ngc;
    S.my = S.ny + 2 * S.ngc;
    S.U.resize((size_t)S.mx * S.my * 8);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(S.U.data(), sizeof(double), S.U.size(), f) != S.U.size()) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 1;
    }
    std::fclose(f);
    double pv[kPvSet] = {0}, qw[(kMaxR + 1) * 4] = {0};
    if (gp_compute(S.R, ell, xbeg, xend, ybeg, yend, S.nx, S.ny, pv, qw) != 0) return 1;
    for (int k = 0; k < kPvSet; k++) pvF[k] = F(pv[k]);
    for (int k = 0; k < (kMaxR + 1) * 4; k++) qwF[k] = F(qw[k]);
    Inst I{};
    I.dx = F((xend - xbeg) / S.nx);
    I.dy = F((yend - ybeg) / S.ny);
    I.mindxdy = I.dy.v < I.dx.v ? I.dy : I.dx;
    I.maxdxdy = I.dy.v > I.dx.v ? I.dy : I.dx;
    const int s = S.ngc, ex = S.ngc + S.nx - 1, ey = S.ngc + S.ny - 1;

    // CFL over the interior (also gives the dt the update uses)
    unsigned long long c0 = fc::n_ops;
    unsigned long long mxb = 0, myb = 0;
    for (int j = s; j <= ey; j++)
        for (int i = s; i <= ex; i++) {
            F u[NCONS];
            for (int v = 0; v < NCONS; v++) u[v] = F(S.u(v, i, j));
            const Prim P = cons2prim(ph, u);
            unsigned long long bx, by;
            cfl_speed_bits(ph.gamma, S.hydro, P.d, P.vx, P.vy, P.p, P.bx, P.by, P.bz, bx, by);
            mxb = bx > mxb ? bx : mxb;
            myb = by > myb ? by : myb;
        }
    const double cfl_per_zone = double(fc::n_ops - c0) / ((double)S.nx * S.ny);
    double xm, ym;
    std::memcpy(&xm, &mxb, 8);
    std::memcpy(&ym, &myb, 8);
    const double dt = cfl * std::min(I.dx.v / xm, I.dy.v / ym);
    const F dtdx = F(dt / I.dx.v), dtdy = F(dt / I.dy.v);

    // windows of 32 x 32 interior cells; around each, the raw fluxes of
    // [wx - 1, wx + 34]^2 so every FLICD stencil of the window is present
    constexpr int W = 32, PADW = W + 4;
    std::mt19937 rng(seed);
    Tally tw;
    unsigned long long cells = 0;
    std::vector<F> FX((size_t)PADW * PADW * NCONS), FY((size_t)PADW * PADW * NCONS);
    for (int w = 0; w < windows; w++) {
        const int wx = s + (int)(rng() % (unsigned)(S.nx - W + 1)), wy = s + (int)(rng() % (unsigned)(S.ny - W + 1));
        auto at = [&](std::vector<F>& A, int i, int j) { return &A[((size_t)(j - (wy - 1)) * PADW + (i - (wx - 1))) * NCONS]; };
        Tally scratch;
        for (int j = wy - 1; j < wy - 1 + PADW; j++)
            for (int i = wx - 1; i < wx - 1 + PADW; i++) {
                const bool inside = i >= wx && i < wx + W && j >= wy && j < wy + W;
                bool hx, hy;
                face_cell(S, ph, i, j, inside ? tw : scratch, at(FX, i, j), at(FY, i, j), hx, hy);
            }
        // the update of the window's cells (first pass: every cell flagged, scheme R > 0)
        for (int j = wy; j < wy + W; j++)
            for (int i = wx; i < wx + W; i++) {
                auto X = [&](int a, int b) { return at(FX, i + a, j + b)[MAGY]; };
                auto Y = [&](int a, int b) { return at(FY, i + a, j + b)[MAGX]; };
                unsigned long long u0 = fc::n_ops;
                const F cx = flicd_corr(X(0, 0), X(1, 0), X(-1, 0), Y(0, 1), Y(0, 0), Y(-1, 0), Y(-1, 1));
                const F cy = flicd_corr(Y(0, 0), Y(0, 1), Y(0, -1), X(1, 0), X(0, 0), X(0, -1), X(1, -1));
                const F cxe = flicd_corr(X(1, 0), X(2, 0), X(0, 0), Y(1, 1), Y(1, 0), Y(0, 0), Y(0, 1));
                const F cyn = flicd_corr(Y(0, 1), Y(0, 2), Y(0, 0), X(1, 1), X(0, 1), X(0, 0), X(1, 0));
                unsigned long long u1 = fc::n_ops;
                F fxm[NCONS], fxp[NCONS], fym[NCONS], fyp[NCONS], uc[NCONS], un[NCONS];
                for (int v = 0; v < NCONS; v++) {
                    fxm[v] = at(FX, i, j)[v];
                    fxp[v] = at(FX, i + 1, j)[v];
                    fym[v] = at(FY, i, j)[v];
                    fyp[v] = at(FY, i, j + 1)[v];
                    uc[v] = F(S.u(v, i, j));
                }
                fxm[MAGX] = fxp[MAGX] = F(0.0);
                fym[MAGX] = cy;
                fyp[MAGX] = cyn;
                fxm[MAGY] = cx;
                fxp[MAGY] = cxe;
                fym[MAGY] = fyp[MAGY] = F(0.0);
                for (int v = 0; v < NCONS; v++) un[v] = fv_update(uc[v], dtdx, fxp[v], fxm[v], dtdy, fyp[v], fym[v]);
                unsigned long long u2 = fc::n_ops;
                Around a;
                for (int v = 0; v < NCONS; v++) {
                    a.e[v] = F(S.u(v, i + 1, j));
                    a.w[v] = F(S.u(v, i - 1, j));
                    a.n[v] = F(S.u(v, i, j + 1));
                    a.s[v] = F(S.u(v, i, j - 1));
                }
                a.d2[0] = F(S.u(DENS, i + 2, j));
                a.d2[1] = F(S.u(DENS, i - 2, j));
                a.d2[2] = F(S.u(DENS, i, j + 2));
                a.d2[3] = F(S.u(DENS, i, j - 2));
                a.d2[4] = F(S.u(DENS, i + 1, j + 1));
                a.d2[5] = F(S.u(DENS, i + 1, j - 1));
                a.d2[6] = F(S.u(DENS, i - 1, j + 1));
                a.d2[7] = F(S.u(DENS, i - 1, j - 1));
                (void)detect_cell(I, ph, a, un, uc);
                unsigned long long u3 = fc::n_ops;
                tw.flicd += u1 - u0;
                tw.fv += u2 - u1;
                tw.detect += u3 - u2;
                cells += 1;
            }
    }
    // the cells of the face kernel's domain outside the interior, in full
    Tally ring;
    F fxr[NCONS], fyr[NCONS];
    for (int j = s - 2; j <= ey + 3; j++)
        for (int i = s - 2; i <= ex + 3; i++) {
            if (i >= s && i <= ex && j >= s && j <= ey) continue;
            bool hx, hy;
            face_cell(S, ph, i, j, ring, fxr, fyr, hx, hy);
        }
    // the RK combinations of stages 1 and 2 of one variable
    unsigned long long r0 = fc::n_ops;
    (void)rk_combine(1, F(1.0), F(2.0));
    (void)rk_combine(2, F(1.0), F(2.0));
    const double rk = double(fc::n_ops - r0) * NCONS;

    const double c = (double)cells, zones = (double)S.nx * S.ny;
    const double face = tw.face() / c, update = (tw.flicd + tw.fv + tw.detect) / c;
    const double per_zone = 3.0 * (face + update) + rk + cfl_per_zone;
    std::printf(
        "{\"nx\": %d, \"ny\": %d, \"R\": %d, \"windows\": %d, \"cells_counted\": %llu, "
        "\"per_zone\": {\"face\": %.3f, \"recon\": %.3f, \"admissibility\": %.3f, \"hllc\": %.3f, \"flux_sums\": %.3f, "
        "\"update\": %.3f, \"flicd\": %.3f, \"fv_update\": %.3f, \"detect\": %.3f, \"rk_per_step\": %.1f, \"cfl\": %.3f}, "
        "\"hllc_calls_per_zone\": %.4f, \"hllc_flops_per_call\": %.2f, "
        "\"flops_per_zone_update\": %.2f, \"face_flops_per_launch\": %.6e}\n",
        S.nx, S.ny, S.R, windows, cells, face, tw.recon / c, tw.admiss / c, tw.hllc / c, tw.sums / c, update,
        tw.flicd / c, tw.fv / c, tw.detect / c, rk, cfl_per_zone, tw.hllc_calls / c,
        tw.hllc_calls ? (double)tw.hllc / tw.hllc_calls : 0.0, per_zone, face * zones + (double)ring.face());
    return 0;
}
