// hmhd_kernels.hip -- gfx950 kernels of one SSP-RK3 stage and the CFL
// reduction.  A stage is a sequence of MOOD passes; each pass is
//
//   k_face     GP reconstruction of both Riemann states of every owned face
//              from an LDS tile of U, admissibility (CAD/PAD) of the
//              reconstructed states, HLLC at the Gauss points and the
//              weighted sum -> FR  (src/reconstruct.f90:77-141,
//              src/getFluxes.f90:26-88, src/riemannSolver.f90:23-47)
//   k_update   FLICD corrections, the unsplit FV update, MOOD detection and
//              the RK combination of the updated cell (src/FLICDdivB.f90,
//              src/timeStep.f90:107-118, src/detect.f90:74-217,
//              src/timeStep.f90:165,169)
//   k_end_iter the order-lowering bookkeeping after detect
//              (src/gridBlock.f90:194-235, written as a gather, so no cell is
//              written by two threads), the periodic MOOD sync
//              (src/boundaryConditions.f90:189-254), the tiles the next
//              iteration needs, the next inner-loop round's lowering and the
//              loop decisions
//
// The first pass of a stage runs k_face_full: every face at the top radius,
// admissibility checked on the way.  Every later iteration (a new
// MOOD pass, or another round of the inner admissibility loop when a
// reconstructed state was inadmissible) runs k_face on the active tiles
// only (the inner round's lowering ran in the scan launch before it): of
// the faces the reference would rebuild (flagged
// sides, fluxes getFluxes_ recomputes), those whose inputs changed are
// recomputed, each side from the radius it was last reconstructed with,
// and the others only repeat the recomputation's side effects.
#include <hip/hip_runtime.h>

#include "hmhd_internal.hpp"
#include "hmhd_kernels.hpp"
#include "hmhd_math.hpp"

namespace hmhd {

__constant__ double c_pv[kMaxGpSets][kPvSet];
__constant__ double c_qw[kMaxGpSets][kMaxR + 1][4];

// padded index -> interior index along one axis (guard fill as index map)
__device__ __forceinline__ int map_axis(int p, int ngc, int n, int bc) {
    int k = p - ngc;
    if (bc == BC_PERIODIC) {
        k += (k < 0) ? n : 0;
        k -= (k >= n) ? n : 0;
    } else {
        k = k < 0 ? 0 : (k > n - 1 ? n - 1 : k);
    }
    return k;
}

__device__ __forceinline__ Phys phys_of(const Inst& I) {
    Phys ph;
    ph.gamma = I.gamma;
    ph.small_p = I.small_p;
    ph.small_rho = I.small_rho;
    ph.hydro = I.hydro;
    return ph;
}

template <int LX, int LP>
struct LdsAcc {
    const double* s;  // points at the centre cell of plane 0
    __device__ __forceinline__ double at(int v, int dx, int dy) const { return s[v * LP + dy * LX + dx]; }
};

struct GlobAcc {
    const double* U;
    size_t plane;
    int i, j, ngc, nx, ny, bc;
    __device__ __forceinline__ double at(int v, int dx, int dy) const {
        int kx = map_axis(i + dx, ngc, nx, bc), ky = map_axis(j + dy, ngc, ny, bc);
        return U[v * plane + (size_t)ky * nx + kx];
    }
};

__device__ __forceinline__ const double* pv_ptr(int set, int rr, int f, int q) {
    return &c_pv[set][gp_pv_off(rr) + (f * gp_nq(rr) + q) * gp_ns(rr)];
}

// gp_point for the later passes: the weights and stencil values of CH
// cells are read from LDS together, then summed in stencil order (the same
// operations in the same order).  Left to itself, the compiler, under the
// register pressure of the later-pass face kernel, waits for each read
// before the next, a round trip per read.
template <int RR, int CH, class A>
__device__ __forceinline__ void gp_point_chunked(const A& a, const double* __restrict__ w, double* out) {
    constexpr Sten<RR> st{};
#pragma unroll
    for (int v = 0; v < NCONS; v++) out[v] = 0.0;
#pragma unroll
    for (int k0 = 0; k0 < st.n; k0 += CH) {
        double x[CH][NCONS], wk[CH];
#pragma unroll
        for (int c = 0; c < CH; c++) {
            if (k0 + c < st.n) {
                wk[c] = w[k0 + c];
#pragma unroll
                for (int v = 0; v < NCONS; v++) x[c][v] = a.at(v, st.x[k0 + c], st.y[k0 + c]);
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int c = 0; c < CH; c++) {
            if (k0 + c < st.n) {
#pragma unroll
                for (int v = 0; v < NCONS; v++) out[v] = out[v] + wk[c] * x[c][v];
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// value of one Riemann state at point q from a side reconstructed with
// radius rr (rr == 0: the cell average itself); pv: a copy of the solve's
// table c_pv[set] (in LDS)
template <int R, class A>
__device__ __forceinline__ void side_value(const A& a, const double* pv, int rr, int f, int q, double* out) {
    constexpr int CH = 5;
    if (rr == 0) {
#pragma unroll
        for (int v = 0; v < NCONS; v++) out[v] = a.at(v, 0, 0);
    } else if (rr == R) {
        gp_point_chunked<R, CH>(a, pv + gp_pv_off(R) + (f * gp_nq(R) + q) * gp_ns(R), out);
    } else if (R >= 3 && rr == 2) {
        gp_point_chunked<(R >= 3 ? 2 : 1), CH>(a, pv + gp_pv_off(2) + (f * gp_nq(2) + q) * gp_ns(2), out);
    } else {
        gp_point_chunked<1, CH>(a, pv + gp_pv_off(1) + (f * gp_nq(1) + q) * gp_ns(1), out);
    }
}

enum : int { FACE_N = 0, FACE_S = 1, FACE_E = 2, FACE_W = 3 };

// Compiler-only memory barrier: the stencil values of a face are re-read
// from LDS for each Gauss point instead of being held in registers across
// the point loop (2 x 13 x 8 doubles at R = 2 would not fit).
__device__ __forceinline__ void reload_barrier() { asm volatile("" ::: "memory"); }

// The control words of one solve, read in one 16-byte load.  Within a
// kernel the words a kernel reads do not change (a kernel writes only
// words it does not read).
struct Ctl {
    int32_t any_inadm, any_lower, run, mode;
};
static_assert(CTL_ANY_INADM == 0 && CTL_ANY_LOWER == 1 && CTL_RUN == 2 && CTL_MODE == 3, "Ctl field order");
__device__ __forceinline__ Ctl ctl_of(const int32_t* c) {
    const int4 v = *reinterpret_cast<const int4*>(c);
    return Ctl{v.x, v.y, v.z, v.w};
}

// the rest of a MOOD pass (fluxes, update, detect, lowering, sync) runs
// only for instances inside their MOOD loop whose inner admissibility loop
// has converged
__device__ __forceinline__ bool pass_ready(const Ctl& c) { return c.run != 0 && c.any_inadm == 0; }

// whether the pass ends with the periodic MOOD sync: a periodic solve whose
// pass ran (its detect lowering did)
__device__ __forceinline__ bool sync_due(const Inst& I, const Ctl& c) {
    return I.bc == BC_PERIODIC && I.active && pass_ready(c);
}

// The record of a tile's solve: with ONE (a single solve) the kernel
// argument, so its fields do not wait for the tile's list entry.
template <bool ONE>
__device__ __forceinline__ const Inst& inst_at(const Inst* __restrict__ insts, const Inst& one, int k) {
    return ONE ? one : insts[k];
}

// Workgroups go to the 8 XCDs round-robin, and each XCD has its own L2.
// When a grid has one block per tile of a row-major list, block b takes
// tile xcd_tile(b, n): every XCD gets one contiguous run of the list (a
// band of tile rows), so neighbouring tiles share their halo lines in one
// L2 instead of fetching them on several XCDs.  A bijection on [0, n).
__device__ __forceinline__ int xcd_tile(int b, int n) {
    constexpr int kXcd = 8;
    const int q = n / kXcd, r = n % kXcd, x = b % kXcd;
    return x * q + (x < r ? x : r) + b / kXcd;
}

// Marks values as used at this point, so their loads are issued here
// (together with whatever else is in flight) rather than sunk past a branch.
__device__ __forceinline__ void use_now(int a) { asm volatile("" ::"v"(a)); }
__device__ __forceinline__ void use_now(double a) { asm volatile("" ::"v"(a)); }
template <class... T>
__device__ __forceinline__ void use_all(T... v) {
    (use_now(v), ...);
}

// The tiles of list `tl` this block takes: entries b, b + G, ... below the
// count.  The block's first entry, the count and the caller's `pin` value
// (its control words) are loaded together, in one round trip; the launch
// has at most as many blocks as the list has room for.  `idle`, a value
// computed from those control words: the launch has nothing to do, and
// the block leaves before walking its entries.
template <class F>
__device__ __forceinline__ void for_tiles(const TileList& tl, F&& body, int pin = 0, bool idle = false) {
    const int b0 = blockIdx.x, G = gridDim.x;
    Item it = tl.items[tl.swz ? xcd_tile(b0, G) : b0];
    const int n = *tl.cnt;
    use_now(it.inst);
    use_now(it.tx);
    use_now(pin);
    if (idle) return;
    for (int b = b0; b < n;) {
        body(it);
        b += G;
        if (b < n) it = tl.items[b];  // only lists without swz have more entries than blocks
    }
}

// The later-pass update runs a tile as kUpdQ blocks of one wave, kUpdRows
// rows of the tile each: a lone wave's loads are bound by the cache lines
// its CU has in flight, so a tile's loads spread over four CUs finish
// sooner than over one.
constexpr int kUpdRows = 4, kUpdQ = kCellT / kUpdRows;
static_assert(kUpdRows * kCellT == 64, "one wave per quarter tile");

// for_tiles for P blocks per entry: block b takes part b % P of entry b / P
// (with swz, of entry xcd_tile(b / P, G / P)), then strides by G / P
// entries; G is a multiple of P.  body(it, part).
template <int P, class F>
__device__ __forceinline__ void for_tile_parts(const TileList& tl, F&& body, int pin = 0, bool idle = false) {
    const int b0 = blockIdx.x, G = gridDim.x / P, e0 = b0 / P, part = b0 % P;
    Item it = tl.items[tl.swz ? xcd_tile(e0, G) : e0];
    const int n = *tl.cnt;
    use_now(it.inst);
    use_now(it.tx);
    use_now(pin);
    if (idle) return;
    for (int e = e0; e < n;) {
        body(it, part);
        e += G;
        if (e < n) it = tl.items[e];  // only lists without swz have more entries than block groups
    }
}

// SSP-RK3 buffer roles from the solve's rotation a = I.rot (U^n in U[a]):
// stage 0 reads U[a] and writes U1 to U[a+1]; stage 1 reads U[a+1] and
// writes U2 to U[a+2]; stage 2 reads U[a+2] and writes U^{n+1} to U[a+1];
// stages 1 and 2 combine with U^n (src/timeStep.f90:161-169).
struct StageBufs {
    int in, out, base;
};
__device__ __forceinline__ StageBufs stage_bufs(int rot, int stage) {
    // in = rot + stage, out = rot + 1 (rot + 2 at stage 1), base = rot, mod 3
    const int in = rot + stage, out = rot + 1 + (stage == 1 ? 1 : 0);
    return {in >= 3 ? in - 3 : in, out >= 3 ? out - 3 : out, rot};
}

// state buffer k of a solve, selected without indexing the pointer array
__device__ __forceinline__ double* ubuf(const Inst& I, int k) { return k == 0 ? I.U[0] : (k == 1 ? I.U[1] : I.U[2]); }

// ------------------------------------------------------------------ //
// stage set-up: rc = rcF = true everywhere (src/timeStep.f90:80-81);
// inadm cleared; every face at radius R; every tile's bad_bits word cleared.  The scheme's reset to R
// (src/timeStep.f90:79) is left to the first-pass k_update, the first
// kernel to read it: the scheme keeps the last step's values until dt is
// known, though cfl() enqueues the start of the next step's first pass.
// Block 0 also starts the stage's MOOD loop, one thread per instance
// (inactive instances stay out of the loop), and empties the active list
// the first pass's scan writes; no other block reads the control words.
__global__ __launch_bounds__(256) void k_stage_init(const Inst* __restrict__ insts, const Item* __restrict__ items,
                                                    int npad, int ninst, int stage, int32_t* __restrict__ scan_count) {
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) *reinterpret_cast<unsigned long long*>(scan_count) = 0;  // both counts of list 0
        for (int k = threadIdx.x; k < ninst; k += blockDim.x) {
            int32_t* c = insts[k].ctl;
            if (!insts[k].active) {
                c[CTL_RUN] = 0;
                continue;
            }
            c[CTL_ANY_INADM] = 0;
            c[CTL_ANY_LOWER] = 0;
            c[CTL_RUN] = 1;
            c[CTL_MODE] = MODE_NEWPASS;
            c[CTL_PASSES + stage] = 1;
            c[CTL_INNER + stage] = 1;
            c[CTL_FIRST] = 1;
        }
    }
    if ((int)blockIdx.x >= npad) return;
    const Item it = items[xcd_tile(blockIdx.x, gridDim.x)];
    const Inst& I = insts[it.inst];
    // the first pass's face kernel ORs into the tile's bad_bits word
    if (threadIdx.x == 0) {
        I.tbad[it.ty * I.ntx + it.tx] = 0;
        I.tbadh[it.ty * I.ntx + it.tx] = 0;
    }
    const int i = it.tx * kCellT + (threadIdx.x % kCellT), j = it.ty * kCellT + (threadIdx.x / kCellT);
    if (i >= I.mx || j >= I.my) return;
    const size_t idx = (size_t)j * I.mx + i, plane = (size_t)I.mx * I.my;
    const uint8_t R = (uint8_t)I.R;
    I.rc[idx] = 1;
    I.rcF[idx] = 1;
#pragma unroll
    for (int f = 0; f < 4; f++) {
        I.faceRR[f * plane + idx] = R;
        I.inadm[f * plane + idx] = 0;
    }
}


// ------------------------------------------------------------------ //
// face kernels.  Padded cell c owns the x-face at i-1/2 (the east state of
// its west neighbour and its own west state) and the y-face at j-1/2
// (north state of the south neighbour, own south state); cells of
// [s-2, e+3]^2 take part.
// The first pass of a stage (k_face_full): every face at the top radius,
// admissibility checked on the way, one cell's two faces per thread, in
// tiles of a tuned shape (see launch_face_full).
// Later passes (k_face, face_tile_iter): one iteration of the inner
// admissibility loop fused with the flux recomputation, over a compacted
// list of the faces whose inputs changed, on the 16 x 16 tiles every
// per-cell kernel uses, a tile as two blocks of eight rows each (blocks
// stride over the tile list), so a tile's loads are in flight on two CUs;
// at R <= 2 the block has 256 threads, one per face of its half.
// Inadmissibility is recorded per face side in four byte planes, so each
// byte has a single writer.  Register budget: two waves per SIMD at R <= 2;
// R = 3 keeps one (its stencils would spill at two).
constexpr int kFaceTYH = kFaceTY / 2;  // rows of a tile per later-pass face block
__host__ __device__ constexpr int face_threads(int R) { return (R >= 3 ? 1 : 2) * kFaceTX * kFaceTYH; }
template <int R>
constexpr int face_lds_cells() {
    return (kFaceTX + 2 * R + 1) * (kFaceTYH + 2 * R + 1);
}

// Where in a tile the cell of an inadmissible face side lies, for the scan
// of the tile and of its eight neighbours: bit 0 for any, then bits for a
// cell within two cells of the tile's west, east, south and north edge
// (BAD_W2 .. BAD_N2), then the same within four cells (BAD_W4 .. BAD_N4).
// (rx, ry): the cell relative to the tile's origin; -1 for a cell of the
// west / south neighbour, whose east / north side belongs to this tile's
// face.
enum : uint32_t {
    BAD_ANY = 1u, BAD_W2 = 2u, BAD_E2 = 4u, BAD_S2 = 8u, BAD_N2 = 16u,
    BAD_W4 = 32u, BAD_E4 = 64u, BAD_S4 = 128u, BAD_N4 = 256u, BAD_NBITS = 9
};
__device__ __forceinline__ uint32_t bad_bits(int rx, int ry) {
    return BAD_ANY | (rx <= 1 ? BAD_W2 : 0u) | (rx >= kCellT - 2 ? BAD_E2 : 0u) | (ry <= 1 ? BAD_S2 : 0u) |
           (ry >= kCellT - 2 ? BAD_N2 : 0u) | (rx <= 3 ? BAD_W4 : 0u) | (rx >= kCellT - 4 ? BAD_E4 : 0u) |
           (ry <= 3 ? BAD_S4 : 0u) | (ry >= kCellT - 4 ? BAD_N4 : 0u);
}
// The end of a face tile: the bad_bits of the wave's active lanes go into
// the tile's word with one atomic, and ANY_INADM is raised, by the first
// active lane (a tile with many inadmissible sides would otherwise queue
// hundreds of atomics on one word).
__device__ __forceinline__ void report_bad(const Inst& I, uint32_t* tbad, uint32_t bits) {
    uint32_t w = 0;
#pragma unroll
    for (int k = 0; k < BAD_NBITS; k++) w |= __ballot((bits >> k) & 1u) != 0 ? 1u << k : 0u;
    if (w && (int)(threadIdx.x & 63u) == __ffsll(__ballot(1)) - 1) {
        I.ctl[CTL_ANY_INADM] = 1;
        atomicOr(tbad, w);
    }
}

// The bad_bits of a wave whose lanes' cells lie in at most two 16 x 16
// tiles (a first-pass face tile is 16 or 32 cells wide, its rows within one
// band of 16): each tile's word gets one atomic OR, and ANY_INADM is raised.
// rx, ry of a lane's bits are relative to its own cell's 16 x 16 tile.
__device__ __forceinline__ void report_bad_cells(const Inst& I, uint32_t bits, int i, int j) {
    const unsigned long long has = __ballot(bits != 0);
    if (!has) return;
    const int tt = (j / kCellT) * I.ntx + i / kCellT;
    const int la = __ffsll(has) - 1;
    const int ta = __shfl(tt, la);
    const bool inA = bits != 0 && tt == ta, inB = bits != 0 && tt != ta;
    uint32_t wa = 0, wb = 0;
#pragma unroll
    for (int k = 0; k < BAD_NBITS; k++) {
        wa |= __ballot(inA && ((bits >> k) & 1u)) != 0 ? 1u << k : 0u;
        wb |= __ballot(inB && ((bits >> k) & 1u)) != 0 ? 1u << k : 0u;
    }
    const unsigned long long hb = __ballot(inB);
    const int lane = (int)(threadIdx.x & 63u);
    if (lane == la) {
        I.ctl[CTL_ANY_INADM] = 1;
        atomicOr(I.tbad + ta, wa);
    }
    if (hb && lane == __ffsll(hb) - 1) atomicOr(I.tbad + tt, wb);
}

// One first-pass face tile of TX x TY cells at (i0, j0), one cell per
// thread (its x face at i-1/2 and its y face at j-1/2): every state of
// [s-2, e+2]^2 at radius R, checked, and every flux of [s-1, e+2]^2 (the
// range src/reconstruct.f90:81-83 and src/getFluxes.f90:26-30 cover).  The
// per-tile bad_bits words were cleared by the stage set-up.  FLB is not
// written here: every cell is in rcF until the stage's first update, which
// writes FLB at every cell of [s, e+1]^2 before anything reads it (the
// update reads a neighbour's FLB only where that neighbour is not in rcF).
template <int R, int TX, int TY>
__device__ __forceinline__ void face_tile_full(const Inst& I, int i0, int j0, int stage, double* __restrict__ sU,
                                               double* __restrict__ sAcc) {
    constexpr int H = R + 1, NT = TX * TY;
    // The flux sums over the Gauss points: in registers, except at radius 1,
    // where the 16 registers they take would cost a wave per SIMD (194
    // VGPRs); there they are a column of LDS per thread (stride NT).
    constexpr bool kAccLds = R == 1;
    constexpr int AS = kAccLds ? NT : 1;
    constexpr int LX = TX + 2 * R + 1, LP = LX * (TY + 2 * R + 1);

    if (I.R != R) return;
    const int s = I.ngc, ex = I.ngc + I.nx - 1, ey = I.ngc + I.ny - 1, mx = I.mx;
    const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
    const int i = i0 + lx, j = j0 + ly;
    const bool in_dom = i >= s - 2 && i <= ex + 3 && j >= s - 2 && j <= ey + 3;
    const bool flux_cell = in_dom && i >= s - 1 && i <= ex + 2 && j >= s - 1 && j <= ey + 2;
    const size_t idx = (size_t)j * mx + i, fplane = (size_t)mx * I.my;
    // reconstruction range [s-2, e+2]^2 for the four cells around the two faces
    const bool cellC = in_dom && i <= ex + 2 && j <= ey + 2;
    const bool cellW = in_dom && i >= s - 1 && j <= ey + 2;
    const bool cellS = in_dom && j >= s - 1 && i <= ex + 2;

    const bool recC = cellC, recW = cellW, recS = cellS, fx = flux_cell, fy = flux_cell;
    const bool do_x = recC || recW || fx, do_y = recC || recS || fy;

    // stage the U tile [i0-H, i0+TX-1+R] x [j0-H, j0+TY-1+R] in LDS
    {
        const double* __restrict__ U = ubuf(I, stage_bufs(I.rot, stage).in);
        const size_t plane = (size_t)I.nx * I.ny;
        const int oi = i0 - H, oj = j0 - H;
        for (int c = threadIdx.x; c < LP; c += NT) {
            const int cy = c / LX, cx = c - cy * LX;
            const int kx = map_axis(oi + cx, s, I.nx, I.bc), ky = map_axis(oj + cy, s, I.ny, I.bc);
            const double* src = U + (size_t)ky * I.nx + kx;
#pragma unroll
            for (int v = 0; v < NCONS; v++) sU[v * LP + c] = src[v * plane];
        }
    }
    __syncthreads();
    uint32_t bits = 0;
    // the cell's position in its 16 x 16 tile, for bad_bits
    const int rx = i % kCellT, ry = j % kCellT;
    if (do_x || do_y) {
        const Phys ph = phys_of(I);
        const int set = I.gpset;
        const int cc = (ly + H) * LX + (lx + H);  // LDS offset of cell (i, j)
        const LdsAcc<LX, LP> aC{sU + cc}, aW{sU + cc - 1}, aS{sU + cc - LX};
        uint8_t* const bad = I.inadm;  // planes [face][my][mx]
        constexpr int NQ = gp_nq(R);
        // x-face: east state of (i-1, j) | west state of (i, j)
        if (do_x) {
            bool badL = false, badR = false;
            double accr[NCONS];
            double* const acc = kAccLds ? sAcc + threadIdx.x : accr;
#pragma unroll 1
            for (int q = 0; q < NQ; q++) {
                double uL[NCONS], uR[NCONS];
                reload_barrier();
                gp_point<R>(aW, pv_ptr(set, R, FACE_E, q), uL);
                gp_point<R>(aC, pv_ptr(set, R, FACE_W, q), uR);
                badL |= !admissible(uL[DENS], cons2pres(ph, uL));
                badR |= !admissible(uR[DENS], cons2pres(ph, uR));
                if (fx) {
                    double F[NCONS];
                    hllc<0>(ph, uL, uR, F);
                    const double wq = c_qw[set][R][q];
#pragma unroll
                    for (int v = 0; v < NCONS; v++) acc[v * AS] = (q == 0 ? 0.0 : acc[v * AS]) + wq * F[v];
                }
            }
            if (recW) bad[FACE_E * fplane + idx - 1] = badL;
            if (recC) bad[FACE_W * fplane + idx] = badR;
            if (recW && badL) bits |= bad_bits(rx - 1, ry);
            if (recC && badR) bits |= bad_bits(rx, ry);
            if (fx) {
#pragma unroll
                for (int v = 0; v < NCONS; v++) I.FR[v * fplane + idx] = acc[v * AS];
                I.fluxRR[idx] = (uint8_t)R;
            }
        }
        // y-face: north state of (i, j-1) | south state of (i, j)
        if (do_y) {
            bool badL = false, badR = false;
            double accr[NCONS];
            double* const acc = kAccLds ? sAcc + threadIdx.x : accr;
#pragma unroll 1
            for (int q = 0; q < NQ; q++) {
                double uL[NCONS], uR[NCONS];
                reload_barrier();
                gp_point<R>(aS, pv_ptr(set, R, FACE_N, q), uL);
                gp_point<R>(aC, pv_ptr(set, R, FACE_S, q), uR);
                badL |= !admissible(uL[DENS], cons2pres(ph, uL));
                badR |= !admissible(uR[DENS], cons2pres(ph, uR));
                if (fy) {
                    double F[NCONS];
                    hllc<1>(ph, uL, uR, F);
                    const double wq = c_qw[set][R][q];
#pragma unroll
                    for (int v = 0; v < NCONS; v++) acc[v * AS] = (q == 0 ? 0.0 : acc[v * AS]) + wq * F[v];
                }
            }
            if (recS) bad[FACE_N * fplane + idx - mx] = badL;
            if (recC) bad[FACE_S * fplane + idx] = badR;
            if (recS && badL) bits |= bad_bits(rx, ry - 1);
            if (recC && badR) bits |= bad_bits(rx, ry);
            if (fy) {
#pragma unroll
                for (int v = 0; v < NCONS; v++) I.FR[(NCONS + v) * fplane + idx] = acc[v * AS];
                I.fluxRR[fplane + idx] = (uint8_t)R;
            }
        }
    }
    report_bad_cells(I, bits, i, j);
}

// One face of a later MOOD iteration on a group of n = rr_now + 1 lanes
// (the face's Gauss points; one for a first-order face), lanes base ..
// base + n - 1 of the wave, lane base + q for point q
// (src/reconstruct.f90:109-133, src/riemannSolver.f90:37-44):
// the lanes rebuild their point of each side to be rebuilt (at the current
// radius) or needed for the flux (at the radius the side was last rebuilt
// with), check its admissibility, and form w_q F_HLLC; lane q = 0 sums the
// points in ascending order from zero (or takes the one first-order flux
// as it is) and writes the face's results.
// rr_now: the face's radius, min of its cells' schemes; storedL/R: the
// radius each side was last reconstructed with; pv, qw: the solve's point
// and quadrature weights (c_pv[set], c_qw[set]) as copied to LDS.
template <int R, int DIR, class A>
__device__ __forceinline__ void face_group(const Inst& I, const Phys& ph, const double* pv, const double* qw,
                                          const A& aL, const A& aR, size_t idx,
                                          size_t nb, bool recL, bool recR, bool flux, int rr_now, int storedL,
                                          int storedR, int q, int base, int n, int lx, int ly, uint32_t& bits) {
    constexpr int G = gp_nq(R);  // the largest group
    const size_t fplane = (size_t)I.mx * I.my;
    constexpr int fL = DIR == 0 ? FACE_E : FACE_N, fR = DIR == 0 ? FACE_W : FACE_S;
    const int rrL = recL ? rr_now : storedL;
    const int rrR = recR ? rr_now : storedR;
    const int nqL = (recL && rrL > 0) ? gp_nq(rrL) : 0;
    const int nqR = (recR && rrR > 0) ? gp_nq(rrR) : 0;
    const int nqF = flux ? (rr_now == 0 ? 1 : gp_nq(rr_now)) : 0;
    double uL[NCONS], uR[NCONS], t[NCONS];
    int badL = 0, badR = 0;
    if (q < nqL || q < nqF) side_value<R>(aL, pv, rrL, fL, q, uL);
    if (q < nqR || q < nqF) side_value<R>(aR, pv, rrR, fR, q, uR);
    if (q < nqL) badL = !admissible(uL[DENS], cons2pres(ph, uL));
    if (q < nqR) badR = !admissible(uR[DENS], cons2pres(ph, uR));
    if (q < nqF) {
        double F[NCONS];
        hllc<DIR>(ph, uL, uR, F);
        const double w = qw[rr_now * 4 + q];
#pragma unroll
        for (int v = 0; v < NCONS; v++) t[v] = rr_now == 0 ? F[v] : w * F[v];
    } else {
#pragma unroll
        for (int v = 0; v < NCONS; v++) t[v] = 0.0;
    }
    // every lane of the group takes part in the exchanges
    double acc[NCONS];
#pragma unroll
    for (int v = 0; v < NCONS; v++) {
        double a = rr_now == 0 ? t[v] : 0.0 + t[v];
#pragma unroll
        for (int k = 1; k < G; k++) {
            // a lane past the group reads its own value, which is not added
            const double tk = __shfl(t[v], k < n ? base + k : base);
            if (nqF > k) a = a + tk;
        }
        acc[v] = a;
    }
    int bl = badL, br = badR;
#pragma unroll
    for (int k = 1; k < G; k++) {
        const int lk = k < n ? base + k : base;
        bl |= __shfl(badL, lk);
        br |= __shfl(badR, lk);
    }
    badL = bl;
    badR = br;
    if (q != 0) return;
    if (recL) {
        I.faceRR[fL * fplane + nb] = (uint8_t)rr_now;
        I.inadm[fL * fplane + nb] = (uint8_t)badL;
    }
    if (recR) {
        I.faceRR[fR * fplane + idx] = (uint8_t)rr_now;
        I.inadm[fR * fplane + idx] = (uint8_t)badR;
    }
    if (recL && badL) bits |= bad_bits(DIR == 0 ? lx - 1 : lx, DIR == 0 ? ly : ly - 1);
    if (recR && badR) bits |= bad_bits(lx, ly);
    if (flux) {
#pragma unroll
        for (int v = 0; v < NCONS; v++) I.FR[(DIR * NCONS + v) * fplane + idx] = acc[v];
        I.FLB[DIR * fplane + idx] = acc[DIR == 0 ? MAGY : MAGX];
        I.fluxRR[DIR * fplane + idx] = (uint8_t)rr_now;
    }
}

// x from the other lane of this lane's pair (lanes 2k, 2k + 1), by DPP;
// both lanes must be active
__device__ __forceinline__ double pair_swap(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0xB1, 0xF, 0xF, false);  // quad_perm [1, 0, 3, 2]
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0xB1, 0xF, 0xF, false);
    return __longlong_as_double((long long)hi << 32 | (unsigned int)lo);
}

// One face of a later MOOD iteration on 2 (R + 1) lanes, lane base + 2q + sd
// for Gauss point q of side sd (0: L, 1: R): face_group with the two sides
// on a pair of lanes instead of one lane, for the single-solve kernel,
// whose sparse passes wait on the dependency chain of one face.  Each lane
// rebuilds its side at its point and checks its admissibility and forms
// the side's part of HLLC (hllc_side); the pair exchanges states and side
// parts, and lane 2q finishes HLLC with its own side as L (the R lane's
// result is not used).  Lane base sums the points in ascending order from
// zero and writes the face's results, as in face_group.
template <int R, int DIR, class A>
__device__ __forceinline__ void face_group_pair(const Inst& I, const Phys& ph, const double* pv, const double* qw,
                                               const A& aS, size_t idx, size_t nb, bool recL, bool recR, bool flux,
                                               int rr_now, int storedL, int storedR, int q, int sd, int base, int lx,
                                               int ly, uint32_t& bits) {
    constexpr int G = gp_nq(R);
    const size_t fplane = (size_t)I.mx * I.my;
    constexpr int fL = DIR == 0 ? FACE_E : FACE_N, fR = DIR == 0 ? FACE_W : FACE_S;
    const int rrL = recL ? rr_now : storedL;
    const int rrR = recR ? rr_now : storedR;
    const int nqL = (recL && rrL > 0) ? gp_nq(rrL) : 0;
    const int nqR = (recR && rrR > 0) ? gp_nq(rrR) : 0;
    const int nqF = flux ? (rr_now == 0 ? 1 : gp_nq(rr_now)) : 0;
    const int rrS = sd ? rrR : rrL, nqS = sd ? nqR : nqL;
    double u[NCONS], t[NCONS];
    int bad = 0;
    if (q < nqS || q < nqF) side_value<R>(aS, pv, rrS, sd ? fR : fL, q, u);
    if (q < nqS) bad = !admissible(u[DENS], cons2pres(ph, u));
    if (q < nqF) {
        // both lanes of the pair are here (q and the face are theirs alike)
        const HSide h = hllc_side<DIR>(ph, u);
        HSide ho;
        double uo[NCONS];
#pragma unroll
        for (int v = 0; v < NCONS; v++) {
            uo[v] = pair_swap(u[v]);
            ho.f[v] = pair_swap(h.f[v]);
        }
        ho.V.d = pair_swap(h.V.d);
        ho.V.vx = pair_swap(h.V.vx);
        ho.V.vy = pair_swap(h.V.vy);
        ho.V.vz = pair_swap(h.V.vz);
        ho.V.p = pair_swap(h.V.p);
        ho.V.bx = pair_swap(h.V.bx);
        ho.V.by = pair_swap(h.V.by);
        ho.V.bz = pair_swap(h.V.bz);
        ho.c = pair_swap(h.c);
        ho.mdv = pair_swap(h.mdv);
        ho.pTot = pair_swap(h.pTot);
        double F[NCONS];
        hllc_from_sides<DIR>(ph, u, uo, h, ho, F);
        const double w = qw[rr_now * 4 + q];
#pragma unroll
        for (int v = 0; v < NCONS; v++) t[v] = rr_now == 0 ? F[v] : w * F[v];
    } else {
#pragma unroll
        for (int v = 0; v < NCONS; v++) t[v] = 0.0;
    }
    // every lane of the group takes part in the exchanges
    double acc[NCONS];
#pragma unroll
    for (int v = 0; v < NCONS; v++) {
        double a = rr_now == 0 ? t[v] : 0.0 + t[v];
#pragma unroll
        for (int k = 1; k < G; k++) {
            const double tk = __shfl(t[v], base + 2 * k);
            if (nqF > k) a = a + tk;
        }
        acc[v] = a;
    }
    int badL = 0, badR = 0;
#pragma unroll
    for (int k = 0; k < G; k++) {
        badL |= __shfl(bad, base + 2 * k);
        badR |= __shfl(bad, base + 2 * k + 1);
    }
    if (q != 0 || sd != 0) return;
    if (recL) {
        I.faceRR[fL * fplane + nb] = (uint8_t)rr_now;
        I.inadm[fL * fplane + nb] = (uint8_t)badL;
    }
    if (recR) {
        I.faceRR[fR * fplane + idx] = (uint8_t)rr_now;
        I.inadm[fR * fplane + idx] = (uint8_t)badR;
    }
    if (recL && badL) bits |= bad_bits(DIR == 0 ? lx - 1 : lx, DIR == 0 ? ly : ly - 1);
    if (recR && badR) bits |= bad_bits(lx, ly);
    if (flux) {
#pragma unroll
        for (int v = 0; v < NCONS; v++) I.FR[(DIR * NCONS + v) * fplane + idx] = acc[v];
        I.FLB[DIR * fplane + idx] = acc[DIR == 0 ? MAGY : MAGX];
        I.fluxRR[DIR * fplane + idx] = (uint8_t)rr_now;
    }
}

// Tiles whose flags this iteration may change, for the scan that ends it:
// within one tile of an active tile (every kernel writes flags of its own
// tile's cells, the face kernel also the east / north sides of the cells
// west / south of it), and, for a periodic solve, around the images of the
// tile by +-N in x and y, where the MOOD sync copies guard-band flags.
// A tile outside this set had no flag in its footprint after the previous
// iteration and gains none in this one, so its scan would come out empty.
__device__ __forceinline__ void mark_candidates(const Inst& I, const Item it, uint32_t* __restrict__ marks,
                                                uint32_t stamp) {
    const int t = threadIdx.x;
    if (t >= 9 || (t > 0 && I.bc != BC_PERIODIC)) return;
    const int sx = (t % 3 == 1) ? I.nx : (t % 3 == 2 ? -I.nx : 0);
    const int sy = (t / 3 == 1) ? I.ny : (t / 3 == 2 ? -I.ny : 0);
    const int nty = (I.my + kCellT - 1) / kCellT;
    const int x0 = it.tx * kCellT + sx - kCellT, x1 = it.tx * kCellT + sx + 2 * kCellT - 1;
    const int y0 = it.ty * kCellT + sy - kCellT, y1 = it.ty * kCellT + sy + 2 * kCellT - 1;
    if (x1 < 0 || y1 < 0 || x0 >= I.mx || y0 >= I.my) return;
    const int ta = x0 < 0 ? 0 : x0 / kCellT, tb = x1 >= I.mx ? I.ntx - 1 : x1 / kCellT;
    const int ua = y0 < 0 ? 0 : y0 / kCellT, ub = y1 >= I.my ? nty - 1 : y1 / kCellT;
    for (int u = ua; u <= ub; u++)
        for (int v = ta; v <= tb; v++) marks[I.tile_off + u * I.ntx + v] = stamp;
}

// A later MOOD iteration on one tile.  Every face the reference's pass
// would touch -- a side to re-reconstruct and check, a flux to recompute
// -- is classified first.  A face whose sides would be rebuilt at the radii
// they already have, and whose flux would be recomputed at the radius it
// was last computed with, would reproduce the stored states, admissibility
// flags and flux bit for bit (the stage input U and the GP tables do not
// change within a stage), so only the recomputation's side effects are
// repeated: the raw-flux copy into FLB and the inadmissibility report.  The
// remaining faces are compacted, x-faces then y-faces from a wave boundary,
// so lanes do not idle behind untouched faces.  BYR: lists by radius as
// well as direction (the multi-solve kernel, whose launches fill the chip);
// the single-solve kernel, latency-bound on sparse tiles, keeps lists by
// direction only.
template <int R, bool BYR>
__device__ __forceinline__ void face_tile_iter(const Inst& I, const Ctl& cw, const Item it, int hh, int stage,
                                               double* __restrict__ sU, uint32_t* __restrict__ lst,
                                               int* __restrict__ lcnt, double* __restrict__ sPV,
                                               double* __restrict__ sQW) {
    // half hh of the tile: its rows hh TY .. hh TY + TY - 1
    constexpr int TX = kFaceTX, TY = kFaceTYH, NC = TX * TY, H = R + 1, NT = face_threads(R);
    constexpr int LX = TX + 2 * R + 1, LP = face_lds_cells<R>();

    if (I.R != R || !cw.run) return;
    const bool inner = cw.mode == MODE_INNER;
    const int s = I.ngc, ex = I.ngc + I.nx - 1, ey = I.ngc + I.ny - 1, mx = I.mx;
    const int i0 = it.tx * TX, j0 = it.ty * kFaceTY + hh * TY, ly0 = hh * TY;  // ly0: the half's first tile row
    const size_t fplane = (size_t)mx * I.my;
    const uint8_t* flag = cw.mode == MODE_NEWPASS ? I.rc : I.rec;
    uint32_t* const tbad = (hh ? I.tbadh : I.tbad) + it.ty * I.ntx + it.tx;
    // the faces to rebuild, in lists by direction and (BYR) radius: bucket
    // d K + rr of lst (K = R + 1 with BYR, else 1 and rr = 0; NC entries
    // each), counts in lcnt
    constexpr int G = gp_nq(R), K = BYR ? G : 1, NB = 2 * K;
    if (threadIdx.x < NB) lcnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) *tbad = 0;  // before the barrier, so the later ORs land on it
    __syncthreads();

    // With one face per thread and a U tile of at most two cells per thread,
    // each thread's cells of the tile are loaded together with its face's
    // classification values, and written to LDS only if the tile has faces
    // to rebuild: the rebuild waits on no further round trip.
    constexpr int KP = (LP + NT - 1) / NT;
    constexpr bool kPre = NT == 2 * NC && KP <= 2;
    const double* __restrict__ Uin = ubuf(I, stage_bufs(I.rot, stage).in);
    const size_t uplane = (size_t)I.nx * I.ny;
    double pre[KP][NCONS];
    uint32_t bits = 0;
    // The solve's GP weights for radii 1..R and its quadrature weights, for
    // LDS: a lane's weights depend on its face's radius and Gauss point, so
    // read from the constant tables they would be a dependent global load
    // per side inside the rebuild; loaded here they arrive with the
    // classification's values.
    constexpr int NPV = gp_pv_off(R + 1), NQW = (R + 1) * 4, KW = (NPV + NQW + NT - 1) / NT;
    const int set = I.gpset;
    double wv[KW];
#pragma unroll
    for (int k = 0; k < KW; k++) {
        const int e = threadIdx.x + k * NT;
        wv[k] = e < NPV ? c_pv[set][e] : (e < NPV + NQW ? (&c_qw[set][0][0])[e - NPV] : 0.0);
    }
    for (int f = threadIdx.x; f < 2 * NC; f += NT) {
        const int d = f / NC, c = f - d * NC;
        const int i = i0 + c % TX, j = j0 + c / TX;
        const bool dom = i >= s - 2 && i <= ex + 3 && j >= s - 2 && j <= ey + 3;
        if constexpr (kPre) {
#pragma unroll
            for (int p = 0; p < KP; p++) {
                const int u0 = f + p * NT, u = u0 < LP ? u0 : LP - 1, uy = u / LX, ux = u - uy * LX;
                const int kxx = map_axis(i0 - H + ux, s, I.nx, I.bc), kyy = map_axis(j0 - H + uy, s, I.ny, I.bc);
                const double* src = Uin + (size_t)kyy * I.nx + kxx;
#pragma unroll
                for (int v = 0; v < NCONS; v++) pre[p][v] = src[v * uplane];
            }
        }
        // every value the classification reads, loaded at once from clamped
        // indices (a cell off the range reads a valid cell and is dropped)
        const int ci = i < 1 ? 1 : (i >= mx ? mx - 1 : i), cj = j < 1 ? 1 : (j >= I.my ? I.my - 1 : j);
        const size_t idx = (size_t)cj * mx + ci, nb = d == 0 ? idx - 1 : idx - mx;
        const int fL = d == 0 ? FACE_E : FACE_N, fR = d == 0 ? FACE_W : FACE_S;
        const uint8_t flagR = flag[idx], flagL = flag[nb], rcR = I.rc[idx], rcL = I.rc[nb];
        const uint8_t sc = I.scheme[idx], sn = I.scheme[nb];
        const uint8_t rrL = I.faceRR[fL * fplane + nb], rrR = I.faceRR[fR * fplane + idx];
        const uint8_t rrF = I.fluxRR[d * fplane + idx];
        const uint8_t badL = I.inadm[fL * fplane + nb], badR = I.inadm[fR * fplane + idx];
        const double raw = I.FR[(d * NCONS + (d == 0 ? MAGY : MAGX)) * fplane + idx];
        use_all((int)flagR, (int)flagL, (int)rcR, (int)rcL, (int)sc, (int)sn, (int)rrL, (int)rrR, (int)rrF, (int)badL,
                (int)badR, raw);
        if (!dom) continue;
        // the face's cells within the reconstruction range [s-2, e+2]^2
        const bool cellR = i <= ex + 2 && j <= ey + 2;
        const bool cellL = d == 0 ? (i >= s - 1 && j <= ey + 2) : (j >= s - 1 && i <= ex + 2);
        const bool recR = cellR && flagR, recL = cellL && flagL;
        const bool fl = i >= s - 1 && i <= ex + 2 && j >= s - 1 && j <= ey + 2 && (rcR || rcL);
        // In an inner-loop round a face with no side to re-reconstruct has
        // nothing to do: no update has run since the face kernel before,
        // which left its flux and FLB current, and its radius can only have
        // changed with a side's cell lowered, which flags that side.
        if (inner ? !(recL || recR) : !(recL || recR || fl)) continue;
        const int rr_now = sc < sn ? sc : sn;
        const bool keepL = !recL || rrL == rr_now;
        const bool keepR = !recR || rrR == rr_now;
        const bool keepF = !fl || rrF == rr_now;
        if (keepL && keepR && keepF) {
            if (fl) I.FLB[d * fplane + idx] = raw;
            if (recL && badL) bits |= bad_bits(d == 0 ? c % TX - 1 : c % TX, ly0 + (d == 0 ? c / TX : c / TX - 1));
            if (recR && badR) bits |= bad_bits(c % TX, ly0 + c / TX);
        } else {
            const int b = d * K + (BYR ? rr_now : 0), slot = atomicAdd(&lcnt[b], 1);
            lst[b * NC + slot] = (uint32_t)c | (recL ? 0x100u : 0u) | (recR ? 0x200u : 0u) | (fl ? 0x400u : 0u) |
                                 (uint32_t)rr_now << 12 | (uint32_t)rrL << 16 | (uint32_t)rrR << 20;
        }
    }
#pragma unroll
    for (int k = 0; k < KW; k++) {
        const int e = threadIdx.x + k * NT;
        if (e < NPV)
            sPV[e] = wv[k];
        else if (e < NPV + NQW)
            sQW[e - NPV] = wv[k];
    }
    __syncthreads();
    // Two layouts of the faces on waves.  By radius: a face of radius rr
    // takes rr + 1 lanes, 64 / (rr + 1) faces to a wave, every bucket in
    // waves of its own (wb[b]: the first wave of bucket b).  By direction
    // only: every face takes R + 1 lanes (wd[d]: the first wave of
    // direction d).  The first saves lanes where faces were lowered; it is
    // taken unless its extra part-filled waves put another wave on one of
    // the CU's four SIMDs (a block's waves go to the SIMDs in turn).
    // By direction, the multi-solve kernel gives a face R + 1 lanes
    // (face_group), the single-solve kernel 2 (R + 1) (face_group_pair).
    constexpr int PG = BYR ? G : 2 * G, FPWG = 64 / PG, WPR = 4;
    int wb[NB + 1], wd[3];
    wb[0] = 0;
    wd[0] = 0;
    wd[1] = (lcnt[0] + FPWG - 1) / FPWG;
    wd[2] = wd[1] + (lcnt[K] + FPWG - 1) / FPWG;
    bool by_radius = false;
    if constexpr (BYR) {
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int fpw = 64 / (b % K + 1);
            wb[b + 1] = wb[b] + (lcnt[b] + fpw - 1) / fpw;
        }
        wd[1] = 0;
#pragma unroll
        for (int r = 0; r < G; r++) wd[1] += lcnt[r];
        wd[2] = 0;
#pragma unroll
        for (int r = 0; r < G; r++) wd[2] += lcnt[G + r];
        wd[1] = (wd[1] + FPWG - 1) / FPWG;
        wd[2] = wd[1] + (wd[2] + FPWG - 1) / FPWG;
        by_radius = (wb[NB] + WPR - 1) / WPR <= (wd[2] + WPR - 1) / WPR;
    }
    const int nwaves = by_radius ? wb[NB] : wd[2];
    if (nwaves > 0) {
        // stage the U tile [i0-H, i0+TX-1+R] x [j0-H, j0+TY-1+R] in LDS
        if constexpr (kPre) {
#pragma unroll
            for (int p = 0; p < KP; p++) {
                const int u = (int)threadIdx.x + p * NT;
                if (u < LP) {
#pragma unroll
                    for (int v = 0; v < NCONS; v++) sU[v * LP + u] = pre[p][v];
                }
            }
        } else {
            const int oi = i0 - H, oj = j0 - H;
            for (int c = threadIdx.x; c < LP; c += NT) {
                const int cy = c / LX, cx = c - cy * LX;
                const int kxx = map_axis(oi + cx, s, I.nx, I.bc), kyy = map_axis(oj + cy, s, I.ny, I.bc);
                const double* src = Uin + (size_t)kyy * I.nx + kxx;
#pragma unroll
                for (int v = 0; v < NCONS; v++) sU[v * LP + c] = src[v * uplane];
            }
        }
    }
    __syncthreads();
    const Phys ph = phys_of(I);
    // A group of n lanes per face, lane q for Gauss point q (n = rr + 1 by
    // radius, R + 1 by direction): every wave does faces of one direction
    // (and, by radius, of one radius), and a sparse pass costs one point's
    // reconstruction and HLLC per lane
    for (int slot = threadIdx.x; slot < 64 * nwaves; slot += NT) {
        const int w = slot >> 6, lane = slot & 63;
        int b, n, f, g;
        if (by_radius) {
            b = 0;
            while (w >= wb[b + 1]) b++;  // the same for the whole wave
            n = b % K + 1;
            g = lane / n;
            f = (w - wb[b]) * (64 / n) + g;
            if (g >= 64 / n || f >= lcnt[b]) continue;  // the whole group skips
        } else {
            const int d = w >= wd[1] ? 1 : 0;
            n = PG;
            g = lane / PG;
            f = (w - wd[d]) * FPWG + g;
            if (g >= FPWG) continue;
            b = d * K;
            if constexpr (BYR)
                while (b < d * K + K - 1 && f >= lcnt[b]) f -= lcnt[b++];  // the radius bucket of face f
            if (f >= lcnt[b]) continue;  // past the direction's last face: the whole group skips
        }
        const int r = lane - g * n;
        const bool isx = b < K;
        const uint32_t e = lst[b * NC + f];
        const int c = e & 0xff, lx = c % TX, ly = c / TX;
        const bool recL = e & 0x100, recR = e & 0x200, fl = e & 0x400;
        const int rr_now = (e >> 12) & 0xf, rrL = (e >> 16) & 0xf, rrR = (e >> 20) & 0xf;
        const size_t idx = (size_t)(j0 + ly) * mx + (i0 + lx);
        const int cc = (ly + H) * LX + (lx + H);
        const LdsAcc<LX, LP> aR{sU + cc};
        if constexpr (BYR) {
            if (isx)
                face_group<R, 0>(I, ph, sPV, sQW, LdsAcc<LX, LP>{sU + cc - 1}, aR, idx, idx - 1, recL, recR, fl, rr_now,
                                rrL, rrR, r, lane - r, n, lx, ly0 + ly, bits);
            else
                face_group<R, 1>(I, ph, sPV, sQW, LdsAcc<LX, LP>{sU + cc - LX}, aR, idx, idx - mx, recL, recR, fl, rr_now,
                                rrL, rrR, r, lane - r, n, lx, ly0 + ly, bits);
        } else {
            // lane r of the group: point r / 2, side r % 2 (1: R, the cell itself)
            const int q = r >> 1, sd = r & 1;
            if (isx)
                face_group_pair<R, 0>(I, ph, sPV, sQW, LdsAcc<LX, LP>{sU + cc - (sd ? 0 : 1)}, idx, idx - 1, recL, recR,
                                      fl, rr_now, rrL, rrR, q, sd, lane - r, lx, ly0 + ly, bits);
            else
                face_group_pair<R, 1>(I, ph, sPV, sQW, LdsAcc<LX, LP>{sU + cc - (sd ? 0 : LX)}, idx, idx - mx, recL,
                                      recR, fl, rr_now, rrL, rrR, q, sd, lane - r, lx, ly0 + ly, bits);
        }
    }
    report_bad(I, tbad, bits);
}

template <int R, bool ONE>
__global__ __launch_bounds__(face_threads(R), R >= 3 ? 1 : 2) void k_face(const Inst* __restrict__ insts,
                                                                                 const Inst one, const TileList tl,
                                                                                 int stage) {
    __shared__ double sU[NCONS * face_lds_cells<R>()];
    __shared__ uint32_t lst[2 * (ONE ? 1 : gp_nq(R)) * kFaceTX * kFaceTYH];
    __shared__ int lcnt[2 * (ONE ? 1 : gp_nq(R))];
    __shared__ double sPV[gp_pv_off(R + 1)], sQW[(R + 1) * 4];
    const Ctl c1 = ONE ? ctl_of(one.ctl) : Ctl{};
    bool later = false;
    // two blocks per tile, one per half (for_tile_parts)
    for_tile_parts<2>(tl, [&](const Item it, int hh) {
        if (later) __syncthreads();  // the previous tile's LDS reads are done
        later = true;
        const Inst& I = inst_at<ONE>(insts, one, it.inst);
        face_tile_iter<R, !ONE>(I, ONE ? c1 : ctl_of(I.ctl), it, hh, stage, sU, lst, lcnt, sPV, sQW);
    }, c1.run | c1.mode);
}

// The first pass's face kernel in one of its tuned forms: TX x TY cells a
// tile, WAVES waves per SIMD the register allocation allows.  Block b takes
// face tile xcd_tile(b, nitems): for a single solve its tiles in row-major
// order, computed from the index; otherwise entry `items[...]` of the list
// of every active solve's face tiles of this shape.
template <int R, bool ONE, int TX, int TY, int WAVES>
__global__ __launch_bounds__(TX * TY, WAVES) void k_face_full(const Inst* __restrict__ insts, const Inst one,
                                                             const Item* __restrict__ items, int stage) {
    __shared__ double sU[NCONS * (TX + 2 * R + 1) * (TY + 2 * R + 1)];
    __shared__ double sAcc[R == 1 ? NCONS * TX * TY : 1];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    Item it;
    if (ONE) {
        const int ntf = (one.mx + TX - 1) / TX;
        it = Item{0, (int16_t)(t % ntf), (int16_t)(t / ntf)};
    } else {
        it = items[t];
    }
    const Inst& I = inst_at<ONE>(insts, one, it.inst);
    face_tile_full<R, TX, TY>(I, it.tx * TX, it.ty * TY, stage, sU, sAcc);
}

// The 13 cells at taxicab distance <= 2 of (i, j), the cross (distance
// <= 1) first.  Gathers read all of them unconditionally from clamped
// indices and mask the ones off the grid, so their loads are in flight
// together instead of one bounds-checked load at a time.
constexpr int kDiamDx[13] = {0, -1, 1, 0, 0, -2, 2, 0, 0, -1, 1, -1, 1};
constexpr int kDiamDy[13] = {0, 0, 0, -1, 1, 0, 0, -2, 2, -1, -1, 1, 1};

struct Nbhd {
    size_t off[13];  // clamped offsets
    bool in[13];     // on the grid
};
__device__ __forceinline__ Nbhd nbhd(int i, int j, int mx, int my) {
    Nbhd n;
#pragma unroll
    for (int k = 0; k < 13; k++) {
        const int ii = i + kDiamDx[k], jj = j + kDiamDy[k];
        n.in[k] = ii >= 0 && ii < mx && jj >= 0 && jj < my;
        const int ci = ii < 0 ? 0 : (ii >= mx ? mx - 1 : ii), cj = jj < 0 ? 0 : (jj >= my ? my - 1 : jj);
        n.off[k] = (size_t)cj * mx + ci;
    }
    return n;
}

// OR of byte plane `a` over the cross and over the diamond of (i, j)
__device__ __forceinline__ void gather_or(const uint8_t* __restrict__ a, const Nbhd& n, uint8_t& cross,
                                          uint8_t& diam) {
    uint8_t v[13];
#pragma unroll
    for (int k = 0; k < 13; k++) v[k] = a[n.off[k]];
    uint8_t c = 0, d = 0;
#pragma unroll
    for (int k = 0; k < 13; k++) {
        const uint8_t x = n.in[k] ? v[k] : 0;
        if (k < 5) c |= x;
        d |= x;
    }
    cross = c;
    diam = c | d;
}

// a cell is inadmissible when any of its four face sides is
__device__ __forceinline__ uint8_t inadm_cell(const Inst& I, size_t idx, size_t fplane) {
    return I.inadm[idx] | I.inadm[fplane + idx] | I.inadm[2 * fplane + idx] | I.inadm[3 * fplane + idx];
}


// The inner-loop lowering of one cell (src/reconstruct.f90:146-171 with
// src/gridBlock.f90:194-235) as a gather: an inadmissible cell lowers its
// scheme; rc |= cross, rcF |= diamond, rec = cross of the inadmissible
// cells.
__device__ __forceinline__ void lower_cell(const Inst& I, int i, int j) {
    const size_t idx = (size_t)j * I.mx + i, fp = (size_t)I.mx * I.my;
    // every load before any store (the byte planes may alias as far as the
    // compiler knows)
    const Nbhd nb = nbhd(i, j, I.mx, I.my);
    const uint8_t bad = inadm_cell(I, idx, fp), sc = I.scheme[idx], rc = I.rc[idx], rcf = I.rcF[idx];
    uint8_t cross = 0, diam = 0;
#pragma unroll
    for (int f = 0; f < 4; f++) {
        uint8_t c, d;
        gather_or(I.inadm + f * fp, nb, c, d);
        cross |= c;
        diam |= d;
    }
    if (bad) I.scheme[idx] = (uint8_t)(sc > 0 ? sc - 1 : 0);
    I.rc[idx] = rc | cross;
    I.rcF[idx] = rcf | diam;
    I.rec[idx] = cross;
}

// Band cells of a periodic solve: the guard columns and rows and their
// interior images N cells away, [0, 2 NGC) and [N, N + 2 NGC) along each
// axis (disjoint for N >= 2 NGC).  The periodic MOOD sync owns them.
__device__ __forceinline__ bool band_cell(const Inst& I, int i, int j) {
    const int s2 = 2 * I.ngc;
    return i < s2 || i >= I.nx || j < s2 || j >= I.ny;
}

// The window of `low` within four cells of a tile, [x0-4, x0+kCellT+4) x
// [y0-4, y0+kCellT+4), row pitch kWinW: byte lane + 64 r of it in v[r], 0
// off the grid (clamped, masked loads, all issued at once).
constexpr int kWinW = kCellT + 8, kWinN = (kWinW * kWinW + 63) / 64;
__device__ __forceinline__ void low_window(const Inst& I, int x0, int y0, int lane, uint8_t (&v)[kWinN]) {
#pragma unroll
    for (int r = 0; r < kWinN; r++) {
        const int c = lane + r * 64, i = x0 - 4 + c % kWinW, j = y0 - 4 + c / kWinW;
        const bool in = c < kWinW * kWinW && i >= 0 && j >= 0 && i < I.mx && j < I.my;
        const int ci = i < 0 ? 0 : (i >= I.mx ? I.mx - 1 : i), cj = j < 0 ? 0 : (j >= I.my ? I.my - 1 : j);
        v[r] = in ? I.low[(size_t)cj * I.mx + ci] : 0;
    }
}

// The post-detect lowering (src/detect.f90:60-72,219-234) of the cells
// lane, lane + 64, lane + 128, lane + 192 of one tile, as a gather: rc =
// cross(low), rcF = diamond(low), the scheme lowered where low.  load()
// reads the cells' scheme; store() gathers cross and diamond from the
// tile's window of low (low_window, in LDS) and writes.  Band cells are
// left to the sync when `skip_bands` (a periodic solve whose pass syncs).
struct LowerDetect {
    static constexpr int NC = kCellT * kCellT / 64;
    uint8_t sc[NC];
    __device__ __forceinline__ void load(const Inst& I, const Item it, int lane);
    __device__ __forceinline__ void store(const Inst& I, const Item it, int lane, const uint8_t* win,
                                          bool skip_bands) const;
};
__device__ __forceinline__ void LowerDetect::load(const Inst& I, const Item it, int lane) {
#pragma unroll
    for (int r = 0; r < NC; r++) {
        const int c = lane + 64 * r, i = it.tx * kCellT + c % kCellT, j = it.ty * kCellT + c / kCellT;
        const int ci = i < I.mx ? i : I.mx - 1, cj = j < I.my ? j : I.my - 1;
        sc[r] = I.scheme[(size_t)cj * I.mx + ci];
    }
}
__device__ __forceinline__ void LowerDetect::store(const Inst& I, const Item it, int lane, const uint8_t* win,
                                                   bool skip_bands) const {
#pragma unroll
    for (int r = 0; r < NC; r++) {
        const int c = lane + 64 * r, lx = c % kCellT, ly = c / kCellT;
        const int i = it.tx * kCellT + lx, j = it.ty * kCellT + ly;
        if (!(i < I.mx && j < I.my && !(skip_bands && band_cell(I, i, j)))) continue;
        const uint8_t* w = win + (ly + 4) * kWinW + (lx + 4);
        uint8_t cr = 0, dm = 0;
#pragma unroll
        for (int k = 0; k < 13; k++) {
            const uint8_t x = w[kDiamDy[k] * kWinW + kDiamDx[k]];
            if (k < 5) cr |= x;
            dm |= x;
        }
        const size_t idx = (size_t)j * I.mx + i;
        I.rc[idx] = cr;
        I.rcF[idx] = dm;
        if (w[0]) I.scheme[idx] = (uint8_t)(sc[r] > 0 ? sc[r] - 1 : 0);
    }
}

// Periodic MOOD sync (src/boundaryConditions.f90:205-252), in the launch
// that ends a pass.  The reference sweeps x over every row, then y over
// every column, pairing each guard cell with its interior image N cells
// away (OR of rc and rcF, MIN of scheme).  The pairs fall into independent
// units: a guard column's pair in a row that is neither a guard row nor a
// guard row's image (x only), a guard row's pair in such a column (y only),
// and the quad of a guard column and a guard row with their images, whose
// two x pairs come before its two y pairs as in the sweeps.  Needs N >= 2
// NGC, so the guard and image bands are disjoint.  sync_gx blocks per solve.
// A unit of K cells ends with rc and rcF the OR over the unit and scheme
// its MIN (for a quad, the x pairs then the y pairs give exactly that).
// The lowering after detect runs in the same launch and leaves band cells
// alone: each member's values are formed here as that lowering would leave
// them (rc = cross(low), rcF = diamond(low), the scheme lowered where low),
// from `low`, which no kernel of the launch writes.
template <int K>
__device__ __forceinline__ void sync_unit(const Inst& I, const int (&ci)[K], const int (&cj)[K]) {
    size_t c[K];
    uint8_t cr[K], dm[K], own[K], sc[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        c[k] = (size_t)cj[k] * I.mx + ci[k];
        const Nbhd nb = nbhd(ci[k], cj[k], I.mx, I.my);
        uint8_t a = 0, d = 0;
#pragma unroll
        for (int q = 0; q < 13; q++) {
            const uint8_t x = nb.in[q] ? I.low[nb.off[q]] : 0;
            if (q == 0) own[k] = x;
            if (q < 5) a |= x;
            d |= x;
        }
        cr[k] = a;
        dm[k] = d;
        sc[k] = I.scheme[c[k]];
    }
    uint8_t orc = 0, orf = 0, m = 0xff;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const uint8_t l = own[k] ? (uint8_t)(sc[k] > 0 ? sc[k] - 1 : 0) : sc[k];
        orc |= cr[k];
        orf |= dm[k];
        m = l < m ? l : m;
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        I.rc[c[k]] = orc;
        I.rcF[c[k]] = orf;
        I.scheme[c[k]] = m;
    }
}

// block bx of gx over the sync units of one solve
__device__ __forceinline__ void sync_block(const Inst& I, int bx, int gx) {
    const int s = I.ngc, g2 = 2 * s, nx = I.nx, ny = I.ny;
    const int nq = g2 * g2, nxo = g2 * (ny - g2), nyo = g2 * (nx - g2);
    // guard column / row index g in [0, 2s) -> its padded index and image
    auto col = [&](int g, int& i, int& ip) {
        i = g < s ? g : nx + g;
        ip = g < s ? i + nx : i - nx;
    };
    auto row = [&](int g, int& j, int& jp) {
        j = g < s ? g : ny + g;
        jp = g < s ? j + ny : j - ny;
    };
    for (int u = bx * blockDim.x + threadIdx.x; u < nq + nxo + nyo; u += gx * blockDim.x) {
        int i, ip, j, jp;
        if (u < nq) {
            col(u % g2, i, ip);
            row(u / g2, j, jp);
            const int xs[4] = {i, ip, i, ip}, ys[4] = {j, j, jp, jp};
            sync_unit<4>(I, xs, ys);
        } else if (u < nq + nxo) {
            const int v = u - nq;
            col(v % g2, i, ip);
            const int r = g2 + v / g2;
            const int xs[2] = {i, ip}, ys[2] = {r, r};
            sync_unit<2>(I, xs, ys);
        } else {
            const int v = u - nq - nxo;
            row(v % g2, j, jp);
            const int cc = g2 + v / g2;
            const int xs[2] = {cc, cc}, ys[2] = {j, jp};
            sync_unit<2>(I, xs, ys);
        }
    }
}

// ------------------------------------------------------------------ //
// FLICD + FV update + detect + RK combination, cells [s, e+1]^2 of the
// padded-grid tiles.
// stage 0: out = Unew; stage 1: out = (3 base + Unew)/4;
// stage 2: out = (base + 2 Unew)/3.


__device__ __forceinline__ void load_around(const Inst& I, const double* __restrict__ U, int i, int j, Around& a) {
    const size_t plane = (size_t)I.nx * I.ny;
    auto off = [&](int di, int dj) {
        return (size_t)map_axis(j + dj, I.ngc, I.ny, I.bc) * I.nx + map_axis(i + di, I.ngc, I.nx, I.bc);
    };
    const size_t oe = off(1, 0), ow = off(-1, 0), on = off(0, 1), os = off(0, -1);
#pragma unroll
    for (int v = 0; v < NCONS; v++) {
        a.e[v] = U[v * plane + oe];
        a.w[v] = U[v * plane + ow];
        a.n[v] = U[v * plane + on];
        a.s[v] = U[v * plane + os];
    }
    a.d2[0] = U[off(2, 0)];
    a.d2[1] = U[off(-2, 0)];
    a.d2[2] = U[off(0, 2)];
    a.d2[3] = U[off(0, -2)];
    a.d2[4] = U[off(1, 1)];
    a.d2[5] = U[off(1, -1)];
    a.d2[6] = U[off(-1, 1)];
    a.d2[7] = U[off(-1, -1)];
}

// One tile of the FV update.  An own (rcF) cell loads everything its
// update reads -- the FLICD stencil of By on x-faces and Bx on y-faces,
// the fluxes of its four faces, the neighbours' flags and FLB values, U
// around it -- before using any of it, so the global loads overlap.  The
// arithmetic and its order are src/FLICDdivB.f90:38-54 and
// src/timeStep.f90:107-118 as written.
__device__ __forceinline__ void update_cell(const Inst& I, const Ctl& cw, int i, int j, int stage) {
    if (!pass_ready(cw)) return;
    const int s = I.ngc, ex = I.ngc + I.nx - 1, ey = I.ngc + I.ny - 1, mx = I.mx;
    if (i < s || j < s || i > ex + 1 || j > ey + 1) return;
    const size_t idx = (size_t)j * mx + i, fp = (size_t)mx * I.my;
    const bool inner = i <= ex && j <= ey;
    if (!I.rcF[idx]) {
        if (inner) I.low[idx] = 0;
        return;
    }
    // By on x-faces (x(i+a, j+b) = X[idx + a + b mx]) and Bx on y-faces
    const double* __restrict__ X = I.FR + MAGY * fp;
    const double* __restrict__ Y = I.FR + (NCONS + MAGX) * fp;
    const double xw = X[idx - 1], x0 = X[idx], x1 = X[idx + 1], x2 = X[idx + 2];
    const double xs0 = X[idx - mx], xs1 = X[idx - mx + 1], xn0 = X[idx + mx], xn1 = X[idx + mx + 1];
    const double ys = Y[idx - mx], y0 = Y[idx], y1 = Y[idx + mx], y2 = Y[idx + 2 * mx];
    const double yw0 = Y[idx - 1], yw1 = Y[idx - 1 + mx], ye0 = Y[idx + 1], ye1 = Y[idx + 1 + mx];
    const uint8_t rcfE = I.rcF[idx + 1], rcfN = I.rcF[idx + mx];
    const double flbE = I.FLB[idx + 1], flbN = I.FLB[fp + idx + mx];
    const StageBufs sb = stage_bufs(I.rot, stage);
    const double* __restrict__ Uin = ubuf(I, sb.in);
    const size_t plane = (size_t)I.nx * I.ny;
    // the cells of the last row / column (not inner) read valid addresses
    // (fluxes of their own faces, cell 0 of U) and drop the values, so the
    // loads need no branch and are in flight together
    const size_t ci = inner ? (size_t)(j - s) * I.nx + (i - s) : 0;
    double fxm[NCONS], fxp[NCONS], fym[NCONS], fyp[NCONS], uc[NCONS], ub[NCONS];
    Around a;
    const int sch = I.scheme[idx];
#pragma unroll
    for (int v = 0; v < NCONS; v++) {
        if (v == MAGX || v == MAGY) continue;
        fxm[v] = I.FR[v * fp + idx];
        fxp[v] = I.FR[v * fp + idx + 1];
        fym[v] = I.FR[(NCONS + v) * fp + idx];
        fyp[v] = I.FR[(NCONS + v) * fp + idx + mx];
    }
    const double* __restrict__ Ub = ubuf(I, sb.base);  // read at stage 0 too, unused there
#pragma unroll
    for (int v = 0; v < NCONS; v++) {
        uc[v] = Uin[v * plane + ci];
        ub[v] = Ub[v * plane + ci];
    }
    load_around(I, Uin, i, j, a);
    // src/FLICDdivB.f90:38-45 and :47-54 at (i, j)
    const double cx = flicd_corr(x0, x1, xw, y1, y0, yw0, yw1);
    const double cy = flicd_corr(y0, y1, ys, x1, x0, xs0, xs1);
    I.FLB[idx] = cx;
    I.FLB[fp + idx] = cy;
    if (!inner) return;
    // the east / north neighbour's correction, or the value its face holds
    const double cxe = rcfE ? flicd_corr(x1, x2, x0, ye1, ye0, y0, y1) : flbE;
    const double cyn = rcfN ? flicd_corr(y1, y2, y0, xn1, xn0, x0, x1) : flbN;
    fxm[MAGX] = 0.0;
    fxp[MAGX] = 0.0;
    fym[MAGX] = cy;
    fyp[MAGX] = cyn;
    fxm[MAGY] = cx;
    fxp[MAGY] = cxe;
    fym[MAGY] = 0.0;
    fyp[MAGY] = 0.0;
    const double dtdx = I.dtdx, dtdy = I.dtdy;
    double un[NCONS];
#pragma unroll
    for (int v = 0; v < NCONS; v++) un[v] = fv_update(uc[v], dtdx, fxp[v], fxm[v], dtdy, fyp[v], fym[v]);
    bool lower = false;
    if (sch > 0) {
        const Phys ph = phys_of(I);
        lower = detect_cell(I, ph, a, un, uc);
    }
    I.low[idx] = lower ? 1 : 0;
    if (lower) I.ctl[CTL_ANY_LOWER] = 1;
    double* __restrict__ Uout = ubuf(I, sb.out);
#pragma unroll
    for (int v = 0; v < NCONS; v++) Uout[v * plane + ci] = rk_combine(stage, ub[v], un[v]);
}

// update_cell for row row0 + lane / kCellT, column lane % kCellT of tile it
__device__ __forceinline__ void update_tile(const Inst& I, const Ctl& cw, const Item it, int row0, int stage) {
    update_cell(I, cw, it.tx * kCellT + (int)(threadIdx.x % kCellT), it.ty * kCellT + row0 + (int)(threadIdx.x / kCellT),
                stage);
}

// The stage's first update of cell (i, j) in a ready first pass: every cell
// is in rc and rcF (they are set everywhere until the first update) and its
// scheme is R.  The arithmetic is update_tile's; the loads come in two phases -- the
// FLICD stencil, the fluxes and U for the FV update and the RK combination,
// then U around the cell for detection -- so that the kernel needs fewer
// registers and more waves hide the loads of a full grid.
__device__ __forceinline__ void update_cell_full(const Inst& I, const Ctl& cw, int i, int j, int stage) {
    if (!pass_ready(cw)) return;
    const int s = I.ngc, ex = I.ngc + I.nx - 1, ey = I.ngc + I.ny - 1, mx = I.mx;
    if (i < s || j < s || i > ex + 1 || j > ey + 1) return;
    const size_t idx = (size_t)j * mx + i, fp = (size_t)mx * I.my;
    const bool inner = i <= ex && j <= ey;
    const double* __restrict__ X = I.FR + MAGY * fp;
    const double* __restrict__ Y = I.FR + (NCONS + MAGX) * fp;
    const double xw = X[idx - 1], x0 = X[idx], x1 = X[idx + 1], x2 = X[idx + 2];
    const double xs0 = X[idx - mx], xs1 = X[idx - mx + 1], xn0 = X[idx + mx], xn1 = X[idx + mx + 1];
    const double ys = Y[idx - mx], y0 = Y[idx], y1 = Y[idx + mx], y2 = Y[idx + 2 * mx];
    const double yw0 = Y[idx - 1], yw1 = Y[idx - 1 + mx], ye0 = Y[idx + 1], ye1 = Y[idx + 1 + mx];
    const StageBufs sb = stage_bufs(I.rot, stage);
    const double* __restrict__ Uin = ubuf(I, sb.in);
    const size_t plane = (size_t)I.nx * I.ny;
    const size_t ci = inner ? (size_t)(j - s) * I.nx + (i - s) : 0;
    double fxm[NCONS], fxp[NCONS], fym[NCONS], fyp[NCONS], uc[NCONS], ub[NCONS];
#pragma unroll
    for (int v = 0; v < NCONS; v++) {
        if (v == MAGX || v == MAGY) continue;
        fxm[v] = I.FR[v * fp + idx];
        fxp[v] = I.FR[v * fp + idx + 1];
        fym[v] = I.FR[(NCONS + v) * fp + idx];
        fyp[v] = I.FR[(NCONS + v) * fp + idx + mx];
    }
    // the RK base state: stage 0 does not combine (U1 = Unew), so it reads none
    const double* __restrict__ Ub = ubuf(I, sb.base);
#pragma unroll
    for (int v = 0; v < NCONS; v++) {
        uc[v] = Uin[v * plane + ci];
        ub[v] = stage != 0 ? Ub[v * plane + ci] : 0.0;
    }
    const double cx = flicd_corr(x0, x1, xw, y1, y0, yw0, yw1);
    const double cy = flicd_corr(y0, y1, ys, x1, x0, xs0, xs1);
    I.FLB[idx] = cx;
    I.FLB[fp + idx] = cy;
    if (!inner) return;
    // the east and north neighbours are in rcF: their corrections
    const double cxe = flicd_corr(x1, x2, x0, ye1, ye0, y0, y1);
    const double cyn = flicd_corr(y1, y2, y0, xn1, xn0, x0, x1);
    fxm[MAGX] = 0.0;
    fxp[MAGX] = 0.0;
    fym[MAGX] = cy;
    fyp[MAGX] = cyn;
    fxm[MAGY] = cx;
    fxp[MAGY] = cxe;
    fym[MAGY] = 0.0;
    fyp[MAGY] = 0.0;
    const double dtdx = I.dtdx, dtdy = I.dtdy;
    double un[NCONS];
#pragma unroll
    for (int v = 0; v < NCONS; v++) un[v] = fv_update(uc[v], dtdx, fxp[v], fxm[v], dtdy, fyp[v], fym[v]);
    double* __restrict__ Uout = ubuf(I, sb.out);
#pragma unroll
    for (int v = 0; v < NCONS; v++) Uout[v * plane + ci] = rk_combine(stage, ub[v], un[v]);
    reload_barrier();  // detection's loads after the first phase's
    Around a;
    load_around(I, Uin, i, j, a);
    const Phys ph = phys_of(I);
    const bool lower = I.scheme[idx] > 0 && detect_cell(I, ph, a, un, uc);
    I.low[idx] = lower ? 1 : 0;
    if (lower) I.ctl[CTL_ANY_LOWER] = 1;
}

// kUpdQ one-wave blocks per tile (for_tile_parts)
template <bool ONE>
__global__ __launch_bounds__(64) void k_update(const Inst* __restrict__ insts, const Inst one, const TileList tl,
                                               int stage, uint32_t* __restrict__ marks,
                                               uint32_t* __restrict__ amarks, uint32_t stamp, int ninst) {
    const Ctl c1 = ONE ? ctl_of(one.ctl) : Ctl{};
    // Block 0 records, for the launch that ends the iteration, whether the
    // pass is ready (its update and detect run here), whether it ends with
    // the periodic sync, and whether the next iteration is an inner-loop
    // round whose lowering that launch does: its loop decisions rewrite the
    // words these come from while its other blocks read them.
    if (blockIdx.x == 0) {
        for (int k = threadIdx.x; k < ninst; k += blockDim.x) {
            const Inst& I = inst_at<ONE>(insts, one, k);
            const Ctl c = ONE ? c1 : ctl_of(I.ctl);
            I.ctl[CTL_READY] = pass_ready(c) ? 1 : 0;
            I.ctl[CTL_SYNC] = sync_due(I, c) ? 1 : 0;
            I.ctl[CTL_LOWER] = c.run != 0 && c.any_inadm != 0 ? 1 : 0;
        }
    }
    // a single solve before its first update, in a pass that is not ready
    // either: the scan that ends the iteration lists every tile (it reads
    // no candidate marks), so none are written
    const bool nomarks = ONE && one.ctl[CTL_FIRST] != 0 && !pass_ready(c1);
    if (ONE && stamp == 0 && !pass_ready(c1)) {
        // the first pass of a single solve whose pass is not ready: only
        // the scheme's reset (below), by tile index, since the solve's
        // tiles are its rows of tiles in order; any grid covers them
        const int nt = one.ntx * ((one.my + kCellT - 1) / kCellT);
        const int lx = threadIdx.x % kCellT, ly = threadIdx.x / kCellT;
        for (int v = blockIdx.x; v < nt * kUpdQ; v += gridDim.x) {
            const int t = v / kUpdQ;
            const int i = (t % one.ntx) * kCellT + lx, j = (t / one.ntx) * kCellT + (v % kUpdQ) * kUpdRows + ly;
            if (i < one.mx && j < one.my) one.scheme[(size_t)j * one.mx + i] = (uint8_t)one.R;
        }
        return;
    }
    if constexpr (ONE) {
        // Every tile of a single solve listed (the stage's full update when
        // its first pass was not ready): the tiles' marks, then the update on
        // runs of 64 cells of a row, a block each, so a wave's load of a plane
        // is one 512-byte run
        const int nt = one.ntx * ((one.my + kCellT - 1) / kCellT);
        if (stamp != 0 && *tl.cnt == nt) {
            if (nomarks) return;  // nothing to mark or update (for_tile_parts' idle launch)
            for (int t = blockIdx.x; t < nt; t += gridDim.x) {
                const Item it{0, (int16_t)(t % one.ntx), (int16_t)(t / one.ntx)};
                if (c1.run && !nomarks) mark_candidates(one, it, marks, stamp);
                if (threadIdx.x == 0 && pass_ready(c1)) amarks[one.tile_off + t] = stamp;
            }
            const int nrx = (one.mx + 63) / 64, nseg = nrx * one.my;
            for (int g = blockIdx.x; g < nseg; g += gridDim.x)
                update_cell(one, c1, (g % nrx) * 64 + (int)threadIdx.x, g / nrx, stage);
            return;
        }
    }
    for_tile_parts<kUpdQ>(tl, [&](const Item it, int q) {
        const int row0 = q * kUpdRows;
        const Inst& I = inst_at<ONE>(insts, one, it.inst);
        const Ctl cw = ONE ? c1 : ctl_of(I.ctl);
        if (stamp != 0) {
            // a later iteration's active tiles: this is the kernel that
            // visits all of them, so it marks the scan candidates, and, in
            // a ready pass, the tiles whose cells the lowering after detect
            // replaces (the tile's first quarter does)
            if (row0 == 0) {
                if (cw.run && !nomarks) mark_candidates(I, it, marks, stamp);
                if (threadIdx.x == 0 && pass_ready(cw)) amarks[I.tile_off + it.ty * I.ntx + it.tx] = stamp;
            }
        } else {
            // the first pass, every tile: the stage set-up's scheme = R
            // (src/timeStep.f90:79), read first here
            const int i = it.tx * kCellT + (threadIdx.x % kCellT), j = it.ty * kCellT + row0 + (int)(threadIdx.x / kCellT);
            if (i < I.mx && j < I.my) I.scheme[(size_t)j * I.mx + i] = (uint8_t)I.R;
        }
        update_tile(I, cw, it, row0, stage);
    }, c1.run | c1.any_inadm | c1.mode, stamp != 0 && nomarks);
}

// The stage's first update (iteration 0, every tile of every solve): the
// control words of k_update's block 0, the scheme reset to R, and, in a
// ready pass, update_cell_full.  At most 168 VGPRs (159 as built): three waves per SIMD.
// a single solve's full update: regions of kFullRX x kFullRY cells a block
constexpr int kFullRX = 64, kFullRY = 4;
static_assert(kFullRX * kFullRY == 256, "a cell per thread");

template <bool ONE>
__global__ __launch_bounds__(256, 3) void k_update_first(const Inst* __restrict__ insts, const Inst one,
                                                         const TileList tl, int stage, int ninst) {
    const Ctl c1 = ONE ? ctl_of(one.ctl) : Ctl{};
    if (blockIdx.x == 0) {
        for (int k = threadIdx.x; k < ninst; k += blockDim.x) {
            const Inst& I = inst_at<ONE>(insts, one, k);
            const Ctl c = ONE ? c1 : ctl_of(I.ctl);
            I.ctl[CTL_READY] = pass_ready(c) ? 1 : 0;
            I.ctl[CTL_SYNC] = sync_due(I, c) ? 1 : 0;
            I.ctl[CTL_LOWER] = c.run != 0 && c.any_inadm != 0 ? 1 : 0;
        }
    }
    if (ONE && !pass_ready(c1)) {
        // a single solve whose pass is not ready: only the scheme's reset,
        // by tile index (its tiles are its rows of tiles in order)
        const int nt = one.ntx * ((one.my + kCellT - 1) / kCellT);
        const int lx = threadIdx.x % kCellT, ly = threadIdx.x / kCellT;
        for (int t = blockIdx.x; t < nt; t += gridDim.x) {
            const int i = (t % one.ntx) * kCellT + lx, j = (t / one.ntx) * kCellT + ly;
            if (i < one.mx && j < one.my) one.scheme[(size_t)j * one.mx + i] = (uint8_t)one.R;
        }
        return;
    }
    if constexpr (ONE) {
        // a single solve: kFullRX x kFullRY-cell regions in row-major order
        // (xcd_tile's bands when the grid has one block per region), so a
        // wave's load of a plane is one 512-byte run of a row
        const int nrx = (one.mx + kFullRX - 1) / kFullRX, nreg = nrx * ((one.my + kFullRY - 1) / kFullRY);
        const int lx = threadIdx.x % kFullRX, ly = threadIdx.x / kFullRX;
        const bool one_each = (int)gridDim.x == nreg;
        for (int b = blockIdx.x; b < nreg; b += gridDim.x) {
            const int r = one_each ? xcd_tile(b, nreg) : b;
            const int i = (r % nrx) * kFullRX + lx, j = (r / nrx) * kFullRY + ly;
            // the stage set-up's scheme = R (src/timeStep.f90:79), read first here
            if (i < one.mx && j < one.my) one.scheme[(size_t)j * one.mx + i] = (uint8_t)one.R;
            update_cell_full(one, c1, i, j, stage);
        }
        return;
    }
    for_tiles(tl, [&](const Item it) {
        const Inst& I = inst_at<ONE>(insts, one, it.inst);
        const Ctl cw = ONE ? c1 : ctl_of(I.ctl);
        // the stage set-up's scheme = R (src/timeStep.f90:79), read first here
        const int i = it.tx * kCellT + (threadIdx.x % kCellT), j = it.ty * kCellT + (threadIdx.x / kCellT);
        if (i < I.mx && j < I.my) I.scheme[(size_t)j * I.mx + i] = (uint8_t)I.R;
        update_cell_full(I, cw, i, j, stage);
    }, c1.run | c1.any_inadm | c1.mode, false);
}

// End of a MOOD iteration, one block.  Each instance still in its loop
// decides from the flags of the iteration just run whether it continues
// the inner admissibility loop (MODE_INNER), starts a new MOOD pass
// (MODE_NEWPASS) or leaves the loop (src/timeStep.f90:86-128 with
// src/reconstruct.f90:146-171); the active list the next iteration's scan
// writes is emptied; then every control word and both list counts are
// copied to host memory, where the host tests RUN.
__device__ __forceinline__ void loop_decisions(int32_t* __restrict__ ctl, int ninst, int stage) {
    for (int k = threadIdx.x; k < ninst; k += blockDim.x) {
        int32_t* c = ctl + CTL_N * k;
        if (!c[CTL_RUN]) continue;
        if (c[CTL_ANY_INADM]) {
            c[CTL_ANY_INADM] = 0;
            c[CTL_MODE] = MODE_INNER;
            c[CTL_INNER + stage] += 1;
        } else if (c[CTL_ANY_LOWER]) {
            c[CTL_ANY_LOWER] = 0;
            c[CTL_MODE] = MODE_NEWPASS;
            c[CTL_PASSES + stage] += 1;
            c[CTL_INNER + stage] += 1;
        } else {
            c[CTL_RUN] = 0;
        }
    }
}

// The host polls a snapshot's sequence word (after the control words and
// the two counts): it is stored last, behind a system-scope fence.
__device__ __forceinline__ void publish_snapshot(int32_t* __restrict__ snap, int ninst, uint32_t seq) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence_system();
        __hip_atomic_store(snap + CTL_N * ninst + 2, (int32_t)seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// The launch that ends a MOOD iteration, k_end_iter:
//
// * the lowering after detect (src/detect.f90:60-72,219-234) on the tiles
//   the iteration's update visited, when the pass was ready;
// * the periodic MOOD sync, computed from `low` (sync_unit);
// * the active tiles of the next iteration: every quantity a pass writes
//   (inner lowering, reconstruction, fluxes, update, detect bookkeeping)
//   lies within two cells of a cell flagged rc, rcF or inadm, so a tile
//   takes part when its footprint grown by two cells holds such a cell;
//   appended in any order, as each tile's work is independent of the
//   others';
// * the next inner-loop round's lowering, when the pass was not ready;
// * the loop decisions and the snapshot for the host.
//
// After a ready pass the lowering replaces rc and rcF by cross(low) and
// diamond(low) on every tile the update visited, and leaves them clear on
// every other (no cell of those has low within two cells: low is only set
// where rcF was, and the update's tiles cover every cell within two of an
// rcF cell).  So every flag after this launch is diamond(low), synced over
// the periodic units, and the scan reads `low` -- final since the update
// ended -- instead of flags the lowering waves are writing: a tile is
// listed when a low cell lies within four cells of it, or of the images of
// its band cells.
//
// kScanTiles tiles per block, one wave each; the block appends its tiles
// with one atomic, so the list count sees one update per block
//
// One more block ends the MOOD iteration (the decisions read only flags
// earlier kernels set, and no other block reads a word they write: the
// words the others read were set by the update's block 0): the decisions,
// the control words copied to `snap`, and, in snap's count slot w and word
// 3, the lengths of the lists this iteration consumed (list 1 - w, complete
// since the previous scan, and its face list), whose count it then clears
// for the next scan; the sequence word `seq` last.  After a first pass,
// which consumed the fixed list of every tile, the count slots hold nothing
// the host uses.

// whether a low cell lies within four cells of the tile at (x0, y0)
__device__ __forceinline__ bool low_near(const Inst& I, int x0, int y0, int lane) {
    uint8_t v[kWinN];
    low_window(I, x0, y0, lane, v);
    uint8_t o = 0;
#pragma unroll
    for (int r = 0; r < kWinN; r++) o |= v[r];
    return o != 0;
}

// Four waves per SIMD in both forms (about 100-120 VGPRs); at five the
// many-solve form spills.
template <bool ONE>
__global__ __launch_bounds__(kEndIterThreads, 1) void k_end_iter(const Inst* __restrict__ insts, const Inst one,
                                                              const Item* __restrict__ items, int npad,
                                                              Item* __restrict__ act, Item* __restrict__ actf,
                                                              int32_t* __restrict__ act_count,
                                                              const uint32_t* __restrict__ marks,
                                                              const uint32_t* __restrict__ amarks, uint32_t stamp,
                                                              int ninst, int stage, int32_t* __restrict__ ctl,
                                                              int wlist, int32_t* __restrict__ snap, uint32_t seq,
                                                              int sync_gx) {
    __shared__ int hit[kScanTiles], hitf[kScanTiles];
    __shared__ int base, basef;
    __shared__ uint8_t swin[kScanTiles][kWinW * kWinW];
    const int nsync = sync_gx * ninst, nfin = snap != nullptr ? 1 : 0;
    if ((int)blockIdx.x < nsync) {
        // the periodic MOOD sync, sync_gx blocks per solve
        const Inst& I = inst_at<ONE>(insts, one, (int)blockIdx.x / sync_gx);
        if (I.ctl[CTL_SYNC]) sync_block(I, (int)blockIdx.x % sync_gx, sync_gx);
        return;
    }
    if ((int)blockIdx.x == nsync && nfin) {
        // the end of the iteration, in a block of its own
        for (int k = threadIdx.x; k < ninst; k += blockDim.x) {
            int32_t* c = ctl + CTL_N * k;
            if (c[CTL_READY]) c[CTL_FIRST] = 0;  // this launch replaced rc and rcF
        }
        loop_decisions(ctl, ninst, stage);
        __syncthreads();
        if (threadIdx.x == 0) {
            int32_t* cnt = ctl + CTL_N * ninst + 2 * (1 - wlist);
            snap[CTL_N * ninst + wlist] = cnt[0];
            snap[CTL_N * ninst + 3] = cnt[1];
            *reinterpret_cast<unsigned long long*>(cnt) = 0;
        }
        for (int t = threadIdx.x; t < CTL_N * ninst; t += blockDim.x) snap[t] = ctl[t];
        publish_snapshot(snap, ninst, seq);
        return;
    }
    const int sb = (int)blockIdx.x - nsync - nfin, nsb = (int)gridDim.x - nsync - nfin;
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    // a single solve's tiles are its rows of tiles in order, so the tile
    // follows from its index
    const int t = xcd_tile(sb, nsb) * kScanTiles + w;
    const int tc = t < npad ? t : (npad > 0 ? npad - 1 : 0);  // the list has room for one entry at least
    const Item it = ONE ? Item{0, (int16_t)(tc % one.ntx), (int16_t)(tc / one.ntx)} : items[tc];
    const Inst& I = inst_at<ONE>(insts, one, it.inst);
    const int tile = I.tile_off + it.ty * I.ntx + it.tx;
    const bool ready = I.ctl[CTL_READY] != 0;
    const bool sync = I.ctl[CTL_SYNC] != 0;
    // Before an inner-loop round (CTL_LOWER, never after a ready pass) this
    // launch lowers for that round: the waves of the face list's tiles raise
    // rc / rcF / rec within two cells of an inadmissible cell, concurrently
    // with the other waves' reads.
    const bool lower = I.ctl[CTL_LOWER] != 0;
    // A single solve before its first update has rc and rcF set everywhere:
    // every tile is active, and takes its own index in the list, without
    // the atomic (with one per block, a full list cost as much as all the
    // rest of the scan)
    const bool idlist = ONE && !ready && I.ctl[CTL_FIRST] != 0;
    // One round of loads for what the wave does: the window of low within
    // four cells of the tile (the scan's test after a ready pass and,
    // through LDS, the gather of the lowering after detect), the tile's
    // scheme, and, in place of the inadmissible sides, the bad_bits of the
    // 3 x 3 tiles around (every inadmissible side of a footprint cell was
    // written by one of them).  Nothing in this launch writes low or tbad.
    const bool dolow = t < npad && ready && (stamp == 0 || amarks[tile] == stamp);
    const bool scanned = t < npad && (idlist || stamp == 0 || marks[tile] == stamp);
    const int X0 = it.tx * kCellT, Y0 = it.ty * kCellT;
    uint8_t win[kWinN];
    if (ready && (dolow || scanned)) low_window(I, X0, Y0, lane, win);
    LowerDetect ld;
    if (dolow) ld.load(I, it, lane);
    const int nty = (I.my + kCellT - 1) / kCellT, ta = it.tx + lane % 3 - 1, tb = it.ty + lane / 3 - 1;
    const bool tin = lane < 9 && ta >= 0 && tb >= 0 && ta < I.ntx && tb < nty;
    const int tw = (tin ? tb : it.ty) * I.ntx + (tin ? ta : it.tx);
    const uint32_t tf = scanned ? I.tbad[tw] | I.tbadh[tw] : 0u;
    if (dolow) {
        // the window to this wave's LDS rows (a wave's LDS accesses complete
        // in order), then the lowering
        uint8_t* sw = swin[w];
#pragma unroll
        for (int r = 0; r < kWinN; r++) {
            const int c = lane + r * 64;
            if (c < kWinW * kWinW) sw[c] = win[r];
        }
        __builtin_amdgcn_wave_barrier();
        ld.store(I, it, lane, sw, sync);
    }
    bool any = false, anyf = false;
    if (scanned) {
        // a neighbour's sides count when they lie within two cells of this
        // tile: the edge (or corner) of the neighbour facing it
        const int dx = lane % 3 - 1, dy = lane / 3 - 1;
        const uint32_t need2 = BAD_ANY | (dx < 0 ? BAD_E2 : 0u) | (dx > 0 ? BAD_W2 : 0u) | (dy < 0 ? BAD_N2 : 0u) |
                               (dy > 0 ? BAD_S2 : 0u);
        const uint32_t need4 = BAD_ANY | (dx < 0 ? BAD_E4 : 0u) | (dx > 0 ? BAD_W4 : 0u) | (dy < 0 ? BAD_N4 : 0u) |
                               (dy > 0 ? BAD_S4 : 0u);
        anyf = tin && (tf & need2) == need2;
        if (ready) {
            uint8_t o = 0;
#pragma unroll
            for (int r = 0; r < kWinN; r++) o |= win[r];
            any = anyf || o != 0;
            // a band cell of the footprint ends the launch with the OR over
            // its sync unit: its images' low cells count too
            if (sync && !__ballot(any)) {
                const int s2 = 2 * I.ngc;
                const int sx[3] = {0, X0 - 2 < s2 ? I.nx : 0, X0 + kCellT + 1 >= I.nx ? -I.nx : 0};
                const int sy[3] = {0, Y0 - 2 < s2 ? I.ny : 0, Y0 + kCellT + 1 >= I.ny ? -I.ny : 0};
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++)
                        if ((a == 0 || sx[a] != 0) && (b == 0 || sy[b] != 0) && (a | b) != 0)
                            any |= low_near(I, X0 + sx[a], Y0 + sy[b], lane);
            }
        } else {
            // rc and rcF of every cell of the footprint, read at once
            // (clamped, masked)
            const int x0 = X0 - 2, y0 = Y0 - 2;
            constexpr int W = kCellT + 4, NR = (W * W + 63) / 64;
            uint8_t f[NR];
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int c = lane + r * 64, i = x0 + c % W, j = y0 + c / W;
                const bool in = c < W * W && i >= 0 && j >= 0 && i < I.mx && j < I.my;
                const int ci = i < 0 ? 0 : (i >= I.mx ? I.mx - 1 : i), cj = j < 0 ? 0 : (j >= I.my ? I.my - 1 : j);
                const size_t idx = (size_t)cj * I.mx + ci;
                f[r] = in ? (uint8_t)(I.rc[idx] | I.rcF[idx]) : 0;
            }
            // The lowering raises flags only within two cells of an
            // inadmissible cell X; the next iteration's update and
            // post-detect lowering reach two cells further.  Whatever this
            // wave reads of flags being raised, a tile with a cell within
            // four cells of X is listed, so every tile the lowered cells
            // make active is.
            const bool any4 = lower && tin && (tf & need4) == need4;
            any = anyf || any4 || idlist;
#pragma unroll
            for (int r = 0; r < NR; r++) any |= f[r] != 0;
        }
    }
    const bool hw = __ballot(any) != 0;
    // The face list.  Before an inner-loop round (CTL_LOWER), the face
    // kernel has work only on faces with a side to re-reconstruct: a cell
    // of the cross of an inadmissible cell, so the tile's footprint holds
    // an inadmissible cell, which the bad_bits show (no kernel of this
    // launch writes those; where they over-report, a tile listed without
    // work changes nothing).  Otherwise it is the active list.
    const bool hf = lower ? __ballot(anyf) != 0 : hw;
    // The lowering of the next (inner-loop) iteration, on the tiles of the
    // face list.  Every cell it changes lies within two cells of an
    // inadmissible cell, so its tile is on that list whatever this scan
    // reads, and every tile within two cells of such a cell is on the
    // active list (any4 above).  Other waves may read flags it has just
    // raised and list a tile the plain order would not; such a tile holds
    // no flag of its own and its kernels change nothing.
    if (hf && lower) {
        const int ci0 = it.tx * kCellT, cj0 = it.ty * kCellT;
        for (int c = lane; c < kCellT * kCellT; c += 64) {
            const int i = ci0 + c % kCellT, j = cj0 + c / kCellT;
            if (i < I.mx && j < I.my) lower_cell(I, i, j);
        }
    }
    if (lane == 0) {
        hit[w] = hw && !idlist ? 1 : 0;
        hitf[w] = hf ? 1 : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // both lists' counts in one 64-bit atomic (each below 2^31)
        unsigned long long n = idlist && sb == 0 ? (unsigned long long)npad : 0ull;
#pragma unroll
        for (int k = 0; k < kScanTiles; k++) n += (unsigned long long)hit[k] | (unsigned long long)hitf[k] << 32;
        const unsigned long long b = n ? atomicAdd(reinterpret_cast<unsigned long long*>(act_count), n) : 0;
        base = (int)(uint32_t)b;
        basef = (int)(uint32_t)(b >> 32);
    }
    __syncthreads();
    if (lane == 0) {
        int r = 0, rf = 0;
        for (int k = 0; k < w; k++) {
            r += hit[k];
            rf += hitf[k];
        }
        if (hw) act[idlist ? t : base + r] = it;
        if (hf) actf[basef + rf] = it;
    }
}

// ------------------------------------------------------------------ //
// CFL: max over the interior of |v| + c_f from V = cons2prim(U), or from the
// uploaded primitives before the first step (src/cfl.f90:28-48)
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        unsigned long long w = __shfl_xor(v, o);
        v = w > v ? w : v;
    }
    return v;
}

template <bool ONE>
__global__ __launch_bounds__(256) void k_cfl(const Inst* __restrict__ insts, const Inst one,
                                             const Item* __restrict__ items, unsigned long long* __restrict__ part) {
    // partials are stored per tile, so each instance's lie at [cell_off, cell_off + cell_cnt)
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const Item it = items[t];
    const Inst& I = inst_at<ONE>(insts, one, it.inst);
    const int ki = it.tx * kCellT + (threadIdx.x % kCellT), kj = it.ty * kCellT + (threadIdx.x / kCellT);
    unsigned long long bx = 0, by = 0;
    if (ki < I.nx && kj < I.ny) {
        const size_t plane = (size_t)I.nx * I.ny, o = (size_t)kj * I.nx + ki;
        double d, vx, vy, p, b1, b2, b3;
        if (I.v0) {
            const double* V = I.V0;
            d = V[o];
            vx = V[plane + o];
            vy = V[2 * plane + o];
            p = V[4 * plane + o];
            b1 = V[5 * plane + o];
            b2 = V[6 * plane + o];
            b3 = V[7 * plane + o];
        } else {
            const Phys ph = phys_of(I);
            double u[NCONS];
#pragma unroll
            for (int v = 0; v < NCONS; v++) u[v] = I.U[I.rot][v * plane + o];
            const Prim V = cons2prim(ph, u);
            d = V.d;
            vx = V.vx;
            vy = V.vy;
            p = V.p;
            b1 = V.bx;
            b2 = V.by;
            b3 = V.bz;
        }
        // negative zero cannot occur (|v| + c >= 0); NaN sorts above +inf
        cfl_speed_bits(I.gamma, I.hydro, d, vx, vy, p, b1, b2, b3, bx, by);
    }
    bx = wave_max_u64(bx);
    by = wave_max_u64(by);
    __shared__ unsigned long long wx[4], wy[4];
    if ((threadIdx.x & 63) == 0) {
        wx[threadIdx.x >> 6] = bx;
        wy[threadIdx.x >> 6] = by;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) {
            bx = wx[w] > bx ? wx[w] : bx;
            by = wy[w] > by ? wy[w] : by;
        }
        part[2 * t] = bx;
        part[2 * t + 1] = by;
    }
}

// second level: one block per instance over its block partials; the two
// maxima of instance k go to out[2k], out[2k + 1] (pinned host memory),
// and the last block to finish stores `seq` at *seqp, after every block's
// maxima (`done` counts the blocks and is left at zero)
__global__ __launch_bounds__(256) void k_cfl_reduce(const Inst* __restrict__ insts,
                                                    const unsigned long long* __restrict__ part,
                                                    unsigned long long* __restrict__ out, uint32_t* __restrict__ done,
                                                    int32_t* __restrict__ seqp, uint32_t seq) {
    const Inst& I = insts[blockIdx.x];
    unsigned long long bx = 0, by = 0;
    // eight partials a thread in flight at once
    constexpr int K = 8;
    const int end = I.cell_off + I.cell_cnt;
    for (int b0 = I.cell_off + threadIdx.x; b0 < end; b0 += K * blockDim.x) {
        ulonglong2 p[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int b = b0 + k * blockDim.x;
            p[k] = b < end ? reinterpret_cast<const ulonglong2*>(part)[b] : ulonglong2{0, 0};
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            bx = p[k].x > bx ? p[k].x : bx;
            by = p[k].y > by ? p[k].y : by;
        }
    }
    bx = wave_max_u64(bx);
    by = wave_max_u64(by);
    __shared__ unsigned long long wx[4], wy[4];
    if ((threadIdx.x & 63) == 0) {
        wx[threadIdx.x >> 6] = bx;
        wy[threadIdx.x >> 6] = by;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) {
            bx = wx[w] > bx ? wx[w] : bx;
            by = wy[w] > by ? wy[w] : by;
        }
        out[2 * blockIdx.x] = bx;
        out[2 * blockIdx.x + 1] = by;
        __threadfence_system();
        if (atomicAdd(done, 1u) == gridDim.x - 1) {
            *done = 0;
            __threadfence_system();
            __hip_atomic_store(seqp, (int32_t)seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// after a step: the active solves' U^{n+1} sits in U[rot+1]; the uploaded
// primitives no longer describe the state
__global__ void k_rotate(Inst* __restrict__ insts, int ninst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ninst && insts[k].active) {
        insts[k].rot = (insts[k].rot + 1) % 3;
        insts[k].v0 = 0;
    }
}

// dt/dx and dt/dy of each instance from a device dt array
__global__ void k_set_dt(Inst* __restrict__ insts, const double* __restrict__ dt, int ninst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ninst) {
        insts[k].dtdx = dt[k] / insts[k].dx;
        insts[k].dtdy = dt[k] / insts[k].dy;
    }
}

// ------------------------------------------------------------------ //
// layout conversion between the reference's padded AoS U(8,mx,my) /
// V(10,mx,my) and the device SoA interior arrays

__global__ void k_unpack(const Inst I, const double* __restrict__ src, int nv, int to_v0, int buf) {
    const size_t plane = (size_t)I.nx * I.ny;
    for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < plane; o += (size_t)gridDim.x * blockDim.x) {
        const int kj = (int)(o / I.nx), ki = (int)(o - (size_t)kj * I.nx);
        const double* s = src + ((size_t)(kj + I.ngc) * I.mx + (ki + I.ngc)) * nv;
        if (to_v0) {
            // d, vx, vy, vz, p, bx, by, bz from V(1:8)
#pragma unroll
            for (int v = 0; v < NCONS; v++) I.V0[v * plane + o] = s[v];
        } else {
#pragma unroll
            for (int v = 0; v < NCONS; v++) I.U[buf][v * plane + o] = s[v];
        }
    }
}

// padded AoS U, V = cons2prim(U) and scheme, guards by the index map
__global__ void k_pack(const Inst I, int buf, double* __restrict__ Uo, double* __restrict__ Vo,
                       int32_t* __restrict__ So) {
    const size_t n = (size_t)I.mx * I.my, plane = (size_t)I.nx * I.ny;
    const Phys ph = phys_of(I);
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
        const int pj = (int)(p / I.mx), pi = (int)(p - (size_t)pj * I.mx);
        const size_t o = (size_t)map_axis(pj, I.ngc, I.ny, I.bc) * I.nx + map_axis(pi, I.ngc, I.nx, I.bc);
        double u[NCONS];
#pragma unroll
        for (int v = 0; v < NCONS; v++) u[v] = I.U[buf][v * plane + o];
        if (Uo) {
#pragma unroll
            for (int v = 0; v < NCONS; v++) Uo[p * NCONS + v] = u[v];
        }
        if (Vo) {
            // src/convert.f90:198-234 including eint and gamma
            double Bp = mag_pressure(ph, u);
            double d = u[DENS], vx = u[MOMX] / d, vy = u[MOMY] / d, vz = u[MOMZ] / d;
            double eint = (u[ENER] - Bp) / d - 0.5 * (vx * vx + vy * vy + vz * vz);
            double pr = eos_pres(ph, d, eint);
            double* V = Vo + p * 10;
            V[0] = d;
            V[1] = vx;
            V[2] = vy;
            V[3] = vz;
            V[4] = pr;
            V[5] = ph.hydro ? 0.0 : u[MAGX];
            V[6] = ph.hydro ? 0.0 : u[MAGY];
            V[7] = ph.hydro ? 0.0 : u[MAGZ];
            V[8] = eint;
            V[9] = ph.gamma;
        }
        if (So) So[p] = I.scheme[p];
    }
}

// ------------------------------------------------------------------ //
// host-side launchers

static inline dim3 g1(int n) { return dim3((unsigned)n); }

void set_gp_tables(int set, const double* pv, const double* qw, hipStream_t st) {
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(c_pv), pv, sizeof(double) * kPvSet, sizeof(double) * kPvSet * set,
                                 hipMemcpyHostToDevice, st);
    (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(c_qw), qw, sizeof(double) * (kMaxR + 1) * 4,
                                 sizeof(double) * (kMaxR + 1) * 4 * set, hipMemcpyHostToDevice, st);
}

void launch_stage_init(const LaunchSet& L, int stage, hipStream_t st) {
    hipLaunchKernelGGL(k_stage_init, g1(L.npad > 0 ? L.npad : 1), dim3(256), 0, st, L.insts, L.pad, L.npad, L.ninst,
                       stage, L.act_count);
}


// the single-solve variant of a kernel when the table holds one solve
#define HMHD_LAUNCH1(K, grid, block, ...)                                                   \
    do {                                                                                    \
        if (L.one)                                                                          \
            hipLaunchKernelGGL(K<true>, grid, block, 0, st, L.insts, L.inst1, __VA_ARGS__); \
        else                                                                                \
            hipLaunchKernelGGL(K<false>, grid, block, 0, st, L.insts, L.inst1, __VA_ARGS__); \
    } while (0)

void launch_end_iter(const LaunchSet& L, int w, uint32_t stamp, int stage, int32_t* snap, uint32_t seq,
                     hipStream_t st) {
    const int nscan = L.npad > 0 ? (L.npad + kScanTiles - 1) / kScanTiles : 1;
    const int sync_gx = L.sync_gx > 0 ? L.sync_gx : 0;
    HMHD_LAUNCH1(k_end_iter, g1(sync_gx * L.ninst + (snap ? 1 : 0) + nscan), dim3(kEndIterThreads), L.pad, L.npad,
                 L.act[w], L.actf[w], L.act_count + 2 * w, L.marks, L.amarks, stamp, L.ninst, stage, L.ctl, w, snap,
                 seq, sync_gx);
}

void launch_face_iter(const LaunchSet& L, TileList tl, int grid, int R, int stage, hipStream_t st) {
    if (grid <= 0) return;
    const dim3 g = g1(2 * grid);  // two blocks per tile, one per half
#define HMHD_FACE(RR, ONE)                                                                                     \
    hipLaunchKernelGGL((k_face<RR, ONE>), g, dim3(face_threads(RR)), 0, st, L.insts, L.inst1, tl, stage);
#define HMHD_FACE_R(ONE)   \
    if (R == 1) {          \
        HMHD_FACE(1, ONE)  \
    } else if (R == 2) {   \
        HMHD_FACE(2, ONE)  \
    } else {               \
        HMHD_FACE(3, ONE)  \
    }
    if (L.one) {
        HMHD_FACE_R(true)
    } else {
        HMHD_FACE_R(false)
    }
#undef HMHD_FACE_R
#undef HMHD_FACE
}

template <int R, bool ONE>
static void launch_ff(const LaunchSet& L, int v, const Item* items, int n, int stage, hipStream_t st) {
    const dim3 g = g1(n);
#define HMHD_FF(TX, TY, W) \
    hipLaunchKernelGGL((k_face_full<R, ONE, TX, TY, W>), g, dim3(TX * TY), 0, st, L.insts, L.inst1, items, stage)
    switch (v) {
        case 0: HMHD_FF(16, 16, 2); return;
        case 1: HMHD_FF(32, 8, 2); return;
        case 2: HMHD_FF(32, 16, 2); return;
        case 3: HMHD_FF(16, 16, 1); return;
        default: HMHD_FF(32, 8, 1); return;
    }
#undef HMHD_FF
}

void launch_face_full(const LaunchSet& L, int R, int v, const Item* items, int n, int stage, hipStream_t st) {
    if (n <= 0) return;
    static_assert(kNumFaceVariants == 5, "launch_ff's cases");
    if (L.one) {
        if (R == 1) launch_ff<1, true>(L, v, items, n, stage, st);
        else if (R == 2) launch_ff<2, true>(L, v, items, n, stage, st);
        else launch_ff<3, true>(L, v, items, n, stage, st);
    } else {
        if (R == 1) launch_ff<1, false>(L, v, items, n, stage, st);
        else if (R == 2) launch_ff<2, false>(L, v, items, n, stage, st);
        else launch_ff<3, false>(L, v, items, n, stage, st);
    }
}

void launch_update(const LaunchSet& L, TileList tl, int grid, int stage, uint32_t stamp, hipStream_t st) {
    // block 0 always runs: it writes the control words the iteration's end reads
    if (stamp == 0) {
        // a single solve's full update strides over its regions (one block
        // each when the grid asks for every tile)
        if (L.one && grid >= L.npad)
            grid = ((L.inst1.mx + kFullRX - 1) / kFullRX) * ((L.inst1.my + kFullRY - 1) / kFullRY);
        HMHD_LAUNCH1(k_update_first, g1(grid > 0 ? grid : 1), dim3(256), tl, stage, L.ninst);
        return;
    }
    HMHD_LAUNCH1(k_update, g1(kUpdQ * (grid > 0 ? grid : 1)), dim3(64), tl, stage, L.marks, L.amarks, stamp, L.ninst);
}
#undef HMHD_LAUNCH1

void launch_cfl(const LaunchSet& L, int ninst, unsigned long long* out, uint32_t* done, int32_t* seqp, uint32_t seq,
                hipStream_t st) {
    if (L.ncell > 0) {
        if (L.one)
            hipLaunchKernelGGL(k_cfl<true>, g1(L.ncell), dim3(256), 0, st, L.insts, L.inst1, L.cell, L.cflpart);
        else
            hipLaunchKernelGGL(k_cfl<false>, g1(L.ncell), dim3(256), 0, st, L.insts, L.inst1, L.cell, L.cflpart);
    }
    hipLaunchKernelGGL(k_cfl_reduce, dim3(ninst), dim3(256), 0, st, L.insts, L.cflpart, out, done, seqp, seq);
}

void launch_set_dt(Inst* insts, const double* dt, int ninst, hipStream_t st) {
    hipLaunchKernelGGL(k_set_dt, dim3((ninst + 63) / 64), dim3(64), 0, st, insts, dt, ninst);
}

void launch_rotate(Inst* insts, int ninst, hipStream_t st) {
    hipLaunchKernelGGL(k_rotate, dim3((ninst + 63) / 64), dim3(64), 0, st, insts, ninst);
}

void launch_unpack(const Inst& I, const double* src, int nv, bool to_v0, int buf, hipStream_t st) {
    const size_t ncell = (size_t)I.nx * I.ny;
    const int blocks = (int)((ncell + 255) / 256 < 4096 ? (ncell + 255) / 256 : 4096);
    hipLaunchKernelGGL(k_unpack, dim3(blocks), dim3(256), 0, st, I, src, nv, to_v0 ? 1 : 0, buf);
}

void launch_pack(const Inst& I, int buf, double* Uo, double* Vo, int32_t* So, hipStream_t st) {
    const size_t npad = (size_t)I.mx * I.my;
    const int blocks = (int)((npad + 255) / 256 < 4096 ? (npad + 255) / 256 : 4096);
    hipLaunchKernelGGL(k_pack, dim3(blocks), dim3(256), 0, st, I, buf, Uo, Vo, So);
}

}  // namespace hmhd
