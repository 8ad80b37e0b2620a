// hmhd_internal.hpp -- data shared by the host side and the kernels.
//
// One solve ("instance") keeps on the device:
//   U[3]    three SoA FP64 state buffers [8][ny][nx], interior cells only.
//           Guard cells are never stored: every read maps a padded index to
//           the interior cell the reference's guard fill would copy
//           (periodic wrap or outflow clamp, src/boundaryConditions.f90).
//   FR      raw face fluxes [2 dirs][8 vars][my][mx] on the padded grid;
//           FR[0](i,j) is the x-face at i-1/2, FR[1](i,j) the y-face at
//           j-1/2 (src/gridBlock.f90:120-131).
//   FLB     the two FLICD-corrected components [2][my][mx]: By on x-faces,
//           Bx on y-faces (src/FLICDdivB.f90:33-58).
//   MOOD state on the padded grid, one byte per cell: scheme, rc
//           (recompute), rcF (recomputeFLICD), low (detect verdict), rec
//           (cells to re-reconstruct in the inner loop), and per face side
//           (N, S, E, W planes): faceRR[4] (radius the side was last
//           reconstructed with), inadm[4] (the side's states failed CAD/PAD).
//   ctl     per-instance flags read back by the host between MOOD passes.
#pragma once

#include <stdint.h>

namespace hmhd {

enum : int { BC_PERIODIC = 0, BC_OUTFLOW = 1 };
enum : int { NCONS = 8 };
// ctl words: flags the kernels raise, the MOOD-loop state k_stage_init and
// the loop decisions (in k_end_iter's launch) keep on the device, and
// per-stage pass / inner-iteration counters
enum : int {
    CTL_ANY_INADM = 0,  // an inadmissible reconstructed state was found
    CTL_ANY_LOWER = 1,  // detect lowered a cell
    CTL_RUN = 2,        // this instance is still inside its MOOD loop
    CTL_MODE = 3,       // MODE_NEWPASS: reconstruct the rc cells; MODE_INNER: the rec cells
    CTL_PASSES = 4,     // [3] MOOD passes per stage
    CTL_INNER = 7,      // [3] inner admissibility iterations per stage
    CTL_SYNC = 10,      // this pass ends with the periodic MOOD sync (set before the loop decisions)
    CTL_LOWER = 11,     // the next iteration is an inner-loop round, lowered in this pass's scan launch
    CTL_FIRST = 12,     // no pass of this stage has run its update yet: rc and rcF are set everywhere
    CTL_READY = 13,     // this iteration's pass ran its update and detect (set before the loop decisions)
    CTL_N = 16
};
enum : int { MODE_NEWPASS = 0, MODE_INNER = 1 };

constexpr int kMaxR = 3;
constexpr int kMaxGpSets = 64;  // 64 x 4.8 KB of constant memory
// prediction-vector block of radius rr inside one set: [4 faces][nq][ns]
constexpr int gp_ns(int rr) { return 1 + 2 * rr * (rr + 1); }
constexpr int gp_nq(int rr) { return rr + 1; }
constexpr int gp_pv_off(int rr) { return rr <= 1 ? 0 : gp_pv_off(rr - 1) + 4 * gp_nq(rr - 1) * gp_ns(rr - 1); }
constexpr int kPvSet = gp_pv_off(kMaxR + 1);  // 596 doubles

struct Inst {
    int32_t nx, ny, ngc, mx, my, R, bc, hydro;
    int32_t gpset, active;
    int32_t cell_off, cell_cnt;  // CFL blocks of this instance
    int32_t rot;                 // U[rot] holds U^n; stage buffers rotate from it
    int32_t v0;                  // the next CFL reads the uploaded primitives V0
    int32_t tile_off, ntx;       // this instance's tiles in the candidate marks; tiles per row
    double dx, dy, mindxdy, maxdxdy;
    double gamma, small_p, small_rho, cfl;
    double dtdx, dtdy;
    double* U[3];
    double* V0;  // uploaded primitives for the first CFL, [8][ny][nx]: d,vx,vy,vz,p,bx,by,bz
    double* FR;
    double* FLB;
    uint8_t* scheme;
    uint8_t* rc;
    uint8_t* rcF;
    uint8_t* inadm;   // [4][my][mx]
    uint8_t* low;
    uint8_t* rec;
    uint8_t* faceRR;  // [4][my][mx]
    uint8_t* fluxRR;  // [2][my][mx] radius of the last flux of the x / y face
    // per 16 x 16 tile ([ty][ntx]): where the inadmissible face sides the
    // face kernels' last visit of the tile left lie (bad_bits); the later
    // passes' face kernel takes a tile in two blocks, one per half (rows
    // 0-7, 8-15), each with its word (tbad, tbadh), and a reader ORs the two
    uint32_t* tbad;
    uint32_t* tbadh;
    int32_t* ctl;
};

// one tile of one instance; the tile grid depends on the kernel family
struct Item {
    int32_t inst;
    int16_t tx, ty;
};

// tile sizes: every per-cell kernel works on the same 16 x 16 tiles of the
// padded grid, so one active-tile list serves a whole MOOD iteration
constexpr int kCellT = 16;
// tiles per block of the launch that ends a MOOD iteration (a wave each),
// and that launch's block size
constexpr int kScanTiles = 4, kEndIterThreads = 64 * kScanTiles;
constexpr int kFaceTX = kCellT, kFaceTY = kCellT;

// A list of tiles whose length is a device count: the active list the
// previous MOOD iteration's scan wrote (its length unknown to the host
// when it enqueues the next iteration), or a fixed list whose length is
// also kept on the device.  swz: the launch has exactly one block per
// entry, and blocks take the entries in xcd_tile order.
struct TileList {
    const Item* items;
    const int32_t* cnt;
    int swz;
};

struct LaunchSet {
    const Inst* insts;
    int32_t* ctl;      // CTL_N words per instance, the list counts (act_count), npad, the CFL block counter
    const Item* pad;   // every tile of the padded grids
    int npad;
    int ninst;
    int sync_gx;       // blocks per instance of the periodic MOOD sync (0: no periodic solve)
    const Item* cell;  // interior tiles for CFL
    int ncell;
    Item* act[2];        // active tiles, written by alternate MOOD iterations
    // the face kernel's tiles of the same iteration: the active tiles, or,
    // before an inner-loop round, those with an inadmissible cell in reach
    Item* actf[2];
    // [4]: list r's counts as one 64-bit word, act[r]'s at 2 r, actf[r]'s at 2 r + 1
    int32_t* act_count;
    const int32_t* npad_dev;  // npad, on the device (the length of `pad` as a TileList)
    uint32_t* marks;     // per tile: stamp of the last iteration that may change its flags
    uint32_t* amarks;    // per tile: stamp of the last iteration whose update visited it
    unsigned long long* cflpart;  // two partial maxima per CFL block
    // One solve in the table: the per-iteration kernels take its record
    // from their arguments (`inst1`, as the device table holds it when the
    // kernel runs) instead of loading it after the tile.
    int one;
    Inst inst1;
};

}  // namespace hmhd
