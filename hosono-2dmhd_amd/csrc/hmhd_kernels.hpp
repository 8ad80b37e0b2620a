// hmhd_kernels.hpp -- host-side launchers of the kernels in hmhd_kernels.hip.
#pragma once

#include <hip/hip_runtime_api.h>

#include "gp_tables.hpp"
#include "hmhd_internal.hpp"

namespace hmhd {

void set_gp_tables(int set, const double* pv, const double* qw, hipStream_t st);
// start of stage `stage`: the MOOD loop's control words (the first pass's
// scan writes list 0) and the flags' stage set-up
void launch_stage_init(const LaunchSet& L, int stage, hipStream_t st);
// The launch that ends a MOOD iteration: the lowering after detect (on the
// update's tiles, when the pass was ready), the periodic MOOD sync (blocks
// ahead of the scan), the scan of the next iteration's tiles, the next
// inner-loop round's lowering, the loop decisions.
// `w`: which of the two active lists the scan of this iteration writes
// stamp 0 scans every tile; otherwise only tiles marked with `stamp`.
// The control words are copied to `snap` (a device address of pinned host
// memory) with, in its count slot w (and word 3), the lengths of the lists
// the iteration consumed (list 1 - w and its face list), then its sequence
// word set to `seq`.
void launch_end_iter(const LaunchSet& L, int w, uint32_t stamp, int stage, int32_t* snap, uint32_t seq,
                     hipStream_t st);
// tile kernels: `grid` blocks stride over the list
// the later passes' face kernel, solves of radius R
void launch_face_iter(const LaunchSet& L, TileList tl, int grid, int R, int stage, hipStream_t st);

// The first pass's face kernel in its tuned forms: a tile shape (tx x ty
// cells, one thread each) and a wave target (the waves per SIMD the
// register budget given to the compiler must allow: 256 VGPRs at two, 512
// at one).  Every radius needs 166 VGPRs or fewer, so the target changes
// the compiler's scheduling, not the occupancy, which LDS bounds.
struct FaceVariant {
    int tx, ty, waves;
};
constexpr int kNumFaceVariants = 5;
constexpr FaceVariant kFaceVariants[kNumFaceVariants] = {{16, 16, 2}, {32, 8, 2}, {32, 16, 2}, {16, 16, 1}, {32, 8, 1}};
constexpr int kNumFaceShapes = 3;  // 16 x 16, 32 x 8, 32 x 16
constexpr int face_shape_of(int v) { return v == 0 || v == 3 ? 0 : (v == 1 || v == 4 ? 1 : 2); }
constexpr bool face_variant_valid(int v) { return v >= 0 && v < kNumFaceVariants; }
// form v for the solves of radius R: `n` blocks, one per face tile (items:
// the list of every active solve's face tiles of the form's shape; unused
// for a single solve, whose tiles follow from the block index)
void launch_face_full(const LaunchSet& L, int R, int v, const Item* items, int n, int stage, hipStream_t st);
// stamp != 0 (a later iteration): also marks the iteration's scan
// candidates and the tiles it visits; stamp == 0 (the first pass): also
// resets the scheme to R.  Block 0 sets the control words the iteration's
// end reads.
void launch_update(const LaunchSet& L, TileList tl, int grid, int stage, uint32_t stamp, hipStream_t st);
// CFL maxima of every solve into `out` (pinned host memory, two per solve);
// the sequence word *seqp is set to `seq` once all are there
void launch_cfl(const LaunchSet& L, int ninst, unsigned long long* out, uint32_t* done, int32_t* seqp, uint32_t seq,
                hipStream_t st);
void launch_set_dt(Inst* insts, const double* dt, int ninst, hipStream_t st);
void launch_rotate(Inst* insts, int ninst, hipStream_t st);
void launch_unpack(const Inst& I, const double* src, int nv, bool to_v0, int buf, hipStream_t st);
void launch_pack(const Inst& I, int buf, double* Uo, double* Vo, int32_t* So, hipStream_t st);

}  // namespace hmhd
