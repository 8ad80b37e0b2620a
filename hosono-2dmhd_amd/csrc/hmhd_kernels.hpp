// hmhd_kernels.hpp -- host-side launchers of the kernels in hmhd_kernels.hip.
#pragma once

#include <hip/hip_runtime_api.h>

#include "gp_tables.hpp"
#include "hmhd_internal.hpp"

namespace hmhd {

void set_gp_tables(int set, const double* pv, const double* qw, hipStream_t st);
void launch_stage_init(const LaunchSet& L, hipStream_t st);
void launch_ctl(const LaunchSet& L, int ninst, int op, int stage, hipStream_t st);
void launch_tile_scan(const LaunchSet& L, hipStream_t st);
void launch_face(const LaunchSet& L, const Item* items, int n, int R, bool full, int stage, hipStream_t st);
void launch_lower_inner(const LaunchSet& L, const Item* items, int n, hipStream_t st);
void launch_update(const LaunchSet& L, const Item* items, int n, int stage, hipStream_t st);
void launch_lower_detect(const LaunchSet& L, const Item* items, int n, hipStream_t st);
void launch_sync(const LaunchSet& L, hipStream_t st);
void launch_cfl(const LaunchSet& L, int ninst, hipStream_t st);
void launch_set_dt(Inst* insts, const double* dt, int ninst, hipStream_t st);
void launch_rotate(Inst* insts, int ninst, hipStream_t st);
void launch_unpack(const Inst& I, const double* src, int nv, bool to_v0, int buf, hipStream_t st);
void launch_pack(const Inst& I, int buf, double* Uo, double* Vo, int32_t* So, hipStream_t st);

}  // namespace hmhd
