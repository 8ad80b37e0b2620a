// hmhd_kernels.hpp -- host-side launchers of the kernels in hmhd_kernels.hip.
#pragma once

#include <hip/hip_runtime_api.h>

#include "gp_tables.hpp"
#include "hmhd_internal.hpp"

namespace hmhd {

void set_gp_tables(int set, const double* pv, const double* qw, hipStream_t st);
void launch_stage_init(const LaunchSet& L, hipStream_t st);
void launch_ctl(const LaunchSet& L, int ninst, int op, int stage, hipStream_t st);
void launch_face(const LaunchSet& L, int R, bool full, int in_buf, hipStream_t st);
void launch_recon_check(const LaunchSet& L, int R, int in_buf, hipStream_t st);
void launch_lower_inner(const LaunchSet& L, hipStream_t st);
void launch_update(const LaunchSet& L, int in_buf, int out_buf, int base_buf, int stage, hipStream_t st);
void launch_lower_detect(const LaunchSet& L, bool periodic_any, hipStream_t st);
void launch_cfl(const LaunchSet& L, int ninst, int buf, bool use_v0, hipStream_t st);
void launch_set_dt(Inst* insts, const double* dt, int ninst, hipStream_t st);
void launch_unpack(const Inst* insts, int inst, const double* src, int nv, bool to_v0, int buf, size_t ncell,
                   hipStream_t st);
void launch_pack(const Inst* insts, int inst, int buf, double* Uo, double* Vo, int32_t* So, size_t npad,
                 hipStream_t st);

}  // namespace hmhd
