// hmhd_solver.hpp -- host-side engine shared by the single-run C ABI
// (hmhd_*) and the batched / served forms: a set of independent instances
// advanced together, one launch per kernel family covering every instance's
// tiles.
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "../../include/hmhd.h"
#include "hmhd_internal.hpp"

namespace hmhd {

// process-wide registry of GP table sets held in __constant__ memory
int gp_acquire(const double* pv, const double* qw, std::string& err);
void gp_release(int set);

struct InstHost {
    hmhd_config cfg;
    int ngc, mx, my;
    double pv[kPvSet];
    double qw[(kMaxR + 1) * 4];
    int gpset = -1;
    Inst dev{};  // device pointers and constants, mirrored into the device array
    std::vector<void*> allocs;
};

enum ProfFamily : int { PF_FACE_FULL = 0, PF_FACE_FLUX, PF_UPDATE, PF_RECON, PF_FLAGS, PF_CFL, PF_N };

class Engine {
   public:
    Engine() = default;
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    // validate cfg (errors as the reference's stop conditions), build GP tables
    static int check_config(const hmhd_config& c, std::string& err);
    int add_instance(const hmhd_config& c);  // returns index or -1 (see err)
    int finalize();                           // allocate device arrays, item lists

    int upload(int k, const double* U, const double* V);
    int cfl(std::vector<double>& dt_out);     // dt of every instance
    int advance(const double* dt_host);       // one SSP-RK3 step per instance
    int download(int k, double* U, double* V, int32_t* scheme);
    int sync();
    int set_gp(int k, const double* pv, const double* qw);

    std::string err;
    std::vector<InstHost> inst;
    hipStream_t st = nullptr;
    int cur = 0;  // state buffer holding U^n (same rotation for all instances)
    std::vector<int32_t> ctl_host;  // CTL_N words per instance (pinned mirror below)
    bool record_schemes = false;
    std::vector<std::vector<int32_t>> stage_schemes;  // per instance, 3*mx*my
    std::vector<bool> v0_valid;
    std::vector<std::vector<double>> uploadedU, uploadedV;

    // profiling
    bool prof_on = false;
    double prof_ms[PF_N] = {0};
    long long prof_n[PF_N] = {0};

   private:
    int hip_check(hipError_t e, const char* what);
    int read_ctl();
    int run_stage(int s, int in, int out, int base);
    void prof_begin(int fam);
    void prof_end(int fam);
    void prof_collect();
    void launch_face_groups(bool full, int in);
    void launch_recon_groups(int in);

    Inst* d_insts = nullptr;
    int32_t* d_ctl = nullptr;
    int32_t* h_ctl = nullptr;  // pinned
    double* d_dt = nullptr;
    double* h_dt = nullptr;  // pinned
    unsigned long long* d_cfl = nullptr;
    unsigned long long* h_cfl = nullptr;  // pinned
    uint8_t* h_scheme = nullptr;          // pinned staging for stage schemes
    size_t h_scheme_n = 0;
    double* d_stage = nullptr;            // staging for upload/download
    size_t d_stage_n = 0;
    int32_t* d_stage_i = nullptr;
    std::vector<Item*> d_items;
    LaunchSet L{};
    // face / recon items per radius
    const Item* face_R[kMaxR + 1] = {nullptr};
    int nface_R[kMaxR + 1] = {0};
    const Item* pad_R[kMaxR + 1] = {nullptr};
    int npad_R[kMaxR + 1] = {0};
    bool any_periodic = false;
    bool finalized = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    size_t ev_next = 0;
    hipEvent_t ev_a = nullptr;
    int ev_fam = -1;
};

}  // namespace hmhd
