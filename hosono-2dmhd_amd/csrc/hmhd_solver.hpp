// hmhd_solver.hpp -- host-side engine shared by the single-run C ABI
// (hmhd_*), the batched form (hmhd_batch_*) and the served form
// (hmhd_server_*): a changing set of independent solves advanced together,
// one launch per kernel family covering every active solve's tiles.
//
// Membership changes (add, remove, activate, deactivate) happen between
// steps; the device-side instance table and tile lists are rebuilt lazily
// before the next launch.  Every solve advances one step per engine step
// with its own dt; all share the rotation of the three state buffers.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <deque>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hmhd.h"
#include "hmhd_internal.hpp"

namespace hmhd {

// The first pass's face kernel comes in tuned forms (kFaceVariants).  The
// form for a launch is chosen per (GP radius, size class of the launch):
// measured on the first first pass that meets the pair, then kept for the
// process.  mode: 1 measure (default), 0 always the 16 x 16 two-wave form,
// 2 + v always form v.
struct FaceTuneEntry {
    int R = 0, size_class = 0, variant = 0;
    double ms[8] = {0};  // milliseconds per form, min over three launches (0: not run)
};
void face_tuning_set_mode(int mode);
std::vector<FaceTuneEntry> face_tuning_table();
void face_tuning_clear();
// size class of a launch over `cells` padded cells: below 2^16, 2^19, 2^21, or above
int face_size_class(long long cells);

// process-wide registry of GP table sets held in __constant__ memory;
// gp_acquire returns the set, -1 on a device error, -2 when every set is in use
int gp_acquire(const double* pv, const double* qw, hipStream_t st, std::string& err);
void gp_release(int set);

struct InstHost {
    int64_t id = -1;
    hmhd_config cfg{};
    int ngc = 0, mx = 0, my = 0;
    double pv[kPvSet];
    double qw[(kMaxR + 1) * 4];
    int gpset = -1;
    bool active = true;
    Inst dev{};              // device pointers and constants, mirrored into the device table
    void* block = nullptr;   // the one device allocation behind every array of this solve
    size_t bytes = 0;
    bool v0_valid = false;   // uploaded primitives feed the next CFL
    std::vector<double> upU, upV;
    std::vector<int32_t> stage_schemes;  // 3*mx*my when recording
    int32_t passes[3] = {0, 0, 0}, inner[3] = {0, 0, 0};
};

enum ProfFamily : int {
    PF_FACE_FULL = 0,
    PF_FACE_FLUX,
    PF_UPDATE,
    PF_RECON,  // no launches (the inner-loop lowering runs in the scan launch)
    PF_FLAGS,
    PF_CFL,
    PF_TILES,  // counters, not times: MOOD iterations / active tiles
    PF_N
};

class Engine {
   public:
    Engine() = default;
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    // validate cfg (errors as the reference's stop conditions)
    static int check_config(const hmhd_config& c, std::string& err);
    // device bytes one solve of this configuration occupies
    static size_t footprint(const hmhd_config& c);

    int64_t add(const hmhd_config& c);  // new solve (GP tables built, device memory zeroed); -1 on error
    int remove(int64_t id);
    int set_active(int64_t id, bool on);
    int upload(int64_t id, const double* U, const double* V);
    int download(int64_t id, double* U, double* V, int32_t* scheme);
    int set_gp(int64_t id, const double* pv, const double* qw);
    // dt of every active solve (cfl_computedt), in the order of active_ids()
    int cfl(std::vector<double>& dt_out);
    // one SSP-RK3 step of every active solve with its own dt (same order)
    int advance(const std::vector<double>& dt);
    int sync();
    const InstHost* find(int64_t id) const;
    InstHost* find(int64_t id);
    std::vector<int64_t> active_ids() const;
    size_t bytes_in_use() const { return bytes_used_; }
    size_t size() const { return inst_.size(); }

    std::string err;
    bool gp_full = false;  // the last add() failed only because every GP table set is in use
    bool record_schemes = false;
    uint32_t prof_mask = 0;  // bit f: time kernel family f
    double prof_ms[PF_N] = {0};
    long long prof_n[PF_N] = {0};

   private:
    int hip_check(hipError_t e, const char* what);
    int ensure_stream();
    int commit();
    int read_snapshot(int w, bool visited);
    int enqueue_iteration(int s, int it, int gface, int gupd);
    int run_stage(int s, bool pre);
    void first_pass_faces(int s);
    void prof_begin(int fam);
    void prof_end(int fam);
    void prof_collect(bool wait_all);
    void launch_face_iter_groups(TileList tl, int grid, int stage);
    int face_variant(int R, int stage);
    template <class T>
    int grow(T** p, size_t& cap, size_t need, bool host, unsigned host_flags = hipHostMallocDefault);

    std::vector<InstHost> inst_;  // dense order = device table order
    std::unordered_map<int64_t, size_t> index_;
    int64_t next_id_ = 0;
    bool dirty_ = true;
    size_t bytes_used_ = 0;
    hipStream_t st_ = nullptr;

    // device-side tables rebuilt by commit()
    Inst* d_insts_ = nullptr;
    size_t cap_insts_ = 0;
    int32_t* d_ctl_ = nullptr;  // CTL_N words per solve, then the active-tile count
    size_t cap_ctl_ = 0;
    int32_t* h_ctl_ = nullptr;  // pinned, two snapshot slots
    int32_t* d_hctl_ = nullptr;  // its device address
    size_t cap_hctl_ = 0;
    // snapshot slot layout: CTL_N words per solve, two list counts, the sequence number
    size_t snap_stride() const { return (size_t)CTL_N * inst_.size() + 4; }
    uint32_t next_seq(int w);
    uint32_t snap_want_[2] = {0, 0};
    // CFL maxima, two per solve, then a sequence word: pinned, written by the device
    unsigned long long* h_cfl_ = nullptr;
    unsigned long long* d_hcfl_ = nullptr;  // its device address
    size_t cap_hcfl_ = 0;
    uint32_t cfl_seq_ = 0;
    // the next advance's first-pass face work (stage 0) is already enqueued,
    // by cfl(), for the state and membership as they are
    bool pre_ok_ = false;
    double* d_dt_ = nullptr;
    size_t cap_dt_ = 0;
    double* h_dt_ = nullptr;
    size_t cap_hdt_ = 0;
    Item* d_pad_ = nullptr;
    size_t cap_pad_ = 0;
    // every active solve's first-pass face tiles of each shape, and their
    // number (for a single solve, its own tiles follow from the block index)
    Item* d_fitems_[3] = {nullptr, nullptr, nullptr};
    size_t cap_fitems_[3] = {0, 0, 0};
    int nfitems_[3] = {0, 0, 0};
    long long rcells_[kMaxR + 1] = {0};  // padded cells of the active solves of each radius
    Item* d_cell_ = nullptr;
    size_t cap_cell_ = 0;
    Item* d_act_ = nullptr;  // the two active lists
    size_t cap_act_ = 0;
    uint32_t* d_marks_ = nullptr;  // scan candidates, one word per tile
    size_t cap_marks_ = 0;
    uint32_t* d_tbad_ = nullptr;  // Inst::tbad of every solve
    size_t cap_tbad_ = 0;
    uint32_t stamp_ = 0;           // last candidate stamp handed out
    unsigned long long* d_part_ = nullptr;
    size_t cap_part_ = 0;
    uint8_t* h_scheme_ = nullptr;
    size_t cap_hscheme_ = 0;
    double* d_stage_ = nullptr;
    size_t cap_stage_ = 0;
    int32_t* d_stage_i_ = nullptr;
    size_t cap_stage_i_ = 0;
    LaunchSet L{};
    int act_n_ = 0;    // length of the active list in the latest snapshot
    // A single solve: per RK stage, the MOOD iteration whose pass first ran
    // its update, in the previous step (the grids of iterations 1 and 2)
    int ready_pred_[3] = {0, 0, 0};
    int32_t npad_val_ = 0;  // source of the device copy of L.npad
    bool has_R_[kMaxR + 1] = {false};
    std::vector<int32_t> ctl_;

    struct EvPair {
        int fam;
        hipEvent_t a, b;
    };
    std::vector<hipEvent_t> ev_free_, ev_all_;
    std::deque<EvPair> ev_pending_;  // in stream order
    hipEvent_t ev_a_ = nullptr;
    int ev_fam_ = -1;
    hipEvent_t ev_get();
};

}  // namespace hmhd
