// hmhd_solver.cpp -- the host engine: device memory of each solve, the
// device-side instance table and tile work lists, and the MOOD-pass loop of
// each SSP-RK3 stage (src/timeStep.f90:45-171) run for every active solve
// at once.
#include "hmhd_solver.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>

#include "gp_tables.hpp"
#include "hmhd_kernels.hpp"

namespace hmhd {

// ------------------------------------------------------------------ //
// GP table registry (__constant__ sets are shared by every engine)
namespace {
struct GpSet {
    int refs = 0;
    double pv[kPvSet];
    double qw[(kMaxR + 1) * 4];
};
std::mutex g_gp_mu;
GpSet g_gp[kMaxGpSets];

constexpr size_t kAlign = 256;
size_t align_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }
}  // namespace

int gp_acquire(const double* pv, const double* qw, hipStream_t st, std::string& err) {
    std::lock_guard<std::mutex> lk(g_gp_mu);
    int free_set = -1;
    for (int s = 0; s < kMaxGpSets; s++) {
        if (g_gp[s].refs > 0 && !std::memcmp(g_gp[s].pv, pv, sizeof g_gp[s].pv) &&
            !std::memcmp(g_gp[s].qw, qw, sizeof g_gp[s].qw)) {
            g_gp[s].refs++;
            return s;
        }
        if (g_gp[s].refs == 0 && free_set < 0) free_set = s;
    }
    if (free_set < 0) {
        err = "more than 64 distinct GP table sets in one process (every set is in use)";
        return -2;
    }
    GpSet& g = g_gp[free_set];
    std::memcpy(g.pv, pv, sizeof g.pv);
    std::memcpy(g.qw, qw, sizeof g.qw);
    // on the engine's own stream: other engines' kernels read other sets meanwhile
    set_gp_tables(free_set, g.pv, g.qw, st);
    if (hipStreamSynchronize(st) != hipSuccess) {
        err = "uploading GP tables to constant memory failed";
        return -1;
    }
    g.refs = 1;
    return free_set;
}

void gp_release(int set) {
    if (set < 0) return;
    std::lock_guard<std::mutex> lk(g_gp_mu);
    if (g_gp[set].refs > 0) g_gp[set].refs--;
}

// ------------------------------------------------------------------ //
// the first-pass face kernel's tuning table (process-wide)
namespace {
std::mutex g_tune_mu;
std::atomic<int> g_tune_mode{1};
std::vector<FaceTuneEntry> g_tune;
}  // namespace

void face_tuning_set_mode(int mode) { g_tune_mode = mode; }

std::vector<FaceTuneEntry> face_tuning_table() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    return g_tune;
}

void face_tuning_clear() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
}

int face_size_class(long long cells) { return cells < (1 << 16) ? 0 : cells < (1 << 19) ? 1 : cells < (1 << 21) ? 2 : 3; }

// The sequence words kernels store in pinned memory for the host to poll
// (CFL results, MOOD snapshots).  They are unique within the process and
// never 0: pinned memory is recycled between engines, and a word an earlier
// engine left there must never equal the one a later engine waits for.
static uint32_t unique_seq() {
    static std::atomic<uint32_t> g_seq{0};
    uint32_t v;
    do {
        v = g_seq.fetch_add(1, std::memory_order_relaxed) + 1;
    } while (v == 0);
    return v;
}

// ------------------------------------------------------------------ //

int Engine::hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return HMHD_OK;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return HMHD_ERR_DEVICE;
}

// Every engine works on a stream of its own that does not synchronize with
// the null stream, and every copy, allocation and release it makes is
// ordered on that stream alone, so engines on one device (the server's
// lanes) overlap.  Solve blocks come from the device's default memory pool
// (stream-ordered), kept reserved once released so that later solves reuse
// it without a device-wide synchronization.
int Engine::ensure_stream() {
    if (st_) return HMHD_OK;
    static std::once_flag pool_once;
    std::call_once(pool_once, [] {
        int dev = 0;
        hipMemPool_t pool;
        if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            uint64_t keep = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
    });
    return hip_check(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking), "hipStreamCreateWithFlags");
}

Engine::~Engine() {
    if (st_) (void)hipStreamSynchronize(st_);
    for (auto& in : inst_) {
        if (in.block) (void)hipFreeAsync(in.block, st_);
        gp_release(in.gpset);
    }
    void* dev[] = {d_insts_, d_ctl_,   d_dt_,    d_pad_,    d_cell_,     d_act_,      d_marks_,
                   d_tbad_,  d_part_,  d_stage_, d_stage_i_, d_fitems_[0], d_fitems_[1], d_fitems_[2]};
    for (void* p : dev)
        if (p) (void)hipFreeAsync(p, st_);
    if (st_) (void)hipStreamSynchronize(st_);
    void* host[] = {h_ctl_, h_cfl_, h_dt_, h_scheme_};
    for (void* p : host)
        if (p) (void)hipHostFree(p);
    for (auto e : ev_all_) (void)hipEventDestroy(e);
    if (st_) (void)hipStreamDestroy(st_);
}

int Engine::check_config(const hmhd_config& c, std::string& err) {
    if (c.R < 1 || c.R > kMaxR) {
        err = "GP_maxRadius must be 1, 2 or 3 (src/GP.f90:150)";
        return HMHD_ERR_ARG;
    }
    const int ngc = c.R + 2;
    if (c.nx < 2 * ngc || c.ny < 2 * ngc) {
        err = "grid_Nx and grid_Ny must be at least 2*(GP_maxRadius+2)";
        return HMHD_ERR_ARG;
    }
    if ((long long)(c.nx + 2 * ngc) * (c.ny + 2 * ngc) > (1ll << 31) ||
        (c.nx + 2 * ngc + kCellT - 1) / kCellT > 32767 || (c.ny + 2 * ngc + kCellT - 1) / kCellT > 32767) {
        err = "grid too large for 32-bit cell indices";
        return HMHD_ERR_ARG;
    }
    if (!(c.xend > c.xbeg) || !(c.yend > c.ybeg)) {
        err = "domain bounds must satisfy xEnd > xBeg and yEnd > yBeg";
        return HMHD_ERR_ARG;
    }
    if (c.bc != HMHD_BC_PERIODIC && c.bc != HMHD_BC_OUTFLOW) {
        err = "Unrecognized choice of boundary condition (src/boundaryConditions.f90:42-49)";
        return HMHD_ERR_ARG;
    }
    if (c.riemann != HMHD_RIEMANN_HLLC) {
        err = "Unrecognized choice of riemann solver (src/simulation.f90:65-71)";
        return HMHD_ERR_ARG;
    }
    if (!(c.gamma > 1.0) || !(c.ell_over_delta > 0.0) || !(c.cfl > 0.0)) {
        err = "sim_gamma must exceed 1; GP_ellOverDelta and sim_cfl must be positive";
        return HMHD_ERR_ARG;
    }
    return HMHD_OK;
}

// layout of one solve's device block: U[3], V0, FR, FLB, then the byte
// planes scheme, rc, rcF, low, rec, faceRR[4], inadm[4], fluxRR[2]
size_t Engine::footprint(const hmhd_config& c) {
    const int ngc = c.R + 2;
    const size_t nin = (size_t)c.nx * c.ny, npad = (size_t)(c.nx + 2 * ngc) * (c.ny + 2 * ngc);
    size_t b = 0;
    b += 4 * align_up(sizeof(double) * NCONS * nin);
    b += align_up(sizeof(double) * 2 * NCONS * npad);
    b += align_up(sizeof(double) * 2 * npad);
    b += 5 * align_up(npad) + 2 * align_up(4 * npad) + align_up(2 * npad);
    return b;
}

int64_t Engine::add(const hmhd_config& c) {
    if (check_config(c, err) != HMHD_OK) return -1;
    if (ensure_stream() != HMHD_OK) return -1;
    InstHost h;
    h.cfg = c;
    h.ngc = c.R + 2;
    h.mx = c.nx + 2 * h.ngc;
    h.my = c.ny + 2 * h.ngc;
    if (gp_compute(c.R, c.ell_over_delta, c.xbeg, c.xend, c.ybeg, c.yend, c.nx, c.ny, h.pv, h.qw) != 0) {
        err = "GP covariance matrix is not positive definite (src/linAlgQuadPrecision.f90:136)";
        return -1;
    }
    h.gpset = gp_acquire(h.pv, h.qw, st_, err);
    gp_full = h.gpset == -2;
    if (h.gpset < 0) return -1;
    h.bytes = footprint(c);
    if (hip_check(hipMallocAsync(&h.block, h.bytes, st_), "hipMallocAsync(solve)") ||
        hip_check(hipMemsetAsync(h.block, 0, h.bytes, st_), "hipMemsetAsync(solve)")) {
        if (h.block) (void)hipFreeAsync(h.block, st_);
        gp_release(h.gpset);
        return -1;
    }
    const size_t nin = (size_t)c.nx * c.ny, npad = (size_t)h.mx * h.my;
    char* p = static_cast<char*>(h.block);
    auto take = [&](size_t bytes) {
        char* q = p;
        p += align_up(bytes);
        return q;
    };
    Inst& d = h.dev;
    for (int b = 0; b < 3; b++) d.U[b] = reinterpret_cast<double*>(take(sizeof(double) * NCONS * nin));
    d.V0 = reinterpret_cast<double*>(take(sizeof(double) * NCONS * nin));
    d.FR = reinterpret_cast<double*>(take(sizeof(double) * 2 * NCONS * npad));
    d.FLB = reinterpret_cast<double*>(take(sizeof(double) * 2 * npad));
    d.scheme = reinterpret_cast<uint8_t*>(take(npad));
    d.rc = reinterpret_cast<uint8_t*>(take(npad));
    d.rcF = reinterpret_cast<uint8_t*>(take(npad));
    d.low = reinterpret_cast<uint8_t*>(take(npad));
    d.rec = reinterpret_cast<uint8_t*>(take(npad));
    d.faceRR = reinterpret_cast<uint8_t*>(take(4 * npad));
    d.inadm = reinterpret_cast<uint8_t*>(take(4 * npad));
    d.fluxRR = reinterpret_cast<uint8_t*>(take(2 * npad));
    d.nx = c.nx;
    d.ny = c.ny;
    d.ngc = h.ngc;
    d.mx = h.mx;
    d.my = h.my;
    d.R = c.R;
    d.bc = c.bc;
    d.hydro = c.force_hydro ? 1 : 0;
    d.gpset = h.gpset;
    d.dx = (c.xend - c.xbeg) / c.nx;
    d.dy = (c.yend - c.ybeg) / c.ny;
    d.mindxdy = d.dy < d.dx ? d.dy : d.dx;
    d.maxdxdy = d.dy > d.dx ? d.dy : d.dx;
    d.gamma = c.gamma;
    d.small_p = c.small_p;
    d.small_rho = c.small_rho;
    d.cfl = c.cfl;
    d.rot = 0;
    d.v0 = 0;
    h.id = next_id_++;
    bytes_used_ += h.bytes;
    index_[h.id] = inst_.size();
    inst_.push_back(std::move(h));
    dirty_ = true;
    pre_ok_ = false;
    return inst_.back().id;
}

const InstHost* Engine::find(int64_t id) const {
    auto it = index_.find(id);
    return it == index_.end() ? nullptr : &inst_[it->second];
}

InstHost* Engine::find(int64_t id) {
    auto it = index_.find(id);
    return it == index_.end() ? nullptr : &inst_[it->second];
}

std::vector<int64_t> Engine::active_ids() const {
    std::vector<int64_t> v;
    for (auto& h : inst_)
        if (h.active) v.push_back(h.id);
    return v;
}

int Engine::remove(int64_t id) {
    auto it = index_.find(id);
    if (it == index_.end()) {
        err = "unknown solve id";
        return HMHD_ERR_ARG;
    }
    int rc = sync();
    if (rc) return rc;
    InstHost& h = inst_[it->second];
    if (h.block) (void)hipFreeAsync(h.block, st_);
    gp_release(h.gpset);
    bytes_used_ -= h.bytes;
    inst_.erase(inst_.begin() + it->second);
    index_.clear();
    for (size_t k = 0; k < inst_.size(); k++) index_[inst_[k].id] = k;
    dirty_ = true;
    pre_ok_ = false;
    return HMHD_OK;
}

int Engine::set_active(int64_t id, bool on) {
    InstHost* h = find(id);
    if (!h) {
        err = "unknown solve id";
        return HMHD_ERR_ARG;
    }
    if (h->active != on) {
        h->active = on;
        dirty_ = true;
        pre_ok_ = false;
    }
    return HMHD_OK;
}

template <class T>
int Engine::grow(T** p, size_t& cap, size_t need, bool host, unsigned host_flags) {
    if (need <= cap && *p) return HMHD_OK;
    size_t n = std::max<size_t>(need, 2 * cap);
    n = std::max<size_t>(n, 16);
    if (*p) {
        if (host)
            (void)hipHostFree(*p);
        else
            (void)hipFreeAsync(*p, st_);
        *p = nullptr;
    }
    int rc = host ? hip_check(hipHostMalloc(reinterpret_cast<void**>(p), sizeof(T) * n, host_flags), "hipHostMalloc")
                  : hip_check(hipMallocAsync(reinterpret_cast<void**>(p), sizeof(T) * n, st_), "hipMallocAsync");
    if (rc) return rc;
    if (host) std::memset(*p, 0, sizeof(T) * n);  // no word of an earlier owner survives
    cap = n;
    return HMHD_OK;
}

static void tiles(std::vector<Item>& v, int k, int w, int h) {
    const int nt_x = (w + kCellT - 1) / kCellT, nt_y = (h + kCellT - 1) / kCellT;
    for (int ty = 0; ty < nt_y; ty++)
        for (int tx = 0; tx < nt_x; tx++) v.push_back(Item{k, (int16_t)tx, (int16_t)ty});
}

// rebuild the device instance table and the tile lists after membership
// or activity changes (between steps only)
int Engine::commit() {
    if (!dirty_) return HMHD_OK;
    int rc;
    if ((rc = ensure_stream())) return rc;
    if ((rc = sync())) return rc;
    const size_t n = inst_.size();
    // CTL_N words per solve, the four list counts, npad, the CFL block counter
    const size_t nctl = (size_t)CTL_N * n + 6;
    if ((rc = grow(&d_insts_, cap_insts_, n, false))) return rc;
    if ((rc = grow(&d_ctl_, cap_ctl_, nctl, false))) return rc;
    // snapshots are polled while the stream runs: coherent (fine-grained) memory
    if ((rc = grow(&h_ctl_, cap_hctl_, 2 * nctl, true, hipHostMallocCoherent | hipHostMallocMapped))) return rc;
    if ((rc = hip_check(hipHostGetDevicePointer(reinterpret_cast<void**>(&d_hctl_), h_ctl_, 0),
                        "hipHostGetDevicePointer")))
        return rc;
    // polled while the stream runs, like the snapshots
    if ((rc = grow(&h_cfl_, cap_hcfl_, 2 * n + 2, true, hipHostMallocCoherent | hipHostMallocMapped))) return rc;
    if ((rc = hip_check(hipHostGetDevicePointer(reinterpret_cast<void**>(&d_hcfl_), h_cfl_, 0),
                        "hipHostGetDevicePointer")))
        return rc;
    if ((rc = grow(&d_dt_, cap_dt_, n, false))) return rc;
    if ((rc = grow(&h_dt_, cap_hdt_, n, true))) return rc;
    if ((rc = hip_check(hipMemsetAsync(d_ctl_, 0, sizeof(int32_t) * nctl, st_), "hipMemsetAsync(ctl)"))) return rc;

    // tiles of the active solves, all 16 x 16 tiles of each padded grid:
    // every tile, the tiles touching a guard band, the interior tiles (CFL)
    std::vector<Item> pad, cell;
    int max_sync_units = 0;
    std::vector<Inst> table(n);
    size_t max_pad = 0, sum_pad = 0, ntiles = 0;
    for (int R = 0; R <= kMaxR; R++) has_R_[R] = false;
    for (size_t k = 0; k < n; k++) {
        InstHost& h = inst_[k];
        Inst d = h.dev;
        d.ctl = d_ctl_ + CTL_N * k;
        d.active = h.active ? 1 : 0;
        d.v0 = h.v0_valid ? 1 : 0;
        d.cell_off = (int32_t)cell.size();
        d.cell_cnt = 0;
        d.ntx = (h.mx + kCellT - 1) / kCellT;
        d.tile_off = (int32_t)ntiles;
        ntiles += (size_t)d.ntx * ((h.my + kCellT - 1) / kCellT);
        const size_t np = (size_t)h.mx * h.my;
        max_pad = std::max(max_pad, np);
        sum_pad += np;
        if (h.active) {
            has_R_[h.cfg.R] = true;
            tiles(pad, (int)k, h.mx, h.my);
            tiles(cell, (int)k, h.cfg.nx, h.cfg.ny);
            d.cell_cnt = (int32_t)cell.size() - d.cell_off;
            if (h.cfg.bc == HMHD_BC_PERIODIC) {
                // units of the periodic MOOD sync: guard quads, x-only and y-only pairs
                const int g2 = 2 * h.ngc;
                const int units = g2 * g2 + g2 * (h.cfg.ny - g2) + g2 * (h.cfg.nx - g2);
                max_sync_units = std::max(max_sync_units, units);
            }
        }
        h.dev = d;
        table[k] = d;
    }
    if ((rc = grow(&d_tbad_, cap_tbad_, 2 * (ntiles + 1), false))) return rc;
    if ((rc = hip_check(hipMemsetAsync(d_tbad_, 0, sizeof(uint32_t) * 2 * (ntiles + 1), st_),
                        "hipMemsetAsync(tbad)")))
        return rc;
    for (size_t k = 0; k < n; k++) {
        inst_[k].dev.tbad = table[k].tbad = d_tbad_ + table[k].tile_off;
        inst_[k].dev.tbadh = table[k].tbadh = d_tbad_ + (ntiles + 1) + table[k].tile_off;
    }
    if (n && (rc = hip_check(hipMemcpyAsync(d_insts_, table.data(), sizeof(Inst) * n, hipMemcpyHostToDevice, st_),
                             "hipMemcpyAsync(table)")))
        return rc;
    auto put = [&](const std::vector<Item>& v, Item** d, size_t& cap) -> int {
        int r = grow(d, cap, v.size() + 1, false);
        if (r || v.empty()) return r;
        return hip_check(hipMemcpyAsync(*d, v.data(), sizeof(Item) * v.size(), hipMemcpyHostToDevice, st_),
                         "hipMemcpyAsync(items)");
    };
    if ((rc = put(pad, &d_pad_, cap_pad_))) return rc;
    // first-pass face tiles of each shape; padded cells per radius
    for (int R = 0; R <= kMaxR; R++) rcells_[R] = 0;
    // the host vectors of the asynchronous copies live until the sync below
    std::vector<Item> fl[kNumFaceShapes];
    for (int k = 0; k < kNumFaceShapes; k++) {
        const int v = k == 0 ? 0 : (k == 1 ? 1 : 2);
        const int tx = kFaceVariants[v].tx, ty = kFaceVariants[v].ty;
        std::vector<Item>& f = fl[k];
        for (size_t q = 0; q < n; q++) {
            const InstHost& h = inst_[q];
            if (!h.active) continue;
            const int ntx = (h.mx + tx - 1) / tx, nty = (h.my + ty - 1) / ty;
            for (int b = 0; b < nty; b++)
                for (int a = 0; a < ntx; a++) f.push_back(Item{(int)q, (int16_t)a, (int16_t)b});
        }
        nfitems_[k] = (int)f.size();
        if ((rc = put(f, &d_fitems_[k], cap_fitems_[k]))) return rc;
    }
    for (size_t q = 0; q < n; q++)
        if (inst_[q].active) rcells_[inst_[q].cfg.R] += (long long)inst_[q].mx * inst_[q].my;
    if ((rc = put(cell, &d_cell_, cap_cell_))) return rc;
    if ((rc = grow(&d_act_, cap_act_, 4 * (pad.size() + 1), false))) return rc;
    // candidate marks, then the update's visit stamps
    if ((rc = grow(&d_marks_, cap_marks_, 2 * (ntiles + 1), false))) return rc;
    if ((rc = hip_check(hipMemsetAsync(d_marks_, 0, sizeof(uint32_t) * 2 * (ntiles + 1), st_),
                        "hipMemsetAsync(marks)")))
        return rc;
    stamp_ = 0;
    if ((rc = grow(&d_part_, cap_part_, 2 * cell.size() + 2, false))) return rc;
    if ((rc = grow(&d_stage_, cap_stage_, 10 * max_pad, false))) return rc;
    if ((rc = grow(&d_stage_i_, cap_stage_i_, max_pad, false))) return rc;
    if ((rc = grow(&h_scheme_, cap_hscheme_, sum_pad, true))) return rc;
    L.insts = d_insts_;
    L.ctl = d_ctl_;
    L.pad = d_pad_;
    L.npad = (int)pad.size();
    L.ninst = (int)n;
    L.sync_gx = std::min((max_sync_units + kEndIterThreads - 1) / kEndIterThreads, 65536 / kEndIterThreads);
    L.cell = d_cell_;
    L.ncell = (int)cell.size();
    L.act[0] = d_act_;
    L.act[1] = d_act_ + pad.size() + 1;
    L.actf[0] = d_act_ + 2 * (pad.size() + 1);
    L.actf[1] = d_act_ + 3 * (pad.size() + 1);
    L.act_count = d_ctl_ + (size_t)CTL_N * n;
    L.npad_dev = d_ctl_ + (size_t)CTL_N * n + 4;
    npad_val_ = (int32_t)pad.size();
    if ((rc = hip_check(hipMemcpyAsync(d_ctl_ + (size_t)CTL_N * n + 4, &npad_val_, sizeof npad_val_,
                                       hipMemcpyHostToDevice, st_),
                        "hipMemcpyAsync(npad)")))
        return rc;
    L.marks = d_marks_;
    L.amarks = d_marks_ + ntiles + 1;
    L.cflpart = d_part_;
    L.one = n == 1 ? 1 : 0;
    L.inst1 = n == 1 ? table[0] : Inst{};
    ctl_.assign((size_t)CTL_N * n, 0);
    dirty_ = false;
    return sync();
}

int Engine::set_gp(int64_t id, const double* pv, const double* qw) {
    InstHost* h = find(id);
    if (!h) {
        err = "unknown solve id";
        return HMHD_ERR_ARG;
    }
    int rc = ensure_stream();
    if (rc) return rc;
    int s = gp_acquire(pv, qw, st_, err);
    if (s < 0) return HMHD_ERR_DEVICE;
    // work already queued (cfl()'s first pass) may still read the old set;
    // once released, another engine may overwrite it from its own stream
    if ((rc = sync())) return rc;
    gp_release(h->gpset);
    h->gpset = s;
    h->dev.gpset = s;
    std::memcpy(h->pv, pv, sizeof h->pv);
    std::memcpy(h->qw, qw, sizeof h->qw);
    dirty_ = true;
    pre_ok_ = false;
    return HMHD_OK;
}

// ------------------------------------------------------------------ //
// profiling: events bracket each launch of a kernel family

hipEvent_t Engine::ev_get() {
    if (ev_free_.empty()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ev_all_.push_back(e);
        return e;
    }
    hipEvent_t e = ev_free_.back();
    ev_free_.pop_back();
    return e;
}

void Engine::prof_begin(int fam) {
    if (!(prof_mask >> fam & 1u)) return;
    ev_a_ = ev_get();
    ev_fam_ = fam;
    if (ev_a_) (void)hipEventRecord(ev_a_, st_);
}

void Engine::prof_end(int fam) {
    if (!(prof_mask >> fam & 1u) || ev_fam_ != fam || !ev_a_) return;
    hipEvent_t b = ev_get();
    if (!b) return;
    (void)hipEventRecord(b, st_);
    ev_pending_.push_back(EvPair{fam, ev_a_, b});
    ev_a_ = nullptr;
    ev_fam_ = -1;
}

// add up the timed launches that have finished (all of them when the
// stream is known to be idle)
void Engine::prof_collect(bool wait_all) {
    while (!ev_pending_.empty()) {
        EvPair& p = ev_pending_.front();
        if (!wait_all && hipEventQuery(p.b) != hipSuccess) break;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            prof_ms[p.fam] += ms;
            prof_n[p.fam] += 1;
        }
        ev_free_.push_back(p.a);
        ev_free_.push_back(p.b);
        ev_pending_.pop_front();
    }
}

// ------------------------------------------------------------------ //

// the sequence number the next snapshot into slot w carries (never 0)
uint32_t Engine::next_seq(int w) {
    snap_want_[w] = unique_seq();
    return snap_want_[w];
}

// wait for snapshot w; its control words become ctl_, and its count slot
// (the list the first pass's scan wrote; for a later iteration, the list
// it consumed) becomes act_n_.  `visited`: the count is the tiles a MOOD
// iteration visited (profiling counter).
int Engine::read_snapshot(int w, bool visited) {
    const size_t n = inst_.size();
    const int32_t* h = h_ctl_ + w * snap_stride();
    // The snapshot's block stores its sequence number last, after a
    // system-scope fence; wait for it in the pinned memory itself (no event
    // in the stream).  An idle stream without it is an error, not a hang.
    const volatile int32_t* seq = h + CTL_N * n + 2;
    const int32_t want = (int32_t)snap_want_[w];
    for (unsigned k = 1; *seq != want; k++) {
        if (k % 1024 == 0) {
            const hipError_t e = hipStreamQuery(st_);
            if (e == hipSuccess && *seq != want) {
                err = "MOOD iteration snapshot missing";
                return HMHD_ERR_DEVICE;
            }
            if (e != hipSuccess && e != hipErrorNotReady) return hip_check(e, "hipStreamQuery");
        }
        __builtin_ia32_pause();
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    std::memcpy(ctl_.data(), h, sizeof(int32_t) * CTL_N * n);
    act_n_ = h[CTL_N * n + w];
    prof_collect(false);
    if (prof_mask >> PF_TILES & 1u) {
        if (visited) prof_n[PF_TILES] += act_n_;  // active tiles summed over MOOD iterations
        prof_ms[PF_TILES] += 1.0;    // MOOD iterations (host round trips)
    }
    return HMHD_OK;
}

// templated kernels run once per GP radius present (tiles of other radii return)
void Engine::launch_face_iter_groups(TileList tl, int grid, int stage) {
    for (int R = 1; R <= kMaxR; R++)
        if (has_R_[R]) launch_face_iter(L, tl, grid, R, stage, st_);
}

// The first-pass face kernel's form for the solves of radius R, from the
// tuning table; on the first launch of a (radius, size class) pair every
// form runs once untimed and then three times timed on this first pass
// (they all compute the same bits: each face's arithmetic does not depend
// on which thread does it), and the fastest is kept.
int Engine::face_variant(int R, int stage) {
    const int dflt = 0, mode = g_tune_mode;
    if (mode == 0) return dflt;
    if (mode >= 2) return face_variant_valid(mode - 2) ? mode - 2 : dflt;
    const int cls = face_size_class(rcells_[R]);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (const auto& e : g_tune)
            if (e.R == R && e.size_class == cls) return e.variant;
    }
    FaceTuneEntry e;
    e.R = R;
    e.size_class = cls;
    e.variant = dflt;
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return dflt;
    double best = 1e30;
    // one untimed launch of every form first: the first launches of a
    // process run on cold caches and clocks, which would rank the forms
    // by launch order rather than by speed
    for (int v = 0; v < kNumFaceVariants; v++) launch_face_full(L, R, v, d_fitems_[face_shape_of(v)],
                                                                 nfitems_[face_shape_of(v)], stage, st_);
    for (int v = 0; v < kNumFaceVariants; v++) {
        const int k = face_shape_of(v);
        double t = 1e30;
        for (int rep = 0; rep < 3; rep++) {
            (void)hipEventRecord(a, st_);
            launch_face_full(L, R, v, d_fitems_[k], nfitems_[k], stage, st_);
            (void)hipEventRecord(b, st_);
            float ms = 0.f;
            if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) ms = 1e30f;
            t = ms < t ? ms : t;
        }
        e.ms[v] = t;
        if (t < best) {
            best = t;
            e.variant = v;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto& x : g_tune)
        if (x.R == R && x.size_class == cls) return x.variant;  // another engine measured it meanwhile
    g_tune.push_back(e);
    return e.variant;
}

// MOOD iteration it >= 1 of stage s: the active tiles listed by iteration
// it - 1 (list (it - 1) & 1), its own scan writing list it & 1, then a
// snapshot into host slot it & 1.  The face kernel's `gface` blocks and
// the update's and lowering's `gupd` blocks stride over their lists.
int Engine::enqueue_iteration(int s, int it, int gface, int gupd) {
    const int r = (it - 1) & 1, w = it & 1;
    const TileList act{L.act[r], L.act_count + 2 * r, 0};
    const TileList actf{L.actf[r], L.act_count + 2 * r + 1, 0};
    // candidate stamps are never 0 (0 asks the scan for every tile)
    if (++stamp_ == 0) stamp_ = 1;
    // (an inner-loop round's lowering ran in the previous scan launch)
    prof_begin(PF_FACE_FLUX);
    launch_face_iter_groups(actf, gface, s);
    prof_end(PF_FACE_FLUX);
    prof_begin(PF_UPDATE);
    launch_update(L, act, gupd, s, stamp_, st_);
    prof_end(PF_UPDATE);
    prof_begin(PF_FLAGS);
    // the lowering after detect, the sync and the scan, which also end the
    // iteration (loop decisions, snapshot)
    launch_end_iter(L, w, stamp_, s, d_hctl_ + w * snap_stride(), next_seq(w), st_);
    prof_end(PF_FLAGS);
    return hip_check(hipGetLastError(), "kernel launch");
}

// The start of stage s: the loop's control words, the flags' stage set-up
// and the first pass's face kernel (every face at the top radius).
// Nothing here reads dt or changes the scheme.
void Engine::first_pass_faces(int s) {
    launch_stage_init(L, s, st_);
    int var[kMaxR + 1];
    for (int R = 1; R <= kMaxR; R++) var[R] = has_R_[R] ? face_variant(R, s) : 0;
    prof_begin(PF_FACE_FULL);
    for (int R = 1; R <= kMaxR; R++) {
        if (!has_R_[R]) continue;
        const int k = face_shape_of(var[R]);
        launch_face_full(L, R, var[R], d_fitems_[k], nfitems_[k], s, st_);
    }
    prof_end(PF_FACE_FULL);
}

// One RK stage: the MOOD loop of src/timeStep.f90:86-128 for every active
// solve.  The loop runs one iteration ahead of the host: iteration it + 1
// is enqueued before the host has read whether iteration it left work.  A
// speculative iteration after the last needed one is a no-op (the loop
// decisions have ended every solve's loop, and no kernel then changes the
// state or its flags), so the results and the pass counts are those of the
// plain loop.
int Engine::run_stage(int s, bool pre) {
    const int n = (int)inst_.size();
    const TileList all{L.pad, L.npad_dev, 1};
    // iteration 0, the first pass: every tile, every face at the top radius
    // (its start already enqueued by cfl() when `pre`)
    if (!pre) first_pass_faces(s);
    // One block per tile, except where a single solve's previous step says
    // the kernel has little to do: its same stage's pass first ran its
    // update at iteration pr.  Iterations it < pr found an inadmissible
    // face side, so their update and lowering did nothing (apart from the
    // first pass's scheme reset, and no candidate marks for a single
    // solve), and iterations 1..pr were inner-loop rounds, whose face list
    // holds only the tiles near an inadmissible cell.  A wrong guess costs
    // time only: the blocks stride over whatever the list holds.
    const int pr = L.one ? ready_pred_[s] : 0, quiet = std::min(L.npad, 512);
    prof_begin(PF_UPDATE);
    launch_update(L, all, pr > 0 ? quiet : L.npad, s, 0, st_);
    prof_end(PF_UPDATE);
    prof_begin(PF_FLAGS);
    launch_end_iter(L, 0, 0, s, d_hctl_, next_seq(0), st_);
    prof_end(PF_FLAGS);
    int rc = hip_check(hipGetLastError(), "kernel launch");
    if (rc) return rc;
    // Grids: iterations 1 and 2 read lists whose counts the host has not
    // seen (the first pass's list has every tile while rc and rcF are set
    // everywhere, until the first update); later ones the latest count.
    auto grid_for = [&](int listed) { return std::min(L.npad, listed + listed / 2 + 16); };
    auto early = [&](int i) { return enqueue_iteration(s, i, pr >= i ? quiet : L.npad, pr > i ? quiet : L.npad); };
    int ready = -1;  // this stage's pr
    if ((rc = early(1))) return rc;
    for (int it = 0;; it++) {
        if ((rc = read_snapshot(it & 1, it > 0))) return rc;
        if (L.one && ready < 0 && ctl_[CTL_FIRST] == 0) ready = it;
        // every continuing iteration lowers at least one scheme value, so
        // the loop ends; the cap only turns a defect into an error
        if (it > 64 * (kMaxR + 1) * 64) {
            err = "MOOD loop did not converge";
            return HMHD_ERR_DEVICE;
        }
        // the snapshot is taken after the loop decisions: RUN set means
        // the solve goes on to iteration it + 1
        bool more = false;
        for (int k = 0; k < n; k++)
            if (ctl_[(size_t)CTL_N * k + CTL_RUN]) more = true;
        if (!more) break;  // iteration it + 1, already enqueued, does nothing
        // iteration it + 2 reads the list iteration it + 1 writes
        if ((rc = it == 0 ? early(2) : enqueue_iteration(s, it + 2, grid_for(act_n_), grid_for(act_n_)))) return rc;
    }
    if (L.one) ready_pred_[s] = ready > 0 ? ready : 0;
    if (record_schemes) {
        size_t off = 0;
        for (auto& h : inst_) {
            const size_t np = (size_t)h.mx * h.my;
            if (h.active &&
                (rc = hip_check(hipMemcpyAsync(h_scheme_ + off, h.dev.scheme, np, hipMemcpyDeviceToHost, st_),
                                "hipMemcpyAsync(scheme)")))
                return rc;
            off += np;
        }
        if ((rc = sync())) return rc;
        off = 0;
        for (auto& h : inst_) {
            const size_t np = (size_t)h.mx * h.my;
            if (h.active) {
                h.stage_schemes.resize(3 * np);
                for (size_t p = 0; p < np; p++) h.stage_schemes[s * np + p] = h_scheme_[off + p];
            }
            off += np;
        }
    }
    return HMHD_OK;
}

int Engine::advance(const std::vector<double>& dt) {
    // cfl()'s first-pass face work stands if nothing changed since
    const bool pre = pre_ok_ && !dirty_;
    pre_ok_ = false;
    int rc = commit();
    if (rc) return rc;
    const int n = (int)inst_.size();
    size_t a = 0;
    for (int k = 0; k < n; k++) h_dt_[k] = inst_[k].active ? (a < dt.size() ? dt[a++] : 0.0) : 0.0;
    if (a != dt.size()) {
        err = "advance: one dt per active solve expected";
        return HMHD_ERR_ARG;
    }
    if (L.one) {
        // the single-solve kernels take the record, dt/dx and dt/dy as
        // k_set_dt computes them, from their arguments; the table's copy is
        // not read
        L.inst1.dtdx = h_dt_[0] / L.inst1.dx;
        L.inst1.dtdy = h_dt_[0] / L.inst1.dy;
    } else {
        if ((rc = hip_check(hipMemcpyAsync(d_dt_, h_dt_, sizeof(double) * n, hipMemcpyHostToDevice, st_),
                            "hipMemcpyAsync(dt)")))
            return rc;
        launch_set_dt(d_insts_, d_dt_, n, st_);
    }
    // U1 = FE(U); U2 = (3U + FE(U1))/4; U^{n+1} = (U + 2 FE(U2))/3
    for (int s = 0; s < 3; s++)
        if ((rc = run_stage(s, s == 0 && pre))) return rc;
    // (the single-solve kernels take rot and v0 from their arguments, and
    // commit() rebuilds the table from the host records)
    if (!L.one) launch_rotate(d_insts_, n, st_);
    if (L.one && L.inst1.active) {
        // as k_rotate
        L.inst1.rot = (L.inst1.rot + 1) % 3;
        L.inst1.v0 = 0;
    }
    for (int k = 0; k < n; k++) {
        InstHost& h = inst_[k];
        if (!h.active) continue;
        h.dev.rot = (h.dev.rot + 1) % 3;
        h.v0_valid = false;
        h.dev.v0 = 0;
        for (int s = 0; s < 3; s++) {
            h.passes[s] = ctl_[(size_t)CTL_N * k + CTL_PASSES + s];
            h.inner[s] = ctl_[(size_t)CTL_N * k + CTL_INNER + s];
        }
    }
    return hip_check(hipGetLastError(), "kernel launch");
}

int Engine::upload(int64_t id, const double* U, const double* V) {
    InstHost* h = find(id);
    if (!h) {
        err = "unknown solve id";
        return HMHD_ERR_ARG;
    }
    int rc = commit();  // staging buffers sized for every solve
    if (rc) return rc;
    const size_t np = (size_t)h->mx * h->my;
    if ((rc = hip_check(hipMemcpyAsync(d_stage_, U, sizeof(double) * NCONS * np, hipMemcpyHostToDevice, st_),
                        "hipMemcpyAsync(U)")))
        return rc;
    launch_unpack(h->dev, d_stage_, NCONS, false, h->dev.rot, st_);
    if ((rc = sync())) return rc;
    h->upU.assign(U, U + NCONS * np);
    if (V) {
        if ((rc = hip_check(hipMemcpyAsync(d_stage_, V, sizeof(double) * 10 * np, hipMemcpyHostToDevice, st_),
                            "hipMemcpyAsync(V)")))
            return rc;
        launch_unpack(h->dev, d_stage_, 10, true, 0, st_);
        if ((rc = sync())) return rc;
        h->upV.assign(V, V + 10 * np);
    } else {
        h->upV.clear();
    }
    h->v0_valid = V != nullptr;
    dirty_ = true;    // the device table carries the v0 flag
    pre_ok_ = false;  // and cfl()'s first-pass face work read the old state
    return HMHD_OK;
}

int Engine::cfl(std::vector<double>& dt_out) {
    int rc = commit();
    if (rc) return rc;
    const int n = (int)inst_.size();
    cfl_seq_ = unique_seq();
    volatile int32_t* seq = reinterpret_cast<volatile int32_t*>(h_cfl_ + 2 * n);
    prof_begin(PF_CFL);
    launch_cfl(L, n, d_hcfl_, reinterpret_cast<uint32_t*>(d_ctl_ + (size_t)CTL_N * n + 5),
               reinterpret_cast<int32_t*>(d_hcfl_ + 2 * n), cfl_seq_, st_);
    prof_end(PF_CFL);
    if ((rc = hip_check(hipGetLastError(), "kernel launch"))) return rc;
    // The next advance's first pass starts with face work that does not
    // depend on dt: enqueue it now, so it runs while dt makes its way
    // through the host.  The state it reads is the one this CFL reads.
    bool any = false;
    for (const auto& h : inst_) any = any || h.active;
    if (any && !pre_ok_) {
        first_pass_faces(0);
        if ((rc = hip_check(hipGetLastError(), "kernel launch"))) return rc;
        pre_ok_ = true;
    }
    // the maxima are in once the sequence word is (stored last, behind a
    // system-scope fence); an idle stream without it is an error
    for (unsigned k = 1; *seq != (int32_t)cfl_seq_; k++) {
        if (k % 1024 == 0) {
            const hipError_t e = hipStreamQuery(st_);
            if (e == hipSuccess && *seq != (int32_t)cfl_seq_) {
                err = "CFL result missing";
                return HMHD_ERR_DEVICE;
            }
            if (e != hipSuccess && e != hipErrorNotReady) return hip_check(e, "hipStreamQuery");
        }
        __builtin_ia32_pause();
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    prof_collect(false);
    dt_out.clear();
    for (int k = 0; k < n; k++) {
        if (!inst_[k].active) continue;
        double xm, ym;
        const unsigned long long bx = reinterpret_cast<volatile unsigned long long*>(h_cfl_)[2 * k];
        const unsigned long long by = reinterpret_cast<volatile unsigned long long*>(h_cfl_)[2 * k + 1];
        std::memcpy(&xm, &bx, 8);
        std::memcpy(&ym, &by, 8);
        const Inst& d = inst_[k].dev;
        // dt = cfl * min(dx/xmax, dy/ymax) (src/cfl.f90:50); a NaN speed
        // gives a NaN dt so the driver's validTimeStep aborts
        const double a = d.dx / xm, b = d.dy / ym;
        double m = a < b ? a : b;
        if (std::isnan(a) || std::isnan(b)) m = std::nan("");
        dt_out.push_back(d.cfl * m);
    }
    return HMHD_OK;
}

int Engine::download(int64_t id, double* U, double* V, int32_t* scheme) {
    InstHost* h = find(id);
    if (!h) {
        err = "unknown solve id";
        return HMHD_ERR_ARG;
    }
    int rc = commit();
    if (rc) return rc;
    const size_t np = (size_t)h->mx * h->my;
    if (h->v0_valid) {
        // no step since the upload: hand back what the caller uploaded
        if (U) std::memcpy(U, h->upU.data(), sizeof(double) * NCONS * np);
        if (V) std::memcpy(V, h->upV.data(), sizeof(double) * 10 * np);
        U = V = nullptr;
    }
    const int buf = h->dev.rot;
    if (U) {
        launch_pack(h->dev, buf, d_stage_, nullptr, nullptr, st_);
        if ((rc = hip_check(hipMemcpyAsync(U, d_stage_, sizeof(double) * NCONS * np, hipMemcpyDeviceToHost, st_),
                            "D2H U")) ||
            (rc = sync()))
            return rc;
    }
    if (V) {
        launch_pack(h->dev, buf, nullptr, d_stage_, nullptr, st_);
        if ((rc = hip_check(hipMemcpyAsync(V, d_stage_, sizeof(double) * 10 * np, hipMemcpyDeviceToHost, st_),
                            "D2H V")) ||
            (rc = sync()))
            return rc;
    }
    if (scheme) {
        launch_pack(h->dev, buf, nullptr, nullptr, d_stage_i_, st_);
        if ((rc = hip_check(hipMemcpyAsync(scheme, d_stage_i_, sizeof(int32_t) * np, hipMemcpyDeviceToHost, st_),
                            "D2H scheme")) ||
            (rc = sync()))
            return rc;
    }
    return HMHD_OK;
}

int Engine::sync() {
    if (!st_) return HMHD_OK;
    const int rc = hip_check(hipStreamSynchronize(st_), "hipStreamSynchronize");
    if (!rc) prof_collect(true);
    return rc;
}

}  // namespace hmhd
