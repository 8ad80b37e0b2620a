// hmhd_solver.cpp -- the host engine: device allocation, tile work lists,
// and the MOOD-pass loop of each SSP-RK3 stage (src/timeStep.f90:45-171).
#include "hmhd_solver.hpp"

#include <cmath>
#include <cstring>
#include <mutex>

#include "gp_tables.hpp"
#include "hmhd_kernels.hpp"

namespace hmhd {

// ------------------------------------------------------------------ //
// GP table registry (__constant__ sets are shared by every handle)
namespace {
struct GpSet {
    int refs = 0;
    double pv[kPvSet];
    double qw[(kMaxR + 1) * 4];
};
std::mutex g_gp_mu;
GpSet g_gp[kMaxGpSets];
}  // namespace

int gp_acquire(const double* pv, const double* qw, std::string& err) {
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
        err = "more than 8 distinct GP table sets in one process";
        return -1;
    }
    GpSet& g = g_gp[free_set];
    std::memcpy(g.pv, pv, sizeof g.pv);
    std::memcpy(g.qw, qw, sizeof g.qw);
    set_gp_tables(free_set, g.pv, g.qw, nullptr);
    if (hipDeviceSynchronize() != hipSuccess) {
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

int Engine::hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return HMHD_OK;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return HMHD_ERR_DEVICE;
}

Engine::~Engine() {
    if (st) (void)hipStreamSynchronize(st);
    for (auto& in : inst) {
        for (void* p : in.allocs) (void)hipFree(p);
        gp_release(in.gpset);
    }
    for (Item* p : d_items) (void)hipFree(p);
    if (d_insts) (void)hipFree(d_insts);
    if (d_ctl) (void)hipFree(d_ctl);
    if (d_dt) (void)hipFree(d_dt);
    if (d_cfl) (void)hipFree(d_cfl);
    if (d_stage) (void)hipFree(d_stage);
    if (d_stage_i) (void)hipFree(d_stage_i);
    if (h_ctl) (void)hipHostFree(h_ctl);
    if (h_dt) (void)hipHostFree(h_dt);
    if (h_cfl) (void)hipHostFree(h_cfl);
    if (h_scheme) (void)hipHostFree(h_scheme);
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    if (st) (void)hipStreamDestroy(st);
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
    if ((long long)(c.nx + 2 * ngc) * (c.ny + 2 * ngc) > (1ll << 31)) {
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

int Engine::add_instance(const hmhd_config& c) {
    if (finalized) {
        err = "engine already finalized";
        return -1;
    }
    if (check_config(c, err) != HMHD_OK) return -1;
    InstHost h;
    h.cfg = c;
    h.ngc = c.R + 2;
    h.mx = c.nx + 2 * h.ngc;
    h.my = c.ny + 2 * h.ngc;
    if (gp_compute(c.R, c.ell_over_delta, c.xbeg, c.xend, c.ybeg, c.yend, c.nx, c.ny, h.pv, h.qw) != 0) {
        err = "GP covariance matrix is not positive definite (src/linAlgQuadPrecision.f90:136)";
        return -1;
    }
    inst.push_back(h);
    return (int)inst.size() - 1;
}

static void tiles(std::vector<Item>& v, int k, int w, int h, int t) {
    const int nt_x = (w + t - 1) / t, nt_y = (h + t - 1) / t;
    for (int ty = 0; ty < nt_y; ty++)
        for (int tx = 0; tx < nt_x; tx++) v.push_back(Item{k, (int16_t)tx, (int16_t)ty});
}

int Engine::finalize() {
    if (finalized) return HMHD_OK;
    if (inst.empty()) {
        err = "no instances";
        return HMHD_ERR_ARG;
    }
    int rc;
    if ((rc = hip_check(hipStreamCreate(&st), "hipStreamCreate"))) return rc;
    const int n = (int)inst.size();
    size_t max_pad = 0;
    for (int k = 0; k < n; k++) {
        InstHost& h = inst[k];
        const hmhd_config& c = h.cfg;
        h.gpset = gp_acquire(h.pv, h.qw, err);
        if (h.gpset < 0) return HMHD_ERR_DEVICE;
        const size_t nin = (size_t)c.nx * c.ny, npad = (size_t)h.mx * h.my;
        max_pad = npad > max_pad ? npad : max_pad;
        auto alloc = [&](size_t bytes, void** p) -> int {
            int r = hip_check(hipMalloc(p, bytes), "hipMalloc");
            if (r) return r;
            h.allocs.push_back(*p);
            return hip_check(hipMemsetAsync(*p, 0, bytes, st), "hipMemsetAsync");
        };
        Inst& d = h.dev;
        d.nx = c.nx;
        d.ny = c.ny;
        d.ngc = h.ngc;
        d.mx = h.mx;
        d.my = h.my;
        d.R = c.R;
        d.bc = c.bc;
        d.hydro = c.force_hydro ? 1 : 0;
        d.gpset = h.gpset;
        d.active = 1;
        d.dx = (c.xend - c.xbeg) / c.nx;
        d.dy = (c.yend - c.ybeg) / c.ny;
        d.mindxdy = d.dy < d.dx ? d.dy : d.dx;
        d.maxdxdy = d.dy > d.dx ? d.dy : d.dx;
        d.gamma = c.gamma;
        d.small_p = c.small_p;
        d.small_rho = c.small_rho;
        d.cfl = c.cfl;
        for (int b = 0; b < 3; b++)
            if ((rc = alloc(sizeof(double) * NCONS * nin, (void**)&d.U[b]))) return rc;
        if ((rc = alloc(sizeof(double) * NCONS * nin, (void**)&d.V0))) return rc;
        if ((rc = alloc(sizeof(double) * 2 * NCONS * npad, (void**)&d.FR))) return rc;
        if ((rc = alloc(sizeof(double) * 2 * npad, (void**)&d.FLB))) return rc;
        if ((rc = alloc(npad, (void**)&d.scheme))) return rc;
        if ((rc = alloc(npad, (void**)&d.rc))) return rc;
        if ((rc = alloc(npad, (void**)&d.rcF))) return rc;
        if ((rc = alloc(npad, (void**)&d.inadm))) return rc;
        if ((rc = alloc(npad, (void**)&d.low))) return rc;
        if ((rc = alloc(npad, (void**)&d.rec))) return rc;
        if ((rc = alloc(4 * npad, (void**)&d.faceRR))) return rc;
        if (c.bc == HMHD_BC_PERIODIC) any_periodic = true;
    }
    if ((rc = hip_check(hipMalloc(&d_ctl, sizeof(int32_t) * CTL_N * n), "hipMalloc"))) return rc;
    if ((rc = hip_check(hipMemsetAsync(d_ctl, 0, sizeof(int32_t) * CTL_N * n, st), "hipMemsetAsync"))) return rc;
    if ((rc = hip_check(hipMalloc(&d_cfl, sizeof(unsigned long long) * 2 * n), "hipMalloc"))) return rc;
    if ((rc = hip_check(hipMalloc(&d_dt, sizeof(double) * n), "hipMalloc"))) return rc;
    if ((rc = hip_check(hipHostMalloc(&h_ctl, sizeof(int32_t) * CTL_N * n), "hipHostMalloc"))) return rc;
    if ((rc = hip_check(hipHostMalloc(&h_dt, sizeof(double) * n), "hipHostMalloc"))) return rc;
    if ((rc = hip_check(hipHostMalloc(&h_cfl, sizeof(unsigned long long) * 2 * n), "hipHostMalloc"))) return rc;
    for (int k = 0; k < n; k++) {
        inst[k].dev.ctl = d_ctl + CTL_N * k;
        inst[k].dev.cflmax = d_cfl + 2 * k;
    }
    std::vector<Inst> hv(n);
    for (int k = 0; k < n; k++) hv[k] = inst[k].dev;
    if ((rc = hip_check(hipMalloc(&d_insts, sizeof(Inst) * n), "hipMalloc"))) return rc;
    if ((rc = hip_check(hipMemcpy(d_insts, hv.data(), sizeof(Inst) * n, hipMemcpyHostToDevice), "hipMemcpy")))
        return rc;

    // tile work lists: face domain [s-2, e+3]^2 per radius, update domain
    // [s, e+1]^2, padded grid per radius, interior
    std::vector<Item> upd, cell, pad_all;
    std::vector<Item> faceR[kMaxR + 1], padR[kMaxR + 1];
    for (int k = 0; k < n; k++) {
        const hmhd_config& c = inst[k].cfg;
        tiles(faceR[c.R], k, c.nx + 5, c.ny + 5, kFaceTX);
        tiles(padR[c.R], k, inst[k].mx, inst[k].my, kCellT);
        tiles(upd, k, c.nx + 1, c.ny + 1, kCellT);
        tiles(cell, k, c.nx, c.ny, kCellT);
    }
    static_assert(kFaceTX == kFaceTY, "square face tiles assumed by the work list");
    auto upload_items = [&](const std::vector<Item>& v, const Item** out, int* count) -> int {
        *count = (int)v.size();
        *out = nullptr;
        if (v.empty()) return HMHD_OK;
        Item* p = nullptr;
        int r = hip_check(hipMalloc(&p, sizeof(Item) * v.size()), "hipMalloc");
        if (r) return r;
        d_items.push_back(p);
        r = hip_check(hipMemcpy(p, v.data(), sizeof(Item) * v.size(), hipMemcpyHostToDevice), "hipMemcpy");
        *out = p;
        return r;
    };
    for (int R = 1; R <= kMaxR; R++) {
        if ((rc = upload_items(faceR[R], &face_R[R], &nface_R[R]))) return rc;
        if ((rc = upload_items(padR[R], &pad_R[R], &npad_R[R]))) return rc;
        pad_all.insert(pad_all.end(), padR[R].begin(), padR[R].end());
    }
    L.insts = d_insts;
    if ((rc = upload_items(upd, &L.upd, &L.nupd))) return rc;
    if ((rc = upload_items(cell, &L.cell, &L.ncell))) return rc;
    if ((rc = upload_items(pad_all, &L.pad, &L.npad))) return rc;

    d_stage_n = max_pad * 10;
    if ((rc = hip_check(hipMalloc(&d_stage, sizeof(double) * d_stage_n), "hipMalloc"))) return rc;
    if ((rc = hip_check(hipMalloc(&d_stage_i, sizeof(int32_t) * max_pad), "hipMalloc"))) return rc;
    h_scheme_n = 0;
    for (auto& h : inst) h_scheme_n += (size_t)h.mx * h.my;
    if ((rc = hip_check(hipHostMalloc(&h_scheme, h_scheme_n), "hipHostMalloc"))) return rc;

    ctl_host.assign((size_t)CTL_N * n, 0);
    stage_schemes.assign(n, {});
    v0_valid.assign(n, false);
    uploadedU.assign(n, {});
    uploadedV.assign(n, {});
    if ((rc = hip_check(hipStreamSynchronize(st), "hipStreamSynchronize"))) return rc;
    finalized = true;
    return HMHD_OK;
}

int Engine::set_gp(int k, const double* pv, const double* qw) {
    InstHost& h = inst[k];
    int s = gp_acquire(pv, qw, err);
    if (s < 0) return HMHD_ERR_DEVICE;
    gp_release(h.gpset);
    h.gpset = s;
    std::memcpy(h.pv, pv, sizeof h.pv);
    std::memcpy(h.qw, qw, sizeof h.qw);
    h.dev.gpset = s;
    return hip_check(hipMemcpy(&d_insts[k].gpset, &s, sizeof(int32_t), hipMemcpyHostToDevice), "hipMemcpy");
}

// ------------------------------------------------------------------ //
// profiling: events bracket each launch of a kernel family

void Engine::prof_begin(int fam) {
    if (!prof_on) return;
    if (ev_next + 2 > ev_pool.size()) {
        for (int k = 0; k < 64; k++) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            ev_pool.push_back(e);
        }
    }
    ev_a = ev_pool[ev_next++];
    ev_fam = fam;
    (void)hipEventRecord(ev_a, st);
}

void Engine::prof_end(int fam) {
    if (!prof_on || ev_fam != fam || !ev_a) return;
    hipEvent_t b = ev_pool[ev_next++];
    (void)hipEventRecord(b, st);
    ev_pending.push_back({fam, {ev_a, b}});
    ev_a = nullptr;
    ev_fam = -1;
}

void Engine::prof_collect() {
    for (auto& p : ev_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
            prof_ms[p.first] += ms;
            prof_n[p.first] += 1;
        }
    }
    ev_pending.clear();
    ev_next = 0;
}

// ------------------------------------------------------------------ //

int Engine::read_ctl() {
    const size_t n = inst.size();
    int rc = hip_check(hipMemcpyAsync(h_ctl, d_ctl, sizeof(int32_t) * CTL_N * n, hipMemcpyDeviceToHost, st),
                       "hipMemcpyAsync(ctl)");
    if (rc) return rc;
    if ((rc = hip_check(hipStreamSynchronize(st), "hipStreamSynchronize"))) return rc;
    std::memcpy(ctl_host.data(), h_ctl, sizeof(int32_t) * CTL_N * n);
    prof_collect();
    return HMHD_OK;
}

void Engine::launch_face_groups(bool full, int in) {
    for (int R = 1; R <= kMaxR; R++) {
        if (!nface_R[R]) continue;
        LaunchSet l = L;
        l.face = face_R[R];
        l.nface = nface_R[R];
        launch_face(l, R, full, in, st);
    }
}

void Engine::launch_recon_groups(int in) {
    for (int R = 1; R <= kMaxR; R++) {
        if (!npad_R[R]) continue;
        LaunchSet l = L;
        l.pad = pad_R[R];
        l.npad = npad_R[R];
        launch_recon_check(l, R, in, st);
    }
}

// one RK stage: the MOOD loop of src/timeStep.f90:86-128 for every instance
int Engine::run_stage(int s, int in, int out, int base) {
    const int n = (int)inst.size();
    launch_ctl(L, n, OP_BEGIN, s, st);
    launch_stage_init(L, st);
    prof_begin(PF_FACE_FULL);
    launch_face_groups(true, in);
    prof_end(PF_FACE_FULL);
    prof_begin(PF_UPDATE);
    launch_update(L, in, out, base, s, st);
    prof_end(PF_UPDATE);
    prof_begin(PF_FLAGS);
    launch_lower_detect(L, any_periodic, st);
    prof_end(PF_FLAGS);
    int rc = hip_check(hipGetLastError(), "kernel launch");
    if (rc) return rc;
    if ((rc = read_ctl())) return rc;
    for (int iter = 0;; iter++) {
        // every continuing iteration lowers at least one scheme value, so
        // the loop ends; the cap only turns a defect into an error
        if (iter > 64 * (kMaxR + 1) * 64) {
            err = "MOOD loop did not converge";
            return HMHD_ERR_DEVICE;
        }
        bool more = false;
        for (int k = 0; k < n; k++) {
            const int32_t* c = &ctl_host[(size_t)CTL_N * k];
            if (c[CTL_RUN] && (c[CTL_ANY_INADM] || c[CTL_ANY_LOWER])) more = true;
        }
        if (!more) break;
        launch_ctl(L, n, OP_NEXT, s, st);
        prof_begin(PF_RECON);
        launch_lower_inner(L, st);
        launch_recon_groups(in);
        prof_end(PF_RECON);
        prof_begin(PF_FACE_FLUX);
        launch_face_groups(false, in);
        prof_end(PF_FACE_FLUX);
        prof_begin(PF_UPDATE);
        launch_update(L, in, out, base, s, st);
        prof_end(PF_UPDATE);
        prof_begin(PF_FLAGS);
        launch_lower_detect(L, any_periodic, st);
        prof_end(PF_FLAGS);
        if ((rc = hip_check(hipGetLastError(), "kernel launch"))) return rc;
        if ((rc = read_ctl())) return rc;
    }
    if (record_schemes) {
        size_t off = 0;
        for (auto& h : inst) {
            const size_t np = (size_t)h.mx * h.my;
            if ((rc = hip_check(hipMemcpyAsync(h_scheme + off, h.dev.scheme, np, hipMemcpyDeviceToHost, st),
                                "hipMemcpyAsync(scheme)")))
                return rc;
            off += np;
        }
        if ((rc = hip_check(hipStreamSynchronize(st), "hipStreamSynchronize"))) return rc;
        off = 0;
        for (int k = 0; k < n; k++) {
            const size_t np = (size_t)inst[k].mx * inst[k].my;
            auto& dst = stage_schemes[k];
            dst.resize(3 * np);
            for (size_t p = 0; p < np; p++) dst[s * np + p] = h_scheme[off + p];
            off += np;
        }
    }
    return HMHD_OK;
}

int Engine::advance(const double* dt_host) {
    const int n = (int)inst.size();
    std::memcpy(h_dt, dt_host, sizeof(double) * n);
    int rc = hip_check(hipMemcpyAsync(d_dt, h_dt, sizeof(double) * n, hipMemcpyHostToDevice, st), "hipMemcpyAsync(dt)");
    if (rc) return rc;
    launch_set_dt(d_insts, d_dt, n, st);
    const int a = cur, b = (cur + 1) % 3, c = (cur + 2) % 3;
    // U1 = FE(U); U2 = (3U + FE(U1))/4; U^{n+1} = (U + 2 FE(U2))/3
    if ((rc = run_stage(0, a, b, a))) return rc;
    if ((rc = run_stage(1, b, c, a))) return rc;
    if ((rc = run_stage(2, c, b, a))) return rc;
    cur = b;
    for (int k = 0; k < n; k++) v0_valid[k] = false;
    return HMHD_OK;
}

int Engine::upload(int k, const double* U, const double* V) {
    InstHost& h = inst[k];
    const size_t np = (size_t)h.mx * h.my, nin = (size_t)h.cfg.nx * h.cfg.ny;
    int rc = hip_check(hipMemcpy(d_stage, U, sizeof(double) * NCONS * np, hipMemcpyHostToDevice), "hipMemcpy(U)");
    if (rc) return rc;
    launch_unpack(d_insts, k, d_stage, NCONS, false, cur, nin, st);
    if ((rc = hip_check(hipStreamSynchronize(st), "unpack U"))) return rc;
    uploadedU[k].assign(U, U + NCONS * np);
    if (V) {
        if ((rc = hip_check(hipMemcpy(d_stage, V, sizeof(double) * 10 * np, hipMemcpyHostToDevice), "hipMemcpy(V)")))
            return rc;
        launch_unpack(d_insts, k, d_stage, 10, true, 0, nin, st);
        if ((rc = hip_check(hipStreamSynchronize(st), "unpack V"))) return rc;
        uploadedV[k].assign(V, V + 10 * np);
        v0_valid[k] = true;
    } else {
        uploadedV[k].clear();
        v0_valid[k] = false;
    }
    return HMHD_OK;
}

int Engine::cfl(std::vector<double>& dt_out) {
    const int n = (int)inst.size();
    // instances with uploaded primitives use them, the rest derive V from U
    bool any_v0 = false, any_u = false;
    for (int k = 0; k < n; k++) (v0_valid[k] ? any_v0 : any_u) = true;
    prof_begin(PF_CFL);
    if (any_v0 && any_u) {
        err = "mixed first-step CFL sources in one batch";
        return HMHD_ERR_ARG;
    }
    launch_cfl(L, n, cur, any_v0, st);
    prof_end(PF_CFL);
    int rc = hip_check(hipMemcpyAsync(h_cfl, d_cfl, sizeof(unsigned long long) * 2 * n, hipMemcpyDeviceToHost, st),
                       "hipMemcpyAsync(cfl)");
    if (rc) return rc;
    if ((rc = hip_check(hipStreamSynchronize(st), "hipStreamSynchronize"))) return rc;
    prof_collect();
    dt_out.resize(n);
    for (int k = 0; k < n; k++) {
        double xm, ym;
        std::memcpy(&xm, &h_cfl[2 * k], 8);
        std::memcpy(&ym, &h_cfl[2 * k + 1], 8);
        const Inst& d = inst[k].dev;
        // dt = cfl * min(dx/xmax, dy/ymax) (src/cfl.f90:50); a NaN speed
        // gives a NaN dt so the driver's validTimeStep aborts
        const double a = d.dx / xm, b = d.dy / ym;
        double m = a < b ? a : b;
        if (std::isnan(a) || std::isnan(b)) m = std::nan("");
        dt_out[k] = d.cfl * m;
    }
    return HMHD_OK;
}

int Engine::download(int k, double* U, double* V, int32_t* scheme) {
    InstHost& h = inst[k];
    const size_t np = (size_t)h.mx * h.my;
    int rc;
    if (v0_valid[k]) {
        // no step since the upload: hand back what the caller uploaded
        if (U) std::memcpy(U, uploadedU[k].data(), sizeof(double) * NCONS * np);
        if (V) std::memcpy(V, uploadedV[k].data(), sizeof(double) * 10 * np);
        U = V = nullptr;
    }
    if (U || V) {
        double* dU = d_stage;
        double* dV = d_stage;
        if (U && V) {
            // two passes through the staging buffer
            launch_pack(d_insts, k, cur, dU, nullptr, nullptr, np, st);
            if ((rc = hip_check(hipMemcpyAsync(U, dU, sizeof(double) * NCONS * np, hipMemcpyDeviceToHost, st),
                                "hipMemcpyAsync(U)")))
                return rc;
            if ((rc = hip_check(hipStreamSynchronize(st), "download U"))) return rc;
            U = nullptr;
        }
        launch_pack(d_insts, k, cur, U ? dU : nullptr, V ? dV : nullptr, nullptr, np, st);
        if (U &&
            (rc = hip_check(hipMemcpyAsync(U, dU, sizeof(double) * NCONS * np, hipMemcpyDeviceToHost, st), "D2H U")))
            return rc;
        if (V && (rc = hip_check(hipMemcpyAsync(V, dV, sizeof(double) * 10 * np, hipMemcpyDeviceToHost, st), "D2H V")))
            return rc;
        if ((rc = hip_check(hipStreamSynchronize(st), "download"))) return rc;
    }
    if (scheme) {
        launch_pack(d_insts, k, cur, nullptr, nullptr, d_stage_i, np, st);
        if ((rc = hip_check(hipMemcpyAsync(scheme, d_stage_i, sizeof(int32_t) * np, hipMemcpyDeviceToHost, st),
                            "D2H scheme")))
            return rc;
        if ((rc = hip_check(hipStreamSynchronize(st), "download scheme"))) return rc;
    }
    return HMHD_OK;
}

int Engine::sync() { return hip_check(hipStreamSynchronize(st), "hipStreamSynchronize"); }

}  // namespace hmhd
