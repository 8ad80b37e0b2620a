// hmhd_api.cpp -- the extern "C" entry points declared in include/hmhd.h:
// the single-run handle (one solve in an Engine), the batch (many solves
// advanced together to their own end times) and the server (a request loop
// that admits solves by HBM budget while others run).
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>

#include "../../include/hmhd.h"
#include "gp_tables.hpp"
#include "hmhd_kernels.hpp"
#include "hmhd_runner.hpp"
#include "hmhd_solver.hpp"

struct hmhd_handle {
    hmhd::Engine eng;
    int64_t id = -1;
    std::string err;
};

struct hmhd_batch_handle {
    hmhd::Engine eng;
    hmhd::Runner run{eng};
    std::string err;
};

namespace {
thread_local std::string g_create_err;

int fail(std::string& where, int code, const std::string& msg) {
    where = msg;
    return code;
}

int create_code(const std::string& msg) {
    if (msg.find("positive definite") != std::string::npos) return HMHD_ERR_GP;
    if (msg.find("hip") != std::string::npos || msg.find("constant memory") != std::string::npos)
        return HMHD_ERR_DEVICE;
    return HMHD_ERR_ARG;
}

// device table layout -> the reference's GP_predictionVectors(nsMax, nqMax,
// 4, R), GP_quadratureWeights(nqMax, R), GP_stencIdxs(2, nsMax, R)
void export_gp(int R, const double* pv, const double* qw, double* predvec, double* quadw, int32_t* stenc) {
    const int nsm = hmhd::gp_ns(R), nqm = hmhd::gp_nq(R);
    if (predvec) {
        std::memset(predvec, 0, sizeof(double) * nsm * nqm * 4 * R);
        for (int rr = 1; rr <= R; rr++)
            for (int f = 0; f < 4; f++)
                for (int q = 0; q < hmhd::gp_nq(rr); q++)
                    for (int k = 0; k < hmhd::gp_ns(rr); k++)
                        predvec[(((size_t)(rr - 1) * 4 + f) * nqm + q) * nsm + k] =
                            pv[hmhd::gp_pv_off(rr) + (f * hmhd::gp_nq(rr) + q) * hmhd::gp_ns(rr) + k];
    }
    if (quadw) {
        std::memset(quadw, 0, sizeof(double) * nqm * R);
        for (int rr = 1; rr <= R; rr++)
            for (int q = 0; q < hmhd::gp_nq(rr); q++) quadw[(rr - 1) * nqm + q] = qw[rr * 4 + q];
    }
    if (stenc) {
        std::memset(stenc, 0, sizeof(int32_t) * 2 * nsm * R);
        for (int rr = 1; rr <= R; rr++) {
            int c = 0;
            for (int i = -rr; i <= rr; i++)
                for (int j = -rr; j <= rr; j++)
                    if ((i < 0 ? -i : i) + (j < 0 ? -j : j) <= rr) {
                        stenc[((rr - 1) * nsm + c) * 2 + 0] = i;
                        stenc[((rr - 1) * nsm + c) * 2 + 1] = j;
                        c++;
                    }
        }
    }
}

int profile_read(hmhd::Engine& e, double* ms, int64_t* launches, int32_t reset) {
    for (int f = 0; f < hmhd::PF_N; f++) {
        if (ms) ms[f] = e.prof_ms[f];
        if (launches) launches[f] = e.prof_n[f];
        if (reset) {
            e.prof_ms[f] = 0.0;
            e.prof_n[f] = 0;
        }
    }
    return HMHD_OK;
}
}  // namespace

namespace hmhd {
void set_free_error(const char* msg) { g_create_err = msg; }
}  // namespace hmhd

// ------------------------------------------------------------------ //
// the request loop of the served form

// The server runs its admitted solves in lanes: a lane is an engine with its
// own stream and worker thread, and lanes overlap on the device (their
// engines share no stream and make no device-wide synchronization).  Lane
// 0 batches any number of solves, stepped together; each further lane
// holds one large solve at a time, which then steps alone with the
// single-solve kernels instead of waiting in lockstep for the batch.  A
// request of at least kSoloCells interior cells takes a free solo lane if
// there is one, otherwise it joins the batch.  Four lanes in all, one per
// hardware queue a process gets by default.
struct hmhd_server_handle {
    static constexpr int kSoloLanes = 3;
    static constexpr long long kSoloCells = 512 * 512;

    struct Req {
        int64_t id = 0;
        hmhd_config cfg{};
        std::vector<double> U, V;
        double tmax = 0.0;
        int32_t nstepmax = 0;
        size_t bytes = 0;
        int lane = -1;
        int64_t gp_wait = -1;  // re-queued for a GP table set: wait until more than this many solves finished
        hmhd_result res{};
        std::vector<double> Uout, Vout;
        std::vector<int32_t> Sout;
        bool done = false;
    };

    struct Lane {
        hmhd::Engine eng;  // touched by this lane's thread only
        hmhd::Runner run{eng};
        bool solo = false;
        std::vector<int64_t> todo;  // assigned requests not yet added (under mu)
        int assigned = 0;           // assigned and not finished (under mu)
        std::unordered_map<int64_t, int64_t> by_solve;  // solve id -> request id (this lane's thread)
        std::condition_variable cv;
        std::thread th;
    };

    uint64_t budget = 0;
    int32_t max_active = 0;
    std::mutex mu;  // guards the queue, the requests' states, the lanes' todo/assigned, the stats
    std::condition_variable cv_done;
    std::deque<int64_t> queue;
    std::unordered_map<int64_t, Req> reqs;  // elements keep their addresses across inserts
    std::vector<std::unique_ptr<Lane>> lanes;
    uint64_t reserved = 0;  // footprint of the assigned requests
    int64_t next_req = 1;
    bool stop = false;
    std::string err;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    hmhd_server_stats st{};
    double busy_s = 0.0;

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

    void finish_locked(Req& r, int32_t state) {
        r.res.state = state;
        r.res.done_s = now();
        r.done = true;
    }

    // Assign queued requests to lanes in arrival order while they fit the
    // budget and the slot limit; a request that does not fit yet is passed
    // over for later, smaller ones.  One larger than the whole budget fails
    // at once.
    void admit_locked() {
        bool changed = false;
        int active = 0;
        for (auto& l : lanes) active += l->assigned;
        for (auto it = queue.begin(); it != queue.end();) {
            Req& r = reqs[*it];
            if (r.bytes > budget) {
                err = "request footprint exceeds the HBM budget";
                finish_locked(r, HMHD_REQ_FAILED);
                it = queue.erase(it);
                changed = true;
                continue;
            }
            if (max_active > 0 && active >= max_active) break;
            if (reserved + r.bytes > budget || r.gp_wait == st.completed) {
                ++it;
                continue;
            }
            int lane = 0;
            if ((long long)r.cfg.nx * r.cfg.ny >= kSoloCells)
                for (size_t k = 1; k < lanes.size(); k++)
                    if (lanes[k]->assigned == 0) {
                        lane = (int)k;
                        break;
                    }
            Lane& L = *lanes[lane];
            r.lane = lane;
            r.res.state = HMHD_REQ_RUNNING;
            r.res.admit_s = now();
            L.todo.push_back(r.id);
            L.assigned += 1;
            active += 1;
            reserved += r.bytes;
            L.cv.notify_one();
            it = queue.erase(it);
        }
        st.active = active;
        st.queued = (int32_t)queue.size();
        if (st.active > st.peak_active) st.peak_active = st.active;
        st.bytes_in_use = reserved;
        if (st.bytes_in_use > st.peak_bytes) st.peak_bytes = st.bytes_in_use;
        if (changed) cv_done.notify_all();
    }

    // a request leaves its lane: its reservation and slot are freed
    void release_locked(Lane& L, Req& r) {
        L.assigned -= 1;
        reserved -= r.bytes;
    }

    void lane_loop(Lane& L) {
        std::vector<int64_t> todo, fin;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu);
                L.cv.wait(lk, [&] { return stop || !L.todo.empty() || !L.by_solve.empty(); });
                if (stop) return;
                todo.swap(L.todo);
            }
            // new solves: GP tables, device memory and upload, outside the lock
            for (int64_t rid : todo) {
                Req* r;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    r = &reqs[rid];
                }
                const int64_t sid = L.eng.add(r->cfg);
                const bool gp_full = sid < 0 && L.eng.gp_full;
                int rc = sid < 0 ? HMHD_ERR_DEVICE : L.eng.upload(sid, r->U.data(), r->V.empty() ? nullptr : r->V.data());
                if (rc != HMHD_OK) {
                    std::lock_guard<std::mutex> lk(mu);
                    if (sid >= 0) (void)L.eng.remove(sid);
                    release_locked(L, *r);
                    bool others = false;
                    for (auto& l : lanes) others = others || l->assigned > 0;
                    if (gp_full && others) {
                        // every GP table set is in use by running solves: wait for one to finish
                        r->lane = -1;
                        r->gp_wait = st.completed;
                        r->res.state = HMHD_REQ_QUEUED;
                        queue.push_front(rid);
                    } else {
                        err = L.eng.err;
                        finish_locked(*r, HMHD_REQ_FAILED);
                        cv_done.notify_all();
                    }
                    admit_locked();
                    continue;
                }
                std::vector<double>().swap(r->U);
                std::vector<double>().swap(r->V);
                L.run.track(sid, r->tmax, r->nstepmax);
                L.by_solve[sid] = rid;
            }
            todo.clear();
            if (L.by_solve.empty()) continue;
            // one engine step of every solve of the lane, outside the lock
            fin.clear();
            const double a = now();
            const int rc = L.run.step(fin);
            const double b = now();
            if (rc != HMHD_OK) {
                // a device error leaves every solve of the lane undefined
                std::lock_guard<std::mutex> lk(mu);
                err = L.eng.err;
                for (auto& kv : L.by_solve) {
                    Req& r = reqs[kv.second];
                    finish_locked(r, HMHD_REQ_FAILED);
                    release_locked(L, r);
                    (void)L.eng.remove(kv.first);
                    L.run.forget(kv.first);
                }
                L.by_solve.clear();
                admit_locked();
                cv_done.notify_all();
                continue;
            }
            // results of the finished solves, outside the lock: nobody reads
            // a request's arrays before it is marked done (its `res`, which
            // poll reads at any time, is written under the lock below)
            struct Done {
                int64_t rid;
                int32_t state, steps;
                int64_t zone_updates;
                double t;
            };
            std::vector<Done> done;
            for (int64_t sid : fin) {
                const int64_t rid = L.by_solve[sid];
                Req* r;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    r = &reqs[rid];
                }
                const hmhd::RunInfo& ri = L.run.info[sid];
                const hmhd::InstHost* h = L.eng.find(sid);
                const size_t np = (size_t)h->mx * h->my;
                r->Uout.resize(8 * np);
                r->Vout.resize(10 * np);
                r->Sout.resize(np);
                int32_t state = ri.state == hmhd::RUN_DONE ? HMHD_REQ_DONE : HMHD_REQ_ABORTED;
                if (L.eng.download(sid, r->Uout.data(), r->Vout.data(), r->Sout.data()) != HMHD_OK)
                    state = HMHD_REQ_FAILED;
                done.push_back(Done{rid, state, ri.steps, ri.zone_updates, ri.t});
                (void)L.eng.remove(sid);
                L.run.forget(sid);
                L.by_solve.erase(sid);
            }
            std::lock_guard<std::mutex> lk(mu);
            busy_s += b - a;
            st.engine_steps += 1;
            for (const Done& d : done) {
                Req& r = reqs[d.rid];
                if (d.state == HMHD_REQ_FAILED) err = L.eng.err;
                r.res.steps = d.steps;
                r.res.t = d.t;
                r.res.zone_updates = d.zone_updates;
                finish_locked(r, d.state);
                release_locked(L, r);
                st.completed += 1;
                st.zone_updates += r.res.zone_updates;
            }
            if (!done.empty()) {
                admit_locked();
                cv_done.notify_all();
            }
        }
    }
};

extern "C" {

// ------------------------------------------------------------------ //
// single run

int hmhd_create(const hmhd_config* cfg, hmhd_t** out) {
    if (!cfg || !out) return fail(g_create_err, HMHD_ERR_ARG, "hmhd_create: null argument");
    *out = nullptr;
    hmhd_t* h = new hmhd_handle;
    h->id = h->eng.add(*cfg);
    if (h->id < 0) {
        g_create_err = h->eng.err;
        delete h;
        return create_code(g_create_err);
    }
    *out = h;
    return HMHD_OK;
}

void hmhd_destroy(hmhd_t* h) { delete h; }

const char* hmhd_last_error(const hmhd_t* h) { return h ? h->err.c_str() : g_create_err.c_str(); }

int hmhd_dims(const hmhd_t* h, int32_t* ngc, int32_t* mx, int32_t* my) {
    if (!h) return HMHD_ERR_ARG;
    const hmhd::InstHost* in = h->eng.find(h->id);
    if (ngc) *ngc = in->ngc;
    if (mx) *mx = in->mx;
    if (my) *my = in->my;
    return HMHD_OK;
}

int hmhd_gp_compute(const hmhd_config* cfg, double* predvec, double* quadw, int32_t* stenc) {
    if (!cfg) return fail(g_create_err, HMHD_ERR_ARG, "hmhd_gp_compute: null config");
    if (hmhd::Engine::check_config(*cfg, g_create_err) != HMHD_OK) return HMHD_ERR_ARG;
    double pv[hmhd::kPvSet], qw[(hmhd::kMaxR + 1) * 4];
    if (hmhd::gp_compute(cfg->R, cfg->ell_over_delta, cfg->xbeg, cfg->xend, cfg->ybeg, cfg->yend, cfg->nx, cfg->ny,
                         pv, qw) != 0)
        return fail(g_create_err, HMHD_ERR_GP,
                    "GP covariance matrix is not positive definite (src/linAlgQuadPrecision.f90:136)");
    export_gp(cfg->R, pv, qw, predvec, quadw, stenc);
    return HMHD_OK;
}

int hmhd_get_gp(const hmhd_t* h, double* predvec, double* quadw, int32_t* stenc) {
    if (!h) return HMHD_ERR_ARG;
    const hmhd::InstHost* in = h->eng.find(h->id);
    export_gp(in->cfg.R, in->pv, in->qw, predvec, quadw, stenc);
    return HMHD_OK;
}

int hmhd_set_gp(hmhd_t* h, const double* predvec, const double* quadw) {
    if (!h) return HMHD_ERR_ARG;
    if (!predvec || !quadw) return fail(h->err, HMHD_ERR_ARG, "hmhd_set_gp: null argument");
    const hmhd::InstHost* in = h->eng.find(h->id);
    const int R = in->cfg.R, nsm = hmhd::gp_ns(R), nqm = hmhd::gp_nq(R);
    double pv[hmhd::kPvSet] = {0};
    double qw[(hmhd::kMaxR + 1) * 4] = {0};
    for (int rr = 1; rr <= R; rr++) {
        for (int f = 0; f < 4; f++)
            for (int q = 0; q < hmhd::gp_nq(rr); q++)
                for (int k = 0; k < hmhd::gp_ns(rr); k++)
                    pv[hmhd::gp_pv_off(rr) + (f * hmhd::gp_nq(rr) + q) * hmhd::gp_ns(rr) + k] =
                        predvec[(((size_t)(rr - 1) * 4 + f) * nqm + q) * nsm + k];
        for (int q = 0; q < hmhd::gp_nq(rr); q++) qw[rr * 4 + q] = quadw[(rr - 1) * nqm + q];
    }
    const int rc = h->eng.set_gp(h->id, pv, qw);
    if (rc) h->err = h->eng.err;
    return rc;
}

int hmhd_upload_state(hmhd_t* h, const double* U, const double* V) {
    if (!h) return HMHD_ERR_ARG;
    if (!U) return fail(h->err, HMHD_ERR_ARG, "hmhd_upload_state: U is required");
    const int rc = h->eng.upload(h->id, U, V);
    if (rc) h->err = h->eng.err;
    return rc;
}

int hmhd_cfl_dt(hmhd_t* h, double* dt) {
    if (!h) return HMHD_ERR_ARG;
    if (!dt) return fail(h->err, HMHD_ERR_ARG, "hmhd_cfl_dt: null argument");
    std::vector<double> d;
    const int rc = h->eng.cfl(d);
    if (rc) h->err = h->eng.err;
    else *dt = d[0];
    return rc;
}

int hmhd_advance(hmhd_t* h, double dt) {
    if (!h) return HMHD_ERR_ARG;
    const int rc = h->eng.advance(std::vector<double>{dt});
    if (rc) h->err = h->eng.err;
    return rc;
}

int hmhd_download(hmhd_t* h, double* U, double* V, int32_t* scheme) {
    if (!h) return HMHD_ERR_ARG;
    const int rc = h->eng.download(h->id, U, V, scheme);
    if (rc) h->err = h->eng.err;
    return rc;
}

int hmhd_stats(const hmhd_t* h, int32_t* passes, int32_t* inner) {
    if (!h) return HMHD_ERR_ARG;
    const hmhd::InstHost* in = h->eng.find(h->id);
    for (int s = 0; s < 3; s++) {
        if (passes) passes[s] = in->passes[s];
        if (inner) inner[s] = in->inner[s];
    }
    return HMHD_OK;
}

int hmhd_record_stage_schemes(hmhd_t* h, int32_t on) {
    if (!h) return HMHD_ERR_ARG;
    h->eng.record_schemes = on != 0;
    return HMHD_OK;
}

int hmhd_stage_schemes(const hmhd_t* h, int32_t* out) {
    if (!h || !out) return HMHD_ERR_ARG;
    const auto& v = h->eng.find(h->id)->stage_schemes;
    if (v.empty()) return HMHD_ERR_ARG;
    std::memcpy(out, v.data(), sizeof(int32_t) * v.size());
    return HMHD_OK;
}

int hmhd_profile(hmhd_t* h, int32_t on) {
    if (!h) return HMHD_ERR_ARG;
    h->eng.prof_mask = (uint32_t)on & ((1u << hmhd::PF_N) - 1u);
    return HMHD_OK;
}

int hmhd_profile_read(hmhd_t* h, double* ms, int64_t* launches, int32_t reset) {
    if (!h) return HMHD_ERR_ARG;
    return profile_read(h->eng, ms, launches, reset);
}

int hmhd_synchronize(hmhd_t* h) {
    if (!h) return HMHD_ERR_ARG;
    const int rc = h->eng.sync();
    if (rc) h->err = h->eng.err;
    return rc;
}

int hmhd_face_tuning(int32_t mode) {
    if (mode < 0 || mode > 1 + HMHD_FACE_FORMS) return fail(g_create_err, HMHD_ERR_ARG, "hmhd_face_tuning: bad mode");
    hmhd::face_tuning_set_mode(mode);
    return HMHD_OK;
}

int hmhd_face_tuning_read(hmhd_face_tune* out, int32_t max, int32_t* n) {
    const std::vector<hmhd::FaceTuneEntry> t = hmhd::face_tuning_table();
    if (n) *n = (int32_t)t.size();
    for (int32_t k = 0; out && k < max && k < (int32_t)t.size(); k++) {
        hmhd_face_tune& o = out[k];
        o.R = t[k].R;
        o.size_class = t[k].size_class;
        o.form = t[k].variant;
        o.tile_x = hmhd::kFaceVariants[t[k].variant].tx;
        o.tile_y = hmhd::kFaceVariants[t[k].variant].ty;
        o.waves = hmhd::kFaceVariants[t[k].variant].waves;
        for (int v = 0; v < 8; v++) o.ms[v] = t[k].ms[v];
    }
    return HMHD_OK;
}

int hmhd_face_tuning_clear(void) {
    hmhd::face_tuning_clear();
    return HMHD_OK;
}

// ------------------------------------------------------------------ //
// batch

int hmhd_batch_create(hmhd_batch_t** out) {
    if (!out) return fail(g_create_err, HMHD_ERR_ARG, "hmhd_batch_create: null argument");
    *out = new hmhd_batch_handle;
    return HMHD_OK;
}

void hmhd_batch_destroy(hmhd_batch_t* b) { delete b; }

const char* hmhd_batch_last_error(const hmhd_batch_t* b) { return b ? b->err.c_str() : g_create_err.c_str(); }

int hmhd_batch_add(hmhd_batch_t* b, const hmhd_config* cfg, const double* U, const double* V, double tmax,
                   int32_t nstepmax, int64_t* id) {
    if (!b) return HMHD_ERR_ARG;
    if (!cfg || !U || !id) return fail(b->err, HMHD_ERR_ARG, "hmhd_batch_add: null argument");
    const int64_t sid = b->eng.add(*cfg);
    if (sid < 0) return fail(b->err, create_code(b->eng.err), b->eng.err);
    const int rc = b->eng.upload(sid, U, V);
    if (rc) {
        b->err = b->eng.err;
        (void)b->eng.remove(sid);
        return rc;
    }
    b->run.track(sid, tmax, nstepmax);
    *id = sid;
    return HMHD_OK;
}

int hmhd_batch_run(hmhd_batch_t* b, int32_t max_steps) {
    if (!b) return HMHD_ERR_ARG;
    std::vector<int64_t> fin;
    for (int32_t s = 0; max_steps <= 0 || s < max_steps; s++) {
        if (b->eng.active_ids().empty()) break;
        const int rc = b->run.step(fin);
        if (rc) {
            b->err = b->eng.err;
            return rc;
        }
    }
    return HMHD_OK;
}

int hmhd_batch_status(const hmhd_batch_t* b, int64_t id, double* t, int32_t* steps, int32_t* state) {
    if (!b) return HMHD_ERR_ARG;
    auto it = b->run.info.find(id);
    if (it == b->run.info.end()) return HMHD_ERR_ARG;
    if (t) *t = it->second.t;
    if (steps) *steps = it->second.steps;
    if (state) *state = it->second.state;
    return HMHD_OK;
}

int hmhd_batch_dims(const hmhd_batch_t* b, int64_t id, int32_t* ngc, int32_t* mx, int32_t* my) {
    if (!b) return HMHD_ERR_ARG;
    const hmhd::InstHost* in = b->eng.find(id);
    if (!in) return HMHD_ERR_ARG;
    if (ngc) *ngc = in->ngc;
    if (mx) *mx = in->mx;
    if (my) *my = in->my;
    return HMHD_OK;
}

int hmhd_batch_download(hmhd_batch_t* b, int64_t id, double* U, double* V, int32_t* scheme) {
    if (!b) return HMHD_ERR_ARG;
    const int rc = b->eng.download(id, U, V, scheme);
    if (rc) b->err = b->eng.err;
    return rc;
}

int hmhd_batch_profile(hmhd_batch_t* b, int32_t on) {
    if (!b) return HMHD_ERR_ARG;
    b->eng.prof_mask = (uint32_t)on & ((1u << hmhd::PF_N) - 1u);
    return HMHD_OK;
}

int hmhd_batch_profile_read(hmhd_batch_t* b, double* ms, int64_t* launches, int32_t reset) {
    if (!b) return HMHD_ERR_ARG;
    return profile_read(b->eng, ms, launches, reset);
}

// ------------------------------------------------------------------ //
// server

uint64_t hmhd_solve_footprint(const hmhd_config* cfg) {
    std::string e;
    if (!cfg || hmhd::Engine::check_config(*cfg, e) != HMHD_OK) return 0;
    return hmhd::Engine::footprint(*cfg);
}

int hmhd_server_create(uint64_t hbm_budget_bytes, int32_t max_active, hmhd_server_t** out) {
    if (!out || hbm_budget_bytes == 0)
        return fail(g_create_err, HMHD_ERR_ARG, "hmhd_server_create: null handle or zero budget");
    hmhd_server_t* s = new hmhd_server_handle;
    s->budget = hbm_budget_bytes;
    s->max_active = max_active;
    for (int k = 0; k <= hmhd_server_handle::kSoloLanes; k++) {
        s->lanes.push_back(std::make_unique<hmhd_server_handle::Lane>());
        s->lanes.back()->solo = k > 0;
    }
    for (auto& l : s->lanes) {
        hmhd_server_handle::Lane* lp = l.get();
        l->th = std::thread([s, lp] { s->lane_loop(*lp); });
    }
    *out = s;
    return HMHD_OK;
}

void hmhd_server_destroy(hmhd_server_t* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (auto& l : s->lanes) l->cv.notify_all();
    }
    for (auto& l : s->lanes)
        if (l->th.joinable()) l->th.join();
    delete s;
}

const char* hmhd_server_last_error(const hmhd_server_t* s) { return s ? s->err.c_str() : g_create_err.c_str(); }

int hmhd_server_submit(hmhd_server_t* s, const hmhd_config* cfg, const double* U, const double* V, double tmax,
                       int32_t nstepmax, int64_t* req) {
    if (!s) return HMHD_ERR_ARG;
    std::string e;
    if (!cfg || !U || !req) e = "hmhd_server_submit: null argument";
    else (void)hmhd::Engine::check_config(*cfg, e);
    if (!e.empty()) {
        std::lock_guard<std::mutex> lk(s->mu);
        return fail(s->err, HMHD_ERR_ARG, e);
    }
    hmhd_server_handle::Req r;
    r.cfg = *cfg;
    const int ngc = cfg->R + 2;
    const size_t np = (size_t)(cfg->nx + 2 * ngc) * (cfg->ny + 2 * ngc);
    r.U.assign(U, U + 8 * np);
    if (V) r.V.assign(V, V + 10 * np);
    r.tmax = tmax;
    r.nstepmax = nstepmax;
    r.bytes = hmhd::Engine::footprint(*cfg);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r.id = s->next_req++;
        r.res.state = HMHD_REQ_QUEUED;
        r.res.submit_s = s->now();
        *req = r.id;
        s->queue.push_back(r.id);
        s->reqs.emplace(r.id, std::move(r));
        s->st.submitted += 1;
        s->admit_locked();
    }
    return HMHD_OK;
}

int hmhd_server_poll(hmhd_server_t* s, int64_t req, hmhd_result* res) {
    if (!s) return HMHD_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(req);
    if (it == s->reqs.end()) return fail(s->err, HMHD_ERR_ARG, "unknown request");
    if (res) *res = it->second.res;
    return HMHD_OK;
}

int hmhd_server_wait(hmhd_server_t* s, int64_t req, double timeout_s, hmhd_result* res, double* U, double* V,
                     int32_t* scheme) {
    if (!s) return HMHD_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(req);
    if (it == s->reqs.end()) return fail(s->err, HMHD_ERR_ARG, "unknown request");
    hmhd_server_handle::Req& r = it->second;  // stays valid: only this call erases it
    auto ready = [&] { return r.done; };
    if (timeout_s > 0) {
        if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
            return fail(s->err, HMHD_ERR_TIMEOUT, "request not finished within the timeout");
    } else {
        s->cv_done.wait(lk, ready);
    }
    if (res) *res = r.res;
    if (U && !r.Uout.empty()) std::memcpy(U, r.Uout.data(), sizeof(double) * r.Uout.size());
    if (V && !r.Vout.empty()) std::memcpy(V, r.Vout.data(), sizeof(double) * r.Vout.size());
    if (scheme && !r.Sout.empty()) std::memcpy(scheme, r.Sout.data(), sizeof(int32_t) * r.Sout.size());
    s->reqs.erase(req);
    return HMHD_OK;
}

int hmhd_server_stats_read(hmhd_server_t* s, hmhd_server_stats* out) {
    if (!s || !out) return HMHD_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->busy_s = s->busy_s;
    out->elapsed_s = s->now();
    return HMHD_OK;
}

}  // extern "C"
