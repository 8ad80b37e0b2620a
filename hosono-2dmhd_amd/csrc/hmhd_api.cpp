// hmhd_api.cpp -- the extern "C" entry points declared in include/hmhd.h:
// a single-run handle is an Engine holding one instance.
#include <cstring>
#include <string>

#include "../../include/hmhd.h"
#include "gp_tables.hpp"
#include "hmhd_solver.hpp"

struct hmhd_handle {
    hmhd::Engine eng;
    std::string err;
};

namespace {
thread_local std::string g_create_err;

int fail(hmhd_t* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    return code;
}

int pass(hmhd_t* h, int code) {
    if (code != HMHD_OK) h->err = h->eng.err;
    return code;
}
}  // namespace

extern "C" {

int hmhd_create(const hmhd_config* cfg, hmhd_t** out) {
    if (!cfg || !out) {
        g_create_err = "hmhd_create: null argument";
        return HMHD_ERR_ARG;
    }
    *out = nullptr;
    hmhd_t* h = new hmhd_handle;
    if (h->eng.add_instance(*cfg) < 0) {
        g_create_err = h->eng.err;
        const int code = g_create_err.find("positive definite") != std::string::npos ? HMHD_ERR_GP : HMHD_ERR_ARG;
        delete h;
        return code;
    }
    int rc = h->eng.finalize();
    if (rc != HMHD_OK) {
        g_create_err = h->eng.err;
        delete h;
        return rc;
    }
    *out = h;
    return HMHD_OK;
}

void hmhd_destroy(hmhd_t* h) { delete h; }

const char* hmhd_last_error(const hmhd_t* h) { return h ? h->err.c_str() : g_create_err.c_str(); }

int hmhd_dims(const hmhd_t* h, int32_t* ngc, int32_t* mx, int32_t* my) {
    if (!h) return HMHD_ERR_ARG;
    const auto& in = h->eng.inst[0];
    if (ngc) *ngc = in.ngc;
    if (mx) *mx = in.mx;
    if (my) *my = in.my;
    return HMHD_OK;
}

}  // extern "C"

namespace {
// device table layout -> the reference's GP_predictionVectors(nsMax, nqMax,
// 4, R), GP_quadratureWeights(nqMax, R), GP_stencIdxs(2, nsMax, R)
void export_gp(int R, const double* pv_dev, const double* qw_dev, double* predvec, double* quadw, int32_t* stenc) {
    const int nsm = hmhd::gp_ns(R), nqm = hmhd::gp_nq(R);
    const double* pv = pv_dev;
    const double* qw = qw_dev;
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
}  // namespace

extern "C" {

int hmhd_gp_compute(const hmhd_config* cfg, double* predvec, double* quadw, int32_t* stenc) {
    if (!cfg) {
        g_create_err = "hmhd_gp_compute: null config";
        return HMHD_ERR_ARG;
    }
    std::string err;
    if (hmhd::Engine::check_config(*cfg, err) != HMHD_OK) {
        g_create_err = err;
        return HMHD_ERR_ARG;
    }
    double pv[hmhd::kPvSet], qw[(hmhd::kMaxR + 1) * 4];
    if (hmhd::gp_compute(cfg->R, cfg->ell_over_delta, cfg->xbeg, cfg->xend, cfg->ybeg, cfg->yend, cfg->nx, cfg->ny,
                         pv, qw) != 0) {
        g_create_err = "GP covariance matrix is not positive definite (src/linAlgQuadPrecision.f90:136)";
        return HMHD_ERR_GP;
    }
    export_gp(cfg->R, pv, qw, predvec, quadw, stenc);
    return HMHD_OK;
}

int hmhd_get_gp(const hmhd_t* h, double* predvec, double* quadw, int32_t* stenc) {
    if (!h) return HMHD_ERR_ARG;
    const auto& in = h->eng.inst[0];
    export_gp(in.cfg.R, in.pv, in.qw, predvec, quadw, stenc);
    return HMHD_OK;
}

int hmhd_set_gp(hmhd_t* h, const double* predvec, const double* quadw) {
    if (!h || !predvec || !quadw) return fail(h, HMHD_ERR_ARG, "hmhd_set_gp: null argument");
    const auto& in = h->eng.inst[0];
    const int R = in.cfg.R, nsm = hmhd::gp_ns(R), nqm = hmhd::gp_nq(R);
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
    return pass(h, h->eng.set_gp(0, pv, qw));
}

int hmhd_upload_state(hmhd_t* h, const double* U, const double* V) {
    if (!h || !U) return fail(h, HMHD_ERR_ARG, "hmhd_upload_state: U is required");
    return pass(h, h->eng.upload(0, U, V));
}

int hmhd_cfl_dt(hmhd_t* h, double* dt) {
    if (!h || !dt) return fail(h, HMHD_ERR_ARG, "hmhd_cfl_dt: null argument");
    std::vector<double> d;
    int rc = h->eng.cfl(d);
    if (rc == HMHD_OK) *dt = d[0];
    return pass(h, rc);
}

int hmhd_advance(hmhd_t* h, double dt) {
    if (!h) return HMHD_ERR_ARG;
    return pass(h, h->eng.advance(&dt));
}

int hmhd_download(hmhd_t* h, double* U, double* V, int32_t* scheme) {
    if (!h) return HMHD_ERR_ARG;
    return pass(h, h->eng.download(0, U, V, scheme));
}

int hmhd_stats(const hmhd_t* h, int32_t* passes, int32_t* inner) {
    if (!h) return HMHD_ERR_ARG;
    const int32_t* c = h->eng.ctl_host.data();
    for (int s = 0; s < 3; s++) {
        if (passes) passes[s] = c[hmhd::CTL_PASSES + s];
        if (inner) inner[s] = c[hmhd::CTL_INNER + s];
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
    const auto& v = h->eng.stage_schemes[0];
    if (v.empty()) return HMHD_ERR_ARG;
    std::memcpy(out, v.data(), sizeof(int32_t) * v.size());
    return HMHD_OK;
}

int hmhd_profile(hmhd_t* h, int32_t on) {
    if (!h) return HMHD_ERR_ARG;
    h->eng.prof_on = on != 0;
    return HMHD_OK;
}

int hmhd_profile_read(hmhd_t* h, double* ms, int64_t* launches, int32_t reset) {
    if (!h) return HMHD_ERR_ARG;
    for (int f = 0; f < hmhd::PF_N; f++) {
        if (ms) ms[f] = h->eng.prof_ms[f];
        if (launches) launches[f] = h->eng.prof_n[f];
        if (reset) {
            h->eng.prof_ms[f] = 0.0;
            h->eng.prof_n[f] = 0;
        }
    }
    return HMHD_OK;
}

int hmhd_synchronize(hmhd_t* h) {
    if (!h) return HMHD_ERR_ARG;
    return pass(h, h->eng.sync());
}

}  // extern "C"
