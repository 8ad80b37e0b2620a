// hmhd_runner.hpp -- the driver loop of src/mhd_driver.f90:76-95 for many
// solves at once: every engine step computes each active solve's CFL dt,
// applies validTimeStep (src/mhd_driver.f90:117-144) per solve, and advances
// all of them together; a solve leaves the active set when it reaches its
// tmax or step limit, or aborts on a negative or NaN dt.
#pragma once

#include <cstdint>
#include <unordered_map>
#include <vector>

#include "hmhd_solver.hpp"

namespace hmhd {

enum RunState : int32_t { RUN_ACTIVE = 0, RUN_DONE = 1, RUN_ABORTED = 2 };

struct RunInfo {
    double t = 0.0, tmax = 0.0;
    int32_t steps = 0, nstepmax = 0;
    int32_t state = RUN_ACTIVE;
    int64_t zone_updates = 0;
    const char* why = "";
};

class Runner {
   public:
    explicit Runner(Engine& e) : eng(e) {}
    void track(int64_t id, double tmax, int32_t nstepmax) {
        RunInfo r;
        r.tmax = tmax;
        r.nstepmax = nstepmax > 0 ? nstepmax : 1000000;
        info[id] = r;
    }
    void forget(int64_t id) { info.erase(id); }

    // One engine step of every active solve; ids that left the active set
    // in this step are appended to `finished`.
    int step(std::vector<int64_t>& finished) {
        std::vector<double> dt;
        int rc = eng.cfl(dt);
        if (rc) return rc;
        const std::vector<int64_t> ids = eng.active_ids();
        std::vector<int64_t> go;
        std::vector<double> go_dt;
        for (size_t k = 0; k < ids.size(); k++) {
            RunInfo& r = info[ids[k]];
            double d = dt[k];
            if (r.t + d > r.tmax) {
                d = r.tmax - r.t;
            } else if (d < 0.0 || d != d || d - d != 0.0) {
                r.state = RUN_ABORTED;
                r.why = d < 0.0 ? "dt is negative" : d != d ? "dt is NAN" : "dt is infinite";
                finished.push_back(ids[k]);
                continue;
            }
            go.push_back(ids[k]);
            go_dt.push_back(d);
        }
        for (int64_t id : finished)
            if ((rc = eng.set_active(id, false))) return rc;
        if (!go.empty()) {
            if ((rc = eng.advance(go_dt))) return rc;
            for (size_t k = 0; k < go.size(); k++) {
                RunInfo& r = info[go[k]];
                const InstHost* h = eng.find(go[k]);
                r.t += go_dt[k];
                r.steps += 1;
                r.zone_updates += (int64_t)h->cfg.nx * h->cfg.ny;
                if (!(r.t < r.tmax) || r.steps >= r.nstepmax) {
                    r.state = RUN_DONE;
                    finished.push_back(go[k]);
                    if ((rc = eng.set_active(go[k], false))) return rc;
                }
            }
        }
        return HMHD_OK;
    }

    Engine& eng;
    std::unordered_map<int64_t, RunInfo> info;
};

}  // namespace hmhd
