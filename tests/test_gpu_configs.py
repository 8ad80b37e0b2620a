"""The benchmark configurations at their own sizes and draws.

* MHD blast 2048^2 (outflow, rho = 1, p = 10 for r < 0.1 else 0.1,
  B = (1, 1, 0)/sqrt 2), the case that drives the MOOD cascade to first
  order (src/detect.f90:88) and exercises FLICD (src/FLICDdivB.f90:13-70)
  at full size: its first two steps bit for bit against the oracle, a GPU
  run to t_end = 0.2, then two more steps bit for bit against the oracle
  started from the GPU's state.
* The batched workload (hosono_2dmhd_amd.workloads.batched_runs, seed 2024,
  as bench.py draws it: sides 64..384, Orszag-Tang / Sedov, 1% noise): its
  first 32 runs advanced together for 20 steps, each against the oracle's
  driver loop (tests/golden/batched_seed2024.json).
* The served workload (workloads.served_requests, seed 2025: Poisson
  arrivals, sides 76..1008, Orszag-Tang / blast / Sedov): its first 12
  requests through a server, each capped at 5 steps, against the oracle
  (tests/golden/served_seed2025.json).
"""
import hashlib
import json
import pathlib
import threading
import time

import numpy as np
import pytest

import cases
import hosono_2dmhd_amd.runtime as rt
import hosono_2dmhd_amd.workloads as wl
import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
BLAST_T_END = 0.2


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def input_sha(U, V):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(U).tobytes())
    h.update(np.ascontiguousarray(V).tobytes())
    return h.hexdigest()


def _step_both(orc, gpu):
    d_o = orc.cfl_dt()
    d_g = gpu.cfl_computedt()
    assert d_o == d_g, f"dt oracle {d_o!r} gpu {d_g!r}"
    orc.advance(d_o)
    gpu.record_stage_schemes(True)
    gpu.advance_solution(d_g)


def _same(orc, gpu, where):
    Uo, Vo = orc.get_state()
    Ug, Vg, Sg = gpu.download()
    if not np.array_equal(Uo, Ug):
        d = np.abs(Uo - Ug)
        pytest.fail(f"{where}: U differs, max |diff| {d.max():.3e}, {int((d > 0).sum())} values")
    assert np.array_equal(Vo, Vg), f"{where}: V differs"
    assert np.array_equal(orc.stage_schemes(), gpu.stage_schemes()), f"{where}: stage scheme masks differ"
    assert all(np.array_equal(a, b) for a, b in zip(orc.stats(), gpu.stats())), f"{where}: MOOD counts differ"


@pytest.mark.timeout(900)
def test_blast2048_first_steps_then_t_end_then_two_more_steps():
    cfg, g, U, V = cases.make("blast", n=2048)
    orc = O.Oracle(cfg)
    orc.set_state(U, V)
    gpu = rt.Solver(cases.runtime_config(cfg))
    gpu.upload(U, V)
    for k in range(2):
        _step_both(orc, gpu)
        _same(orc, gpu, f"blast 2048^2 step {k + 1}")
    del orc
    gpu.record_stage_schemes(False)
    t = 0.0
    # the two steps above restart the clock: the run to t_end is the driver loop from here
    steps, passes, t0 = 0, 0, time.perf_counter()
    while t < BLAST_T_END:
        dt = rt.valid_time_step(gpu.cfl_computedt(), t, BLAST_T_END)
        gpu.advance_solution(dt)
        t += dt
        steps += 1
        passes += int(gpu.stats()[0].sum())
    el = time.perf_counter() - t0
    Ug, Vg, Sg = gpu.download()
    ngc = cfg.R + 2
    inner = Vg[ngc:-ngc, ngc:-ngc, :]
    print(f"blast 2048^2: {steps} steps to t = {t} in {el:.1f} s "
          f"({steps * 2048 * 2048 / el / 1e6:.1f} M zone-updates/s), {passes / (3 * steps):.2f} passes per stage")
    assert np.isfinite(Ug).all() and np.isfinite(Vg).all()
    assert (inner[:, :, 0] > 0).all() and (inner[:, :, 4] > 0).all(), "non-positive density or pressure"
    first_order = int((Sg[ngc:-ngc, ngc:-ngc] == 0).sum())
    assert first_order > 0, "no first-order cells at t_end"
    orc = O.Oracle(cfg)
    orc.set_state(Ug, Vg)
    late_passes = 0
    for k in range(2):
        _step_both(orc, gpu)
        _same(orc, gpu, f"blast 2048^2 from t={t}, step {k + 1}")
        late_passes += int(gpu.stats()[0].sum())
    assert late_passes > 6, "expected MOOD passes beyond the first late in the run"
    print(f"first-order cells at t_end: {first_order}; passes in the two late steps: {late_passes}")
    gpu.close()


def _mix(kind, seed):
    path = GOLDEN / f"{kind}_seed{seed}.json"
    if not path.exists():
        pytest.skip(f"{path.name} not generated (tests/golden/make_golden.py {kind} ...)")
    return json.loads(path.read_text())


def _check_run(e, r, t, steps, U, V, where):
    assert input_sha(r.U, r.V) == e["input_sha256"], f"{where}: initial state differs from the fixture's"
    assert steps == e["steps"], f"{where}: {steps} steps, oracle {e['steps']}"
    assert t == e["t"], f"{where}: t {t!r}, oracle {e['t']!r}"
    assert sha(U) == e["U_sha256"], f"{where}: final U differs from the oracle's"
    assert sha(V) == e["V_sha256"], f"{where}: final V differs from the oracle's"


def test_batched_workload_against_the_oracle():
    fx = _mix("batched", 2024)
    runs = wl.batched_runs(fx["seed"], fx["count"])
    with rt.Batch() as b:
        ids = [b.add(r.cfg, r.U, r.V, r.tmax, fx["nstepmax"]) for r in runs]
        b.run(0)
        for k, (r, i, e) in enumerate(zip(runs, ids, fx["runs"])):
            assert (r.kind, r.n) == (e["kind"], e["n"])
            t, steps, state = b.status(i)
            assert state == rt.RUN_DONE
            U, V, _ = b.download(i)
            _check_run(e, r, t, steps, U, V, f"batched run {k} ({r.kind} {r.n}^2)")


def test_served_workload_against_the_oracle():
    fx = _mix("served", 2025)
    reqs = wl.served_requests(fx["seed"], fx["count"], 12.0)
    ids = [None] * len(reqs)
    with rt.Server(8 << 30) as srv:
        def feeder():
            t0 = time.perf_counter()
            for k, r in enumerate(reqs):
                # the Poisson arrival times, compressed 20x so the requests overlap
                delay = 0.05 * r.arrival_s - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                ids[k] = srv.submit(r.cfg, r.U, r.V, r.tmax, fx["nstepmax"])
        th = threading.Thread(target=feeder)
        th.start()
        th.join()
        res = [srv.wait(i, timeout_s=300.0) for i in ids]
        st = srv.stats()
    assert st["peak_active"] >= 2, "the requests never overlapped"
    for k, (r, x, e) in enumerate(zip(reqs, res, fx["runs"])):
        assert (r.kind, r.n) == (e["kind"], e["n"])
        assert x.state == rt.REQ_DONE
        _check_run(e, r, x.t, x.steps, x.U, x.V, f"request {k} ({r.kind} {r.n}^2)")
