#!/usr/bin/env python3
"""FP64 zone-updates/s of the hosono-2DMHD step on one MI355X.

Workload: the Orszag-Tang vortex on a 1024 x 1024 periodic grid (rho =
gamma^2, p = gamma, v = (-sin 2 pi y, sin 2 pi x), B = (-sin 2 pi y,
sin 4 pi x)/sqrt(4 pi), gamma = 5/3), GP-MOOD radius 2 with l = 12 dx, HLLC,
SSP-RK3, CFL 0.8, one instance, state resident in HBM.  A step is one
driver iteration: the CFL reduction (cfl_computedt) and one full SSP-RK3
step with every MOOD pass (timeStep_advanceSolution).  A zone-update is one
interior cell advanced one step.

The solve runs from t = 0 to t = 0.5 and is reported whole as
`full_solve`.  With --steps K the line's value is exactly K of its N
steps, one from each K-th of the solve (by step index), the one whose MOOD
pass count is nearest that stretch's mean, timed in a second solve from
the same initial state, each bracketed by stream synchronizations: the
MOOD loop's depth varies over the solve (1 pass per stage early, about 8
near t = 0.2, 20.5 on average, with large step-to-step swings), so neither
a contiguous window nor a handful of evenly spaced steps stands for the
configuration.  With no --steps the value is the whole solve.
Warm-up steps run on a separate copy before the timed region.

One JSON line on stdout.  `roofline` covers the dominant kernel, the face
kernel in its first-pass form (GP reconstruction + admissibility + HLLC),
timed with HIP events on the solver's stream over the whole solve.
`cpu_baseline` times the reference's own Fortran step (compiled from its
sources by oracle/Makefile, OpenMP on the host cores, then one thread) from
four of the GPU solve's states (t = 0.0625, 0.1875, 0.3125, 0.4375), in a
child process started before the GPU is touched and idle until the GPU leg
has finished.

Beside the plain form the line carries three more objects (--forms):
`batched`, 256 seeded runs of mixed sizes advanced together in one engine,
against the same runs launched one after another on their own handles;
`served`, Poisson arrivals of mixed requests into a server with an HBM
budget, with p50/p99 submit-to-result latency and aggregate zone-updates/s;
and `blast`, the MHD blast at 2048^2 from t = 0 to t = 0.2 (outflow, the
MOOD cascade down to first order), one solve.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import math
import os
import pathlib
import subprocess
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
PKG_DIR = ROOT / "hosono-2dmhd_amd"

METRIC = "FP64 zone-updates/s at 1024^2 (plain vs batched); served p50/p99 time-to-solution"
UNIT = "zone-updates/s"

# MI355X vendor peaks
PEAK_FP64_TFLOPS = 78.6  # FP64 vector == FP64 matrix on CDNA4
PEAK_HBM_GBS = 8000.0

# Algorithmic FP64 flops come from hmhd_flop_count (built with the library):
# the kernels' own formulas compiled on the host with a counting scalar
# (+ - * / sqrt = 1; min/max/abs/compare = 0), run over the solve's states.
FLOP_COUNT_EXE = PKG_DIR / "hmhd_flop_count"
# compulsory HBM bytes of the face kernel per interior cell: read U (8 FP64),
# write the raw x and y fluxes (16 FP64), the four inadmissible-side flags
# and the two last-flux radii (one byte each)
BYTES_FACE_KERNEL = 8 * 8 + 16 * 8 + 6


def count_flops(cfg, states, windows: int = 64) -> dict | None:
    """Instrumented FP64 operation counts of the first MOOD pass over each
    padded state U (my, mx, 8) in `states` (hmhd_flop_count: face and update
    work averaged over `windows` seeded 32 x 32 windows, the face kernel's
    domain outside the interior in full), averaged over the states."""
    import numpy as np
    import tempfile
    if not FLOP_COUNT_EXE.exists():
        return None
    outs = []
    with tempfile.TemporaryDirectory() as d:
        for k, U in enumerate(states):
            path = pathlib.Path(d) / f"U{k}.bin"
            np.ascontiguousarray(U, dtype=np.float64).tofile(path)
            args = [str(FLOP_COUNT_EXE), str(path), str(cfg.nx), str(cfg.ny), str(cfg.R), repr(cfg.ell_over_delta),
                    repr(cfg.gamma), repr(cfg.small_p), repr(cfg.small_rho), str(int(cfg.force_hydro)), str(cfg.bc),
                    repr(cfg.xbeg), repr(cfg.xend), repr(cfg.ybeg), repr(cfg.yend), repr(cfg.cfl), str(windows),
                    str(k + 1)]
            r = subprocess.run(args, capture_output=True, text=True, timeout=300)
            if r.returncode != 0:
                return {"error": r.stderr.strip()[-300:]}
            outs.append(json.loads(r.stdout))
    mean = lambda key: sum(o[key] for o in outs) / len(outs)
    per_zone = {k: sum(o["per_zone"][k] for o in outs) / len(outs) for k in outs[0]["per_zone"]}
    return {"flops_per_zone_update": mean("flops_per_zone_update"),
            "face_flops_per_launch": mean("face_flops_per_launch"),
            "hllc_flops_per_call": mean("hllc_flops_per_call"), "per_zone": per_zone,
            "states": len(outs), "windows_per_state": windows,
            "how": "hmhd_flop_count: the kernels' formulas (hmhd_math.hpp) on the host with a counting FP64 "
                   "type; first MOOD pass of a stage (every cell at radius R), x3 stages, plus the RK "
                   "combinations and the CFL pass; + - * / sqrt count 1, min/max/abs/compare 0"}


def load_package():
    if "hosono_2dmhd_amd" in sys.modules:
        return sys.modules["hosono_2dmhd_amd"]
    spec = importlib.util.spec_from_file_location(
        "hosono_2dmhd_amd", PKG_DIR / "__init__.py", submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hosono_2dmhd_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def workload(n: int):
    load_package()
    import hosono_2dmhd_amd.ics as ics
    g = ics.Grid(n, n, 0.0, 1.0, 0.0, 1.0, R=2)
    U, V = ics.orszag_tang(g)
    return g, U, V


def _cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_child(n: int, max_steps: int, max_seconds: float) -> None:
    """Runs in a child process started before the GPU is touched: waits on
    stdin for a JSON list of saved states (GPU states at sampled steps of
    the timed solve), then times the reference's Fortran step from each
    (max_steps steps in all, split evenly, or until max_seconds) with every
    host core, and one step with one thread from the middle state."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import ctypes
    import numpy as np
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    cfg = O.Config(n, n, R=2, ell_over_delta=12.0)
    line = sys.stdin.readline().strip()
    paths = json.loads(line) if line else []
    if not paths:
        return
    per = max(1, max_steps // len(paths))
    ts, steps, el = [], 0, 0.0
    el1 = None
    if O.ref_available("fast"):
        kind = "reference"
        # the Fortran runtime prints banners on fd 1: keep stdout for the result
        saved = os.dup(1)
        devnull = os.open(os.devnull, os.O_WRONLY)
        os.dup2(devnull, 1)
        try:
            ref = O.RefHarness(cfg, "fast")
            for p in paths:
                st = np.load(p)
                ts.append(float(st["t"]))
                ref.set_state(st["U"], st["V"])
                ref.cfl_dt()
                t0 = time.perf_counter()
                for _ in range(per):
                    ref.advance(ref.cfl_dt())
                    steps += 1
                    if time.perf_counter() - t0 > max_seconds / len(paths):
                        break
                el += time.perf_counter() - t0
            # one thread: the OpenMP runtime the reference build links
            omp = ctypes.CDLL("libomp.so", mode=ctypes.RTLD_GLOBAL)
            omp.omp_set_num_threads(1)
            st = np.load(paths[len(paths) // 2])
            ref.set_state(st["U"], st["V"])
            dt = ref.cfl_dt()
            t1 = time.perf_counter()
            ref.advance(dt)
            el1 = time.perf_counter() - t1
            t1_state = float(st["t"])
            ref.flush()
        finally:
            os.dup2(saved, 1)
            os.close(devnull)
        what = ("reference Fortran (amdflang -O3 -fopenmp; timeStep.f90 serial, as its default(none) "
                "clause does not compile under OpenMP)")
    else:
        kind = "port"
        for p in paths:
            st = np.load(p)
            ts.append(float(st["t"]))
            orc = O.Oracle(cfg)
            orc.set_state(st["U"], st["V"])
            t0 = time.perf_counter()
            for _ in range(per):
                orc.advance(orc.cfl_dt())
                steps += 1
            el += time.perf_counter() - t0
        what = "C restatement (hmhd_oracle.c), OpenMP"
    print(json.dumps({
        "value": steps * n * n / el, "unit": UNIT, "cores": threads, "kind": kind,
        "sample": f"{steps} steps of the {n}x{n} Orszag-Tang solve, {per} from each of the GPU's states at "
                  f"t = {', '.join(f'{x:.4f}' for x in ts)}, {what}, "
                  f"{el:.2f} s",
        "value_1thread": (n * n / el1) if el1 else None,
        "sample_1thread": f"1 step from the state at t = {t1_state:.4f} with 1 OpenMP thread, {el1:.2f} s"
        if el1 else None,
        "cpu_model": _cpu_model(),
        "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
    }))


def _big_stack():
    # the reference keeps its three RK stage arrays as automatic arrays
    # (src/timeStep.f90:154-157): 3 x 68 MB on the stack at 1024^2
    import resource
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def start_cpu_baseline(n: int, steps: int, seconds: float):
    """The CPU child, started before the GPU is touched; it idles on its
    stdin until the GPU leg hands it a state (finish_cpu_baseline)."""
    env = dict(os.environ)
    env.setdefault("OMP_NUM_THREADS", str(len(os.sched_getaffinity(0))))
    env.setdefault("OMP_PROC_BIND", "close")
    env.setdefault("OMP_PLACES", "cores")
    env.setdefault("OMP_STACKSIZE", "512M")
    return subprocess.Popen([sys.executable, str(pathlib.Path(__file__).resolve()), "--cpu-child",
                             "--n", str(n), "--cpu-steps", str(steps), "--cpu-seconds", str(seconds)],
                            stdin=subprocess.PIPE, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                            text=True, env=env, preexec_fn=_big_stack)


def finish_cpu_baseline(proc, state_paths):
    """Hand the child the states (none: nothing to time) and collect its line."""
    try:
        out, err = proc.communicate(input=json.dumps([str(p) for p in state_paths]) + "\n", timeout=900)
    except subprocess.TimeoutExpired:
        proc.kill()
        return {"value": None, "unit": UNIT, "cores": None, "kind": None, "sample": "timed out"}
    line = [ln for ln in out.splitlines() if ln.startswith("{")]
    if proc.returncode != 0 or not line:
        return {"value": None, "unit": UNIT, "cores": None, "kind": None,
                "sample": f"failed: {err.strip()[-300:]}"}
    return json.loads(line[-1])


# ---------------------------------------------------------------------- #
# batched and served workloads

def solo_run(rt, cfg, U, V, tmax: float, nstepmax: int):
    """One run on its own handle, the plain form; returns (seconds, steps)."""
    with rt.Solver(cfg) as s:
        s.upload(U, V)
        s.synchronize()
        t, steps = 0.0, 0
        t0 = time.perf_counter()
        while t < tmax and steps < nstepmax:
            dt = rt.valid_time_step(s.cfl_computedt(), t, tmax)
            s.advance_solution(dt)
            t += dt
            steps += 1
        s.synchronize()
        return time.perf_counter() - t0, steps


def batched_form(rt, args) -> dict:
    """256 runs (hosono_2dmhd_amd.workloads.batched_runs: sides from
    BATCH_SIDES, Orszag-Tang or Sedov, 1% noise, all seeded) advanced
    together, each for --batch-steps steps or to its end time; then the
    same runs one after another on their own handles."""
    import numpy as np
    import hosono_2dmhd_amd.workloads as wl
    runs = wl.batched_runs(args.seed, args.batch_runs)
    # The face kernel's form is tuned once per process and launch size
    # class, on the first launch that meets it: warm every class this form
    # and its per-instance comparison launch (one untimed step each).
    with rt.Batch() as b:
        for r in runs:
            b.add(r.cfg, r.U, r.V, r.tmax, 1)
        b.run(1)
    for n in sorted({r.n for r in runs}):
        r = next(x for x in runs if x.n == n)
        solo_run(rt, r.cfg, r.U, r.V, r.tmax, 1)
    with rt.Batch() as b:
        ids = [b.add(r.cfg, r.U, r.V, r.tmax, args.batch_steps) for r in runs]
        t0 = time.perf_counter()
        b.run(0)
        batch_s = time.perf_counter() - t0
        st = [b.status(i) for i in ids]
        zu = sum(steps * r.n * r.n for (_, steps, _), r in zip(st, runs))
        check = [b.download(i) for i in ids[:2]]
    solo_s, solo_zu = 0.0, 0
    for r in runs:
        sec, steps = solo_run(rt, r.cfg, r.U, r.V, r.tmax, args.batch_steps)
        solo_s += sec
        solo_zu += steps * r.n * r.n
    same = True
    for k in range(len(check)):
        r = runs[k]
        with rt.Solver(r.cfg) as s:
            s.upload(r.U, r.V)
            rt.run(s, r.tmax, args.batch_steps)
            Us, Vs, _ = s.download()
        same = same and bool(np.array_equal(Us, check[k][0]) and np.array_equal(Vs, check[k][1]))
    return {
        "workload": f"{len(runs)} runs, sides {list(wl.BATCH_SIDES)} and Orszag-Tang/Sedov drawn with seed "
                    f"{args.seed}, 1% seeded noise, {args.batch_steps} steps each (or to t_end "
                    f"{wl.T_END_BATCH})",
        "runs": len(runs), "cells": int(sum(r.n * r.n for r in runs)), "zone_updates": int(zu),
        "value": zu / batch_s, "seconds": batch_s,
        "per_instance_plain": {"value": solo_zu / solo_s, "seconds": solo_s,
                               "note": "the same runs one after another, each on its own handle"},
        "speedup_vs_per_instance": (zu / batch_s) / (solo_zu / solo_s),
        "first_runs_bitwise_equal_to_solo": same,
    }


def blast_form(rt, args) -> dict:
    """The MHD blast at 2048^2 (hosono_2dmhd_amd.workloads.make_run:
    outflow, rho = 1, p = 10 for r < 0.1 else 0.1, B = (1, 1, 0)/sqrt 2),
    one solve from t = 0 to --blast-tmax, the state resident in HBM."""
    import hosono_2dmhd_amd.workloads as wl
    n = args.blast_n
    cfg, U, V = wl.make_run("blast", n)
    with rt.Solver(cfg) as s:
        s.upload(U, V)
        s.synchronize()
        t, steps, passes = 0.0, 0, 0
        t0 = time.perf_counter()
        while t < args.blast_tmax:
            dt = rt.valid_time_step(s.cfl_computedt(), t, args.blast_tmax)
            s.advance_solution(dt)
            t += dt
            steps += 1
            passes += int(s.stats()[0].sum())
        s.synchronize()
        el = time.perf_counter() - t0
    return {"workload": f"mhd_blast_{n}", "nx": n, "ny": n, "t_end": t, "steps": steps, "seconds": el,
            "value": steps * n * n / el, "unit": UNIT, "ms_per_step": el / steps * 1e3,
            "mood_passes_per_stage": passes / (3 * steps)}


def served_form(rt, args) -> dict:
    """Poisson arrivals of requests (hosono_2dmhd_amd.workloads.
    served_requests: sides log-uniform in [64, 1024], Orszag-Tang / blast /
    Sedov, each to its end time in T_END_SERVED) into a server with an HBM
    budget; latency is submit -> result."""
    import threading
    import numpy as np
    import hosono_2dmhd_amd.workloads as wl
    reqs = [(r.arrival_s, r.kind, r.n, r.cfg, r.U, r.V)
            for r in wl.served_requests(args.seed + 1, args.served_requests, args.served_rate)]
    T_END_SERVED = wl.T_END_SERVED
    budget = int(args.served_budget_gib * (1 << 30))
    ids = [None] * len(reqs)
    with rt.Server(budget) as srv:
        def feeder():
            t0 = time.perf_counter()
            for k, (ta, kind, n, cfg, U, V) in enumerate(reqs):
                delay = ta - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                ids[k] = srv.submit(cfg, U, V, T_END_SERVED[kind], 0)
        th = threading.Thread(target=feeder)
        th.start()
        th.join()
        res = [srv.wait(i, timeout_s=600.0, arrays=False) for i in ids]
        stats = srv.stats()
    lat = np.array([r.latency_s for r in res])
    sides = np.array([n for _, _, n, _, _, _ in reqs])
    classes = {}
    for name, lo, hi in (("small (<256^2)", 0, 256), ("medium (256^2..511^2)", 256, 512), ("large (>=512^2)", 512, 1 << 30)):
        sel = (sides >= lo) & (sides < hi)
        if sel.any():
            classes[name] = {"requests": int(sel.sum()), "p50_latency_s": float(np.percentile(lat[sel], 50)),
                             "p99_latency_s": float(np.percentile(lat[sel], 99))}
    zu = sum(r.zone_updates for r in res)
    span = max(r.done_s for r in res) - min(r.submit_s for r in res)
    solo_s, solo_zu = 0.0, 0
    if args.served_plain:
        for ta, kind, n, cfg, U, V in reqs:
            sec, steps = solo_run(rt, cfg, U, V, T_END_SERVED[kind], 1_000_000)
            solo_s += sec
            solo_zu += steps * n * n
    out = {
        "workload": f"{len(reqs)} requests, Poisson rate {args.served_rate}/s, sides log-uniform in [64, 1024], "
                    f"Orszag-Tang/blast/Sedov to t_end {T_END_SERVED}, seed {args.seed + 1}",
        "hbm_budget_bytes": budget, "completed": int(sum(r.state == rt.REQ_DONE for r in res)),
        "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()), "latency_by_size": classes,
        "zone_updates": int(zu), "value": zu / span, "seconds": span,
        "peak_active": stats["peak_active"], "peak_bytes": stats["peak_bytes"],
        "engine_steps": stats["engine_steps"], "busy_lane_seconds_per_second": stats["busy_s"] / span,
    }
    if args.served_plain:
        out["per_instance_plain"] = {"value": solo_zu / solo_s, "seconds": solo_s,
                                     "note": "the same requests one after another, each on its own handle"}
        out["speedup_vs_per_instance"] = out["value"] / out["per_instance_plain"]["value"]
    return out


# ---------------------------------------------------------------------- #

def representative_steps(step_passes: list[int], K: int) -> list[int]:
    """Indices of K steps of a solve whose steps took `step_passes` MOOD
    passes: the solve cut by step index into K stretches as equal as
    possible, and from each the first step whose pass count is nearest the
    stretch's mean.  Sorted; fewer than K (no repeats) when the solve has
    fewer than K steps."""
    n = len(step_passes)
    want = set()
    for k in range(K):
        a = k * n // K
        b = max((k + 1) * n // K, a + 1)
        if a >= n:
            continue
        m = sum(step_passes[a:b]) / (b - a)
        want.add(min(range(a, b), key=lambda i: (abs(step_passes[i] - m), i)))
    return sorted(want)


def face_form_used(rt, n: int, forced: str = "tune") -> str | None:
    """The first-pass face kernel's form ("TXxTY/wW") for radius 2 at an
    n x n grid's launch size class: the forced one, or the one this
    process's tuning table holds."""
    if forced != "tune":
        tx, ty, w = rt.FACE_FORMS[int(forced)]
        return f"{tx}x{ty}/w{w}"
    padded = (n + 8) * (n + 8)
    cls = 0 if padded < (1 << 16) else 1 if padded < (1 << 19) else 2 if padded < (1 << 21) else 3
    for e in rt.face_tuning_table():
        if e["R"] == 2 and e["size_class"] == cls:
            return f"{e['tile'][0]}x{e['tile'][1]}/w{e['waves']}"
    return None


def plain_form(rt, args, cpu_proc=None):
    """The metric's workload: one Orszag-Tang solve from t = 0 to t_end.

    The whole solve is `full_solve`.  With --steps K the line's value is
    exactly K of its N steps: a second solve from the same initial state
    times, from each K-th of the solve by step index, the step whose MOOD
    pass count is nearest that stretch's mean, each bracketed by stream
    synchronizations, so the sample has the solve's mix of quiet and
    MOOD-heavy steps (stretches of steps, not of simulated time: the
    MOOD-heavy phases take more, shorter steps);
    without --steps the value is the whole solve.  The GPU states at t =
    (j + 1/2) t_end / 4, j = 0..3, go to the CPU baseline's child (the
    downloads are outside every timed region)."""
    import numpy as np
    g, U, V = workload(args.n)
    cfg = rt.Config(args.n, args.n, R=2, ell_over_delta=12.0)

    # warm-up on a throwaway copy (module load, clocks, first launches)
    with rt.Solver(cfg) as w:
        w.upload(U, V)
        for _ in range(args.warmup):
            w.advance_solution(w.cfl_computedt())
        w.synchronize()

    # HIP events bracket only the first-pass face kernel, except in the
    # breakdown window, where every kernel family is timed
    s = rt.Solver(cfg)
    s.upload(U, V)
    s.profile(["face_full"])
    s.synchronize()
    s.profile_read(reset=True)
    tmax, K, zones = args.tmax, args.steps, args.n * args.n
    # states at t = (j + 1/2) t_end / 4: the CPU baseline's start states and
    # the operation count's inputs
    cpu_t = [(j + 0.5) * tmax / 4 for j in range(4)]
    ci = 0
    state_paths = []
    samp_s, samp_n, samp_passes, samp_first = 0.0, 0, 0, []
    t, steps, passes = 0.0, 0, 0
    step_passes = []    # MOOD passes of every step of the whole solve
    bd = None           # [first step, profile snapshot, wall start] of the breakdown window
    breakdown, bd_range, bd_wall = {}, None, None
    full_s = 0.0
    seg0 = time.perf_counter()
    while t < tmax:
        if ci < len(cpu_t) and t >= cpu_t[ci]:
            s.synchronize()
            full_s += time.perf_counter() - seg0
            while ci < len(cpu_t) and t >= cpu_t[ci]:
                if ci == 0 or state_paths[-1][1] != steps:
                    Um, Vm, _ = s.download(want_scheme=False)
                    path = pathlib.Path(f"/tmp/hmhd_bench_state_{os.getpid()}_{ci}.npz")
                    np.savez(path, U=Um, V=Vm, t=np.array(t))
                    del Um, Vm
                    state_paths.append((path, steps))
                ci += 1
            s.synchronize()
            seg0 = time.perf_counter()
        if bd is None and steps == args.breakdown_at and args.breakdown_steps > 0:
            s.profile(True)
            bd = [steps, s.profile_read(), time.perf_counter()]
        elif bd is not None and bd_range is None and steps == bd[0] + args.breakdown_steps:
            snap1 = s.profile_read()
            bd_wall = time.perf_counter() - bd[2]
            s.profile(["face_full"])
            breakdown = {k: (snap1[k][0] - bd[1][k][0], snap1[k][1] - bd[1][k][1]) for k in snap1}
            bd_range = [bd[0], steps]
        dt = rt.valid_time_step(s.cfl_computedt(), t, tmax)
        s.advance_solution(dt)
        t += dt
        steps += 1
        step_passes.append(int(s.stats()[0].sum()))
        passes += step_passes[-1]
    s.synchronize()
    full_s += time.perf_counter() - seg0
    prof = s.profile_read()
    if K > 0:
        # the timed sample: a second solve from the same initial state, one
        # step from each K-th of the N steps the whole solve took, the one
        # whose MOOD pass count is nearest its stretch's mean
        want = representative_steps(step_passes, K)
        s.upload(U, V)
        s.profile([])
        s.synchronize()
        t2, n2 = 0.0, 0
        for idx in want:
            while n2 < idx:
                dt = rt.valid_time_step(s.cfl_computedt(), t2, tmax)
                s.advance_solution(dt)
                t2 += dt
                n2 += 1
            s.synchronize()
            t_step = time.perf_counter()
            dt = rt.valid_time_step(s.cfl_computedt(), t2, tmax)
            s.advance_solution(dt)
            p = int(s.stats()[0].sum())
            s.synchronize()
            samp_s += time.perf_counter() - t_step
            samp_n += 1
            samp_passes += p
            samp_first.append(n2)
            t2 += dt
            n2 += 1
    s.close()
    cpu = finish_cpu_baseline(cpu_proc, [q for q, _ in state_paths]) if cpu_proc is not None else None
    flops = count_flops(cfg, [np.load(q)["U"] for q, _ in state_paths] or [U])
    for q, _ in state_paths:
        q.unlink(missing_ok=True)

    full = {"value": steps * zones / full_s, "unit": UNIT, "steps": steps, "seconds": full_s, "t_end": t,
            "ms_per_step": full_s / steps * 1e3, "mood_passes_per_stage": passes / (3 * steps),
            "note": "the whole solve from t=0 in the same run; it contains the breakdown window, whose "
                    "events cost it a little"}
    if K > 0 and samp_n > 0:
        value, n_timed, sec = samp_n * zones / samp_s, samp_n, samp_s
        sampled = {"steps": samp_first, "passes_per_stage": samp_passes / (3 * samp_n),
                   "rule": f"one step from each of {K} equal stretches of the {steps}-step solve, the one whose "
                           f"MOOD pass count is nearest the stretch's mean, timed in a second solve from t = 0, "
                           f"each bracketed by stream synchronizations"}
    else:
        value, n_timed, sec = full["value"], steps, full_s
        sampled = {"steps": "all", "passes_per_stage": full["mood_passes_per_stage"], "rule": "the whole solve"}

    ms_full, n_full = prof["face_full"]
    avg_s = ms_full / n_full / 1e3 if n_full else float("nan")
    if not flops or "error" in flops:
        raise RuntimeError(f"hmhd_flop_count failed or is not built: {flops}")
    flops_launch = flops["face_flops_per_launch"]
    achieved = flops_launch / avg_s / 1e12
    # the form the face kernel ran in, and its HBM bytes per launch as two
    # rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) measured it for that form
    form = face_form_used(rt, args.n, args.face_form)
    traffic = None
    tf = ROOT / "profiles" / "face_kernel_traffic_ot1024.json"
    if tf.exists() and args.n == 1024 and form is not None:
        traffic = json.loads(tf.read_text()).get("forms", {}).get(form, {}).get("bytes_per_launch")
    roofline = {
        "bound": "fp64_valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
        "kernel": f"k_face_full<2> in form {form}", "flops_per_launch": flops_launch, "avg_launch_ms": avg_s * 1e3,
        "launches": n_full,
        "peak_note": "FP64 vector (VALU) peak; the kernel uses no MFMA (the FP64 matrix peak is the same on MI355X)",
        "hbm_achieved_gbs": zones * BYTES_FACE_KERNEL / avg_s / 1e9, "hbm_peak_gbs": PEAK_HBM_GBS,
    }
    mood_iters, tiles = breakdown.pop("tiles", (0, 0))
    line = {
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": n_timed,
        "warmup": args.warmup, "ms_per_step": sec / n_timed * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"orszag_tang_{args.n}", "nx": args.n, "ny": args.n, "gp_radius": 2,
                   "ell_over_delta": 12.0, "riemann": "hllc", "cfl": 0.8, "t_end": tmax,
                   "form": "plain", "timed_steps": sampled["rule"],
                   "mood_passes_per_stage": sampled["passes_per_stage"]},
        "timed_sample_steps": sampled["steps"],
        "full_solve": full,
        "roofline": roofline,
        "flops_per_zone_update": flops["flops_per_zone_update"],
        "flop_count": flops,
        "fp64_tflops_full_solve": flops["flops_per_zone_update"] * full["value"] / 1e12,
        "kernel_breakdown": {"steps": bd_range,
                             "ms": {k: v[0] for k, v in breakdown.items()},
                             "launches": {k: v[1] for k, v in breakdown.items()},
                             "mood_iterations": int(mood_iters), "active_tiles_visited": int(tiles),
                             # host wall time of the window's steps, to set against the
                             # kernel families' event times (the gap is GPU idle time)
                             "window_wall_ms": bd_wall * 1e3 if bd_wall is not None else None},
        "cpu_baseline": cpu,
    }
    return line


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0,
                    help="timed steps, spread evenly over the solve (0: the whole solve)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--tmax", type=float, default=0.5)
    ap.add_argument("--breakdown-at", type=int, default=700,
                    help="first step of the all-kernel timing window in the whole solve")
    ap.add_argument("--breakdown-steps", type=int, default=40)
    ap.add_argument("--cpu-steps", type=int, default=20, help="steps in the CPU sample (split over its four states)")
    ap.add_argument("--cpu-seconds", type=float, default=30.0, help="cap on the CPU sample's seconds")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (profiling runs)")
    ap.add_argument("--forms", default="plain,batched,served,blast",
                    help="comma list of plain, batched, served, blast (the line's value is always the plain form's)")
    ap.add_argument("--blast-n", type=int, default=2048)
    ap.add_argument("--blast-tmax", type=float, default=0.2)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--face-form", default="tune",
                    help="first-pass face kernel form: tune (default) or an index into runtime.FACE_FORMS")
    ap.add_argument("--batch-runs", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--served-requests", type=int, default=48)
    ap.add_argument("--served-rate", type=float, default=12.0, help="mean arrivals per second")
    ap.add_argument("--served-budget-gib", type=float, default=8.0)
    ap.add_argument("--served-plain", type=int, default=1, help="also run each request on its own handle")
    ap.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.cpu_child:
        cpu_baseline_child(args.n, args.cpu_steps, args.cpu_seconds)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this one-GPU build "
              f"(a multi-GPU form would place independent solves on separate devices, or shard one "
              f"solve into blocks with halo exchange)", file=sys.stderr)
        return 2
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]

    # the CPU child starts before anything touches the GPU; it waits for a state
    cpu = None if args.no_cpu else start_cpu_baseline(args.n, args.cpu_steps, args.cpu_seconds)

    load_package()
    import hosono_2dmhd_amd.runtime as rt
    if args.face_form != "tune":
        rt.face_tuning(int(args.face_form))
    line = plain_form(rt, args, cpu)
    print(f"plain: {line['value']:.4g} zone-updates/s; full solve {line['full_solve']['value']:.4g}",
          file=sys.stderr, flush=True)
    if "batched" in forms:
        line["batched"] = batched_form(rt, args)
        print(f"batched: {line['batched']['value']:.4g} zone-updates/s", file=sys.stderr, flush=True)
    if "served" in forms:
        line["served"] = served_form(rt, args)
        print(f"served: p50 {line['served']['p50_latency_s']:.3f} s", file=sys.stderr, flush=True)
    if "blast" in forms:
        line["blast"] = blast_form(rt, args)
        print(f"blast: {line['blast']['value']:.4g} zone-updates/s", file=sys.stderr, flush=True)
    # the first-pass face kernel's forms, as tuned in this process: per GP
    # radius and launch size class, the form kept and every form's time
    line["face_tuning"] = rt.face_tuning_table()
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
