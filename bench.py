#!/usr/bin/env python3
"""FP64 zone-updates/s of the hosono-2DMHD step on one MI355X.

Workload: the Orszag-Tang vortex on a 1024 x 1024 periodic grid (rho =
gamma^2, p = gamma, v = (-sin 2 pi y, sin 2 pi x), B = (-sin 2 pi y,
sin 4 pi x)/sqrt(4 pi), gamma = 5/3), GP-MOOD radius 2 with l = 12 dx, HLLC,
SSP-RK3, CFL 0.8, one instance, state resident in HBM.  A step is one
driver iteration: the CFL reduction (cfl_computedt) and one full SSP-RK3
step with every MOOD pass (timeStep_advanceSolution).  A zone-update is one
interior cell advanced one step.

With no --steps the timed region is the whole solve from t = 0 to t = 0.5;
with --steps K it is exactly K steps from t = 0.  Warm-up steps run on a
separate copy before the timed region.

One JSON line on stdout.  `roofline` covers the dominant kernel, the face
kernel in its first-pass form (GP reconstruction + admissibility + HLLC),
timed with HIP events on the solver's stream inside the timed region.
`cpu_baseline` times the reference's own Fortran step (compiled from its
sources by oracle/Makefile, OpenMP on the host cores) on a bounded sample of
the same workload, in a child process started before the GPU is touched.

Beside the plain form the line carries two more objects (--forms):
`batched`, 256 seeded runs of mixed sizes advanced together in one engine,
against the same runs launched one after another on their own handles; and
`served`, Poisson arrivals of mixed requests into a server with an HBM
budget, with p50/p99 submit-to-result latency and aggregate zone-updates/s.
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

# Algorithmic FP64 flops (+ - * / sqrt = 1; min/max/abs/compare = 0) of the
# first MOOD pass at GP radius 2, counted from the arithmetic each kernel
# performs, summed per zone below.
FLOP_CONS2PRIM = 21
FLOP_CONS2PRES = 17
FLOP_FAST_SPEED = 21
FLOP_CONS2FLUX = 33      # given the primitives
FLOP_HLLC_STAR = 2 * FLOP_CONS2PRIM + 20 + 3 * FLOP_FAST_SPEED + 4 + 2 * FLOP_CONS2FLUX + 49 + 18 + 14 + 2 + 13 + 8 + 27 + 24
FLOP_RECON_R2 = 4 * 3 * 13 * 8 * 2      # 4 face sides x 3 points x 13 cells x 8 vars, mul + add
FLOP_ADMISSIBILITY_R2 = 4 * 3 * FLOP_CONS2PRES
FLOP_FLUX_SUM_R2 = 2 * 3 * 16           # 2 owned faces x 3 points x (8 mul + 8 add)
FLOP_FACE_KERNEL_R2 = FLOP_RECON_R2 + FLOP_ADMISSIBILITY_R2 + 2 * 3 * FLOP_HLLC_STAR + FLOP_FLUX_SUM_R2
FLOP_UPDATE_KERNEL = 2 * 8 + 8 * 6 + (FLOP_CONS2PRES + 5 * FLOP_CONS2PRIM + 7 + 7 + 5)
FLOP_RK_COMBINE = 2 * 8 * 3
FLOP_CFL = FLOP_CONS2PRIM + 5 + 16 + 2
FLOP_ZONE_UPDATE_R2 = 3 * (FLOP_FACE_KERNEL_R2 + FLOP_UPDATE_KERNEL) + FLOP_RK_COMBINE + FLOP_CFL
# compulsory HBM bytes of the face kernel per interior cell: read U (8 FP64),
# write the raw x and y fluxes (16 FP64) and the two FLICD components (2 FP64)
BYTES_FACE_KERNEL = 8 * 8 + 16 * 8 + 2 * 8


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


def cpu_baseline_child(n: int, steps: int) -> None:
    """Runs in a child process: the reference's Fortran step on the host."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    g, U, V = workload(n)
    cfg = O.Config(n, n, R=2, ell_over_delta=12.0)
    if O.ref_available("fast"):
        kind = "reference"
        # the Fortran runtime prints banners on fd 1: keep stdout for the result
        saved = os.dup(1)
        devnull = os.open(os.devnull, os.O_WRONLY)
        os.dup2(devnull, 1)
        try:
            ref = O.RefHarness(cfg, "fast")
            ref.set_state(U, V)
            ref.cfl_dt()
            t0 = time.perf_counter()
            ref.run(steps, 1.0e30)
            el = time.perf_counter() - t0
            ref.flush()
        finally:
            os.dup2(saved, 1)
            os.close(devnull)
    else:
        kind = "port"
        orc = O.Oracle(cfg)
        orc.set_state(U, V)
        t0 = time.perf_counter()
        orc.run(steps, 1.0e30)
        el = time.perf_counter() - t0
    print(json.dumps({"value": steps * n * n / el, "unit": UNIT, "cores": threads, "kind": kind,
                      "sample": f"{steps} steps of the {n}x{n} Orszag-Tang solve from t=0 "
                                f"({'reference Fortran, amdflang -O3 -fopenmp' if kind == 'reference' else 'C restatement, OpenMP'}), "
                                f"{el:.2f} s"}))


def _big_stack():
    # the reference keeps its three RK stage arrays as automatic arrays
    # (src/timeStep.f90:154-157): 3 x 68 MB on the stack at 1024^2
    import resource
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))


def run_cpu_baseline(n: int, steps: int):
    try:
        p = subprocess.run([sys.executable, str(pathlib.Path(__file__).resolve()), "--cpu-child",
                            "--n", str(n), "--cpu-steps", str(steps)],
                           capture_output=True, text=True, timeout=600, preexec_fn=_big_stack)
        line = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
        if p.returncode != 0 or not line:
            return {"value": None, "unit": UNIT, "cores": None, "kind": None,
                    "sample": f"failed: {p.stderr.strip()[-300:]}"}
        return json.loads(line[-1])
    except subprocess.TimeoutExpired:
        return {"value": None, "unit": UNIT, "cores": None, "kind": None, "sample": "timed out"}


# ---------------------------------------------------------------------- #
# batched and served workloads

BATCH_SIDES = (64, 96, 128, 192, 256, 384)
T_END_BATCH = {"ot": 0.5, "sedov": 0.05}
T_END_SERVED = {"ot": 0.05, "blast": 0.01, "sedov": 0.005}


def make_run(rt, kind: str, n: int, noise_seed: int | None):
    """(config, U, V) of one run: the Orszag-Tang vortex above, the MHD blast
    (rho = 1, p = 10 for r < 0.1 else 0.1, B = (1, 1, 0)/sqrt 2, outflow) or
    the Sedov blast (E = 1 in the central 3x3 cells, rho = 1, p = 1e-5,
    B = (0.5, 0, 0), outflow), optionally with 1% seeded uniform noise on
    rho, v, p and B."""
    import hosono_2dmhd_amd.ics as ics
    gamma = 5.0 / 3.0
    if kind == "ot":
        g = ics.Grid(n, n, 0.0, 1.0, 0.0, 1.0, R=2)
        U, V = ics.orszag_tang(g, gamma)
        bc = rt.BC_PERIODIC
    elif kind == "blast":
        g = ics.Grid(n, n, -0.5, 0.5, -0.5, 0.5, R=2)
        U, V = ics.mhd_blast(g, gamma)
        bc = rt.BC_OUTFLOW
    else:
        g = ics.Grid(n, n, -0.5, 0.5, -0.5, 0.5, R=2)
        U, V = ics.sedov(g, gamma, energy=1.0, p_bg=1e-5, B=(0.5, 0.0, 0.0))
        bc = rt.BC_OUTFLOW
    if noise_seed is not None:
        U, V = ics.perturb(V, g, gamma, bc, noise_seed, 0.01)
    cfg = rt.Config(n, n, g.xbeg, g.xend, g.ybeg, g.yend, R=2, ell_over_delta=12.0, gamma=gamma, bc=bc)
    return cfg, U, V


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
    """256 runs (sides from BATCH_SIDES, Orszag-Tang or Sedov, 1% noise, all
    seeded) advanced together, each for --batch-steps steps or to its end
    time; then the same runs one after another on their own handles."""
    import numpy as np
    rng = np.random.default_rng(args.seed)
    runs = []
    for k in range(args.batch_runs):
        n = int(rng.choice(BATCH_SIDES))
        kind = "ot" if rng.random() < 0.5 else "sedov"
        runs.append((kind, n, make_run(rt, kind, n, args.seed * 100003 + k)))
    with rt.Batch() as b:
        ids = [b.add(cfg, U, V, T_END_BATCH[kind], args.batch_steps) for kind, n, (cfg, U, V) in runs]
        t0 = time.perf_counter()
        b.run(0)
        batch_s = time.perf_counter() - t0
        st = [b.status(i) for i in ids]
        zu = sum(steps * n * n for (_, steps, _), (_, n, _) in zip(st, runs))
        check = [b.download(i) for i in ids[:2]]
    solo_s, solo_zu = 0.0, 0
    for k, (kind, n, (cfg, U, V)) in enumerate(runs):
        sec, steps = solo_run(rt, cfg, U, V, T_END_BATCH[kind], args.batch_steps)
        solo_s += sec
        solo_zu += steps * n * n
    same = True
    for k in range(len(check)):
        kind, n, (cfg, U, V) = runs[k]
        with rt.Solver(cfg) as s:
            s.upload(U, V)
            rt.run(s, T_END_BATCH[kind], args.batch_steps)
            Us, Vs, _ = s.download()
        same = same and bool(np.array_equal(Us, check[k][0]) and np.array_equal(Vs, check[k][1]))
    sides = [n for _, n, _ in runs]
    return {
        "workload": f"{len(runs)} runs, sides {list(BATCH_SIDES)} and Orszag-Tang/Sedov drawn with seed "
                    f"{args.seed}, 1% seeded noise, {args.batch_steps} steps each (or to t_end "
                    f"{T_END_BATCH})",
        "runs": len(runs), "cells": int(sum(n * n for n in sides)), "zone_updates": int(zu),
        "value": zu / batch_s, "seconds": batch_s,
        "per_instance_plain": {"value": solo_zu / solo_s, "seconds": solo_s,
                               "note": "the same runs one after another, each on its own handle"},
        "speedup_vs_per_instance": (zu / batch_s) / (solo_zu / solo_s),
        "first_runs_bitwise_equal_to_solo": same,
    }


def served_form(rt, args) -> dict:
    """Poisson arrivals of requests (sides log-uniform in [64, 1024],
    Orszag-Tang / blast / Sedov, each to its end time in T_END_SERVED) into
    a server with an HBM budget; latency is submit -> result."""
    import threading
    import numpy as np
    rng = np.random.default_rng(args.seed + 1)
    reqs = []
    t_arr = 0.0
    for k in range(args.served_requests):
        n = int(round(math.exp(rng.uniform(math.log(64.0), math.log(1024.0)))))
        kind = ("ot", "blast", "sedov")[int(rng.integers(3))]
        t_arr += float(rng.exponential(1.0 / args.served_rate))
        cfg, U, V = make_run(rt, kind, n, None)
        reqs.append((t_arr, kind, n, cfg, U, V))
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
        "mean_latency_s": float(lat.mean()), "zone_updates": int(zu), "value": zu / span, "seconds": span,
        "peak_active": stats["peak_active"], "peak_bytes": stats["peak_bytes"],
        "engine_steps": stats["engine_steps"], "busy_fraction": stats["busy_s"] / span,
    }
    if args.served_plain:
        out["per_instance_plain"] = {"value": solo_zu / solo_s, "seconds": solo_s,
                                     "note": "the same requests one after another, each on its own handle"}
        out["speedup_vs_per_instance"] = out["value"] / out["per_instance_plain"]["value"]
    return out


# ---------------------------------------------------------------------- #

def plain_form(rt, args):
    """The metric's workload: one Orszag-Tang solve, timed from t = 0."""
    g, U, V = workload(args.n)
    cfg = rt.Config(args.n, args.n, R=2, ell_over_delta=12.0)

    # warm-up on a throwaway copy (module load, clocks, first launches)
    with rt.Solver(cfg) as w:
        w.upload(U, V)
        for _ in range(args.warmup):
            w.advance_solution(w.cfl_computedt())
        w.synchronize()

    # timed region: HIP events bracket only the first-pass face kernel
    s = rt.Solver(cfg)
    s.upload(U, V)
    s.profile(["face_full"])
    s.synchronize()
    s.profile_read(reset=True)
    tmax = args.tmax if args.steps == 0 else float("inf")
    limit = args.steps if args.steps > 0 else 10_000_000
    t = 0.0
    steps = 0
    passes = 0
    # every kernel family is timed over a window of steps from the middle
    # of the solve (shocks formed, many MOOD passes); events around every
    # launch slow those steps slightly
    win0 = args.breakdown_at if args.steps == 0 else max(0, limit // 2 - args.breakdown_steps // 2)
    win1 = win0 + args.breakdown_steps
    snap0 = snap1 = None
    w_t0 = w_t1 = None
    t0 = time.perf_counter()
    while t < tmax and steps < limit:
        if steps == win0:
            s.profile(True)
            snap0 = s.profile_read()
            w_t0 = time.perf_counter()
        elif steps == win1:
            snap1 = s.profile_read()
            w_t1 = time.perf_counter()
            s.profile(["face_full"])
        dt = rt.valid_time_step(s.cfl_computedt(), t, tmax)
        s.advance_solution(dt)
        t += dt
        steps += 1
        passes += int(s.stats()[0].sum())
    s.synchronize()
    elapsed = time.perf_counter() - t0
    prof = s.profile_read()
    s.close()
    if snap1 is None:
        snap1 = prof
        win1 = steps
        w_t1 = time.perf_counter()
    breakdown = {k: (snap1[k][0] - snap0[k][0], snap1[k][1] - snap0[k][1]) for k in snap1} if snap0 else {}

    zones = args.n * args.n
    value = steps * zones / elapsed
    ms_full, n_full = prof["face_full"]
    avg_s = ms_full / n_full / 1e3 if n_full else float("nan")
    flops_launch = zones * FLOP_FACE_KERNEL_R2
    achieved = flops_launch / avg_s / 1e12
    traffic = None
    tf = ROOT / "profiles" / "face_kernel_traffic_ot1024.json"
    if tf.exists() and args.n == 1024:
        traffic = json.loads(tf.read_text()).get("bytes_per_launch")
    roofline = {
        "bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
        "kernel": "k_face<2,FULL>", "flops_per_launch": flops_launch, "avg_launch_ms": avg_s * 1e3,
        "launches": n_full, "peak_note": "FP64 vector peak (equal to the FP64 MFMA peak on MI355X)",
        "hbm_achieved_gbs": zones * BYTES_FACE_KERNEL / avg_s / 1e9, "hbm_peak_gbs": PEAK_HBM_GBS,
    }
    mood_iters, tiles = breakdown.pop("tiles", (0, 0))
    line = {
        "metric": METRIC, "value": value, "unit": UNIT, "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": elapsed / steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"orszag_tang_{args.n}", "nx": args.n, "ny": args.n, "gp_radius": 2,
                   "ell_over_delta": 12.0, "riemann": "hllc", "cfl": 0.8, "t_end": t,
                   "form": "plain", "mood_passes_per_stage": passes / (3 * steps)},
        "roofline": roofline,
        "flops_per_zone_update": FLOP_ZONE_UPDATE_R2,
        "kernel_breakdown": {"steps": [win0, win1],
                             "ms": {k: v[0] for k, v in breakdown.items()},
                             "launches": {k: v[1] for k, v in breakdown.items()},
                             "mood_iterations": int(mood_iters), "active_tiles_visited": int(tiles),
                             # host wall time of the window's steps, to set against the
                             # kernel families' event times (the gap is GPU idle time)
                             "window_wall_ms": (w_t1 - w_t0) * 1e3 if w_t0 is not None else None},
    }
    return line


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (0: the whole solve to t=0.5)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--tmax", type=float, default=0.5)
    ap.add_argument("--breakdown-at", type=int, default=700, help="first step of the all-kernel timing window")
    ap.add_argument("--breakdown-steps", type=int, default=40)
    ap.add_argument("--cpu-steps", type=int, default=3, help="steps in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (profiling runs)")
    ap.add_argument("--forms", default="plain,batched,served",
                    help="comma list of plain, batched, served (the line's value is always the plain form's)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--batch-runs", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--served-requests", type=int, default=48)
    ap.add_argument("--served-rate", type=float, default=12.0, help="mean arrivals per second")
    ap.add_argument("--served-budget-gib", type=float, default=8.0)
    ap.add_argument("--served-plain", type=int, default=1, help="also run each request on its own handle")
    ap.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.cpu_child:
        cpu_baseline_child(args.n, args.cpu_steps)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported on this one-GPU build "
              f"(a multi-GPU form would place independent solves on separate devices, or shard one "
              f"solve into blocks with halo exchange)", file=sys.stderr)
        return 2
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]

    cpu = None if args.no_cpu else run_cpu_baseline(args.n, args.cpu_steps)

    load_package()
    import hosono_2dmhd_amd.runtime as rt
    line = plain_form(rt, args)
    line["cpu_baseline"] = cpu
    if "batched" in forms:
        line["batched"] = batched_form(rt, args)
        print(f"batched: {line['batched']['value']:.4g} zone-updates/s", file=sys.stderr, flush=True)
    if "served" in forms:
        line["served"] = served_form(rt, args)
        print(f"served: p50 {line['served']['p50_latency_s']:.3f} s", file=sys.stderr, flush=True)
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
