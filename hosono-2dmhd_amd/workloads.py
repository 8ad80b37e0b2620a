"""The seeded workloads of the batched and served forms, drawn the same
way wherever they are run (bench.py, the parity tests, the fixture writer).

A run is the Orszag-Tang vortex (periodic [0,1]^2, rho = gamma^2, p = gamma,
v = (-sin 2 pi y, sin 2 pi x), B = (-sin 2 pi y, sin 4 pi x)/sqrt(4 pi)),
the MHD blast ([-0.5,0.5]^2, outflow, rho = 1, p = 10 for r < 0.1 else 0.1,
B = (1, 1, 0)/sqrt 2) or the Sedov blast ([-0.5,0.5]^2, outflow, E = 1 in
the central 3x3 cells, rho = 1, p = 1e-5, B = (0.5, 0, 0)); gamma = 5/3,
GP radius 2 with l = 12 dx, HLLC, CFL 0.8, FP64.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import ics
from . import runtime as rt

BATCH_SIDES = (64, 96, 128, 192, 256, 384)
T_END_BATCH = {"ot": 0.5, "sedov": 0.05}
T_END_SERVED = {"ot": 0.05, "blast": 0.01, "sedov": 0.005}
GAMMA = 5.0 / 3.0


@dataclasses.dataclass
class Run:
    kind: str
    n: int
    cfg: rt.Config
    U: np.ndarray
    V: np.ndarray
    tmax: float
    noise_seed: int | None = None
    arrival_s: float = 0.0


def make_run(kind: str, n: int, noise_seed: int | None = None):
    """(config, U, V) of one run, optionally with 1% seeded uniform noise on
    rho, v, p and B (ics.perturb)."""
    if kind == "ot":
        g = ics.Grid(n, n, 0.0, 1.0, 0.0, 1.0, R=2)
        U, V = ics.orszag_tang(g, GAMMA)
        bc = rt.BC_PERIODIC
    elif kind == "blast":
        g = ics.Grid(n, n, -0.5, 0.5, -0.5, 0.5, R=2)
        U, V = ics.mhd_blast(g, GAMMA)
        bc = rt.BC_OUTFLOW
    elif kind == "sedov":
        g = ics.Grid(n, n, -0.5, 0.5, -0.5, 0.5, R=2)
        U, V = ics.sedov(g, GAMMA, energy=1.0, p_bg=1e-5, B=(0.5, 0.0, 0.0))
        bc = rt.BC_OUTFLOW
    else:
        raise KeyError(kind)
    if noise_seed is not None:
        U, V = ics.perturb(V, g, GAMMA, bc, noise_seed, 0.01)
    cfg = rt.Config(n, n, g.xbeg, g.xend, g.ybeg, g.yend, R=2, ell_over_delta=12.0, gamma=GAMMA, bc=bc)
    return cfg, U, V


def batched_runs(seed: int, count: int, build: bool = True) -> list[Run]:
    """`count` runs: side drawn uniformly from BATCH_SIDES, Orszag-Tang or
    Sedov with equal odds, 1% noise seeded by seed * 100003 + k, each to
    its T_END_BATCH.  The first k runs of a larger count are the same runs.
    build=False leaves U and V empty (the draw only)."""
    rng = np.random.default_rng(seed)
    runs = []
    for k in range(count):
        n = int(rng.choice(BATCH_SIDES))
        kind = "ot" if rng.random() < 0.5 else "sedov"
        ns = seed * 100003 + k
        if build:
            cfg, U, V = make_run(kind, n, ns)
        else:
            cfg, U, V = None, None, None
        runs.append(Run(kind, n, cfg, U, V, T_END_BATCH[kind], ns))
    return runs


def served_requests(seed: int, count: int, rate: float, build: bool = True,
                    sides: tuple[int, int] = (64, 1024)) -> list[Run]:
    """`count` requests arriving as a Poisson process of `rate` per second:
    side log-uniform in `sides`, kind uniform over Orszag-Tang, blast and
    Sedov, no noise, each to its T_END_SERVED."""
    rng = np.random.default_rng(seed)
    out = []
    t_arr = 0.0
    lo, hi = math.log(float(sides[0])), math.log(float(sides[1]))
    for _ in range(count):
        n = int(round(math.exp(rng.uniform(lo, hi))))
        kind = ("ot", "blast", "sedov")[int(rng.integers(3))]
        t_arr += float(rng.exponential(1.0 / rate))
        if build:
            cfg, U, V = make_run(kind, n, None)
        else:
            cfg, U, V = None, None, None
        out.append(Run(kind, n, cfg, U, V, T_END_SERVED[kind], None, t_arr))
    return out
