#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/.

1. reconstruct_R3_ot128.json -- the numbers of the reference's only numeric
   golden, test/test_reconstruct_faceValsWithGP/output.txt (x-momentum face
   values of the 3x3 cells around interior cell (5, 5), R = 3, l = 6 dx,
   Orszag-Tang 128^2, printed ES12.5), parsed into data.
2. ref_<case>.npz -- the reference's own Fortran step (compiled from its
   sources by oracle/Makefile, oracle/_ref/libref_exact.so) run on the
   cases of tests/cases.py: the dt sequence, the final U and V, the final
   scheme of each stage of the last step and its MOOD pass counts.

Run from the repository root after `make -C oracle ref`:
    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import pathlib
import re
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT / "tests"))
import conftest  # noqa: E402,F401  (loads the package, oracle path)
import cases  # noqa: E402
import oracle as O  # noqa: E402

REF_PRINT = pathlib.Path("/root/reference/test/test_reconstruct_faceValsWithGP/output.txt")

# (case name, size override, radius override, steps)
RUNS = [
    ("ot", 32, None, 40),
    ("blast", 32, None, 12),
    ("sedov", 32, None, 10),
    ("ot_ref", 24, None, 10),
    ("explosion_ref", 24, None, 12),
    ("sedov_ref", 24, None, 8),
]


def parse_reference_print():
    """output.txt lays out each cell as: north values k = nq..1, then rows
    k = 1..nq of (west, east), the cell average in the middle, south values
    k = nq..1 (test/test_reconstruct_faceValsWithGP.f90:83-149)."""
    num = re.compile(r"-?\d\.\d{5}E[+-]\d{2}")
    rows = [ln for ln in REF_PRINT.read_text().splitlines() if ln.startswith("|")]
    assert len(rows) == 24
    cells = {}
    nq = 4
    for r in range(3):
        dj = 1 - r
        blk = rows[8 * r:8 * r + 8]
        seg = [ln.split("|")[1:4] for ln in blk]
        for c in range(3):
            di = c - 1
            north = [float(x) for x in num.findall(seg[0][c])]
            south = [float(x) for x in num.findall(seg[7][c])]
            west, east, avg = [], [], None
            for k in range(nq):
                vals = [float(x) for x in num.findall(seg[2 + k][c])]
                west.append(vals[0])
                east.append(vals[-1])
                if len(vals) == 3:
                    avg = vals[1]
            cells[f"{di},{dj}"] = {
                "north": north[::-1], "south": south[::-1],  # now k = 1..nq
                "west": west, "east": east, "average": avg,
                "strings": {"north": num.findall(seg[0][c])[::-1], "south": num.findall(seg[7][c])[::-1]},
            }
    return {"source": "test/test_reconstruct_faceValsWithGP/output.txt",
            "variable": "momx (U index 2)", "radius": 3, "ell_over_delta": 6.0, "n": 128,
            "centre_interior_cell": [5, 5], "cells": cells}


def input_digest(U, V):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(U).tobytes())
    h.update(np.ascontiguousarray(V).tobytes())
    return h.hexdigest()


def run_reference(name, n, R, steps):
    cfg, g, U, V = cases.make(name, n=n, R=R)
    ref = O.RefHarness(cfg, "exact")
    ref.set_state(U, V)
    dts = []
    for _ in range(steps - 1):
        dt = ref.cfl_dt()
        ref.advance(dt)
        dts.append(dt)
    dt = ref.cfl_dt()
    Um, masks, passes = ref.step_masks(dt)
    ref.advance(dt)
    dts.append(dt)
    Uf, Vf, S = ref.get_state()
    assert np.array_equal(Um, Uf), "re-assembled step differs from timeStep_advanceSolution"
    ref.flush()
    return dict(dts=np.array(dts), U=Uf, V=Vf, stage_schemes=masks.astype(np.int8),
                passes=passes, input_sha256=np.array(input_digest(U, V)))


def main():
    if REF_PRINT.exists():
        (HERE / "reconstruct_R3_ot128.json").write_text(json.dumps(parse_reference_print(), indent=1) + "\n")
    if not O.ref_available("exact"):
        sys.exit("oracle/_ref/libref_exact.so missing: run `make -C oracle ref` first")
    # the reference prints banners through the Fortran runtime on fd 1
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    results = {}
    try:
        os.dup2(devnull, 1)
        for name, n, R, steps in RUNS:
            results[name] = run_reference(name, n, R, steps)
    finally:
        os.dup2(saved, 1)
    for name, (case, n, R, steps) in zip(results, RUNS):
        np.savez_compressed(HERE / f"ref_{name}.npz", n=np.array(n or -1), steps=np.array(steps),
                            **results[name])
        print(name, "steps", steps, "passes of last step", results[name]["passes"])


if __name__ == "__main__":
    main()
