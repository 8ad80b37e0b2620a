"""Final-time parity: whole runs through the driver loop of
src/mhd_driver.f90:64-95 (cfl_computedt, validTimeStep's clamp to sim_tmax,
timeStep_advanceSolution) on the GPU, against the oracle run alongside and
against fixtures of the reference's own build (tests/golden/final_*.npz).

The stated final-time tolerance is zero: every dt, and the final U, V and
scheme, bit for bit.  The per-variable relative L1 distance over the
interior, sum |U_gpu - U_ref| / sum |U_ref|, is printed (and is 0).
"""
import hashlib
import math
import pathlib

import numpy as np
import pytest

import cases
import hosono_2dmhd_amd.runtime as rt
import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
VARS = ("dens", "momx", "momy", "momz", "ener", "magx", "magy", "magz")


def rel_l1(U, Uref, ngc):
    a = U[ngc:-ngc, ngc:-ngc, :]
    b = Uref[ngc:-ngc, ngc:-ngc, :]
    out = {}
    for v, name in enumerate(VARS):
        den = np.abs(b[:, :, v]).sum()
        num = np.abs(a[:, :, v] - b[:, :, v]).sum()
        out[name] = float(num / den) if den > 0 else float(num)
    return out


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_ot128_to_final_time_against_the_oracle():
    """Orszag-Tang 128^2, R = 2, l = 12 dx to t = 0.5: every dt, the MOOD
    pass counts of every step, the final U, V and scheme bit for bit."""
    tmax = 0.5
    cfg, g, U, V = cases.make("ot", n=128)
    orc = O.Oracle(cfg)
    orc.set_state(U, V)
    gpu = rt.Solver(cases.runtime_config(cfg))
    gpu.upload(U, V)
    assert gpu.cfl_computedt() == orc.cfl_dt()  # the driver's first call (src/mhd_driver.f90:66)
    t, steps, passes = 0.0, 0, 0
    while t < tmax:
        d_o = rt.valid_time_step(orc.cfl_dt(), t, tmax)
        d_g = rt.valid_time_step(gpu.cfl_computedt(), t, tmax)
        assert d_o == d_g, f"step {steps + 1}: dt oracle {d_o!r} gpu {d_g!r}"
        orc.advance(d_o)
        gpu.advance_solution(d_g)
        po, qo = orc.stats()
        pg, qg = gpu.stats()
        assert np.array_equal(po, pg) and np.array_equal(qo, qg), \
            f"step {steps + 1}: MOOD passes/inner oracle {po}/{qo} gpu {pg}/{qg}"
        passes += int(pg.sum())
        t += d_g
        steps += 1
    Uo, Vo = orc.get_state()
    Ug, Vg, Sg = gpu.download()
    l1 = rel_l1(Ug, Uo, cfg.R + 2)
    print(f"OT 128^2 to t = {t}: {steps} steps, {passes / (3 * steps):.2f} MOOD passes per stage, rel L1 {l1}")
    assert np.array_equal(Ug, Uo), f"final U differs: rel L1 {l1}"
    assert np.array_equal(Vg, Vo), "final V differs"
    assert np.array_equal(Sg, orc.stage_schemes()[2]), "final scheme differs"
    assert passes > 6 * steps, "expected MOOD passes beyond the first"
    gpu.close()


@pytest.mark.parametrize("n", [128, 256, 1024])
def test_ot_final_time_against_reference_fixture(n):
    """Orszag-Tang n^2 to t = 0.5 on the GPU against the fixture of the
    reference build (the oracle's for 1024^2): the dt sequence, the sha256
    of the final U, V and scheme, and the exact interior totals."""
    path = GOLDEN / f"final_ot{n}.npz"
    if not path.exists():
        pytest.skip(f"{path.name} not generated (tests/golden/make_golden.py final ot {n} 0.5 ...)")
    fx = np.load(path)
    tmax = float(fx["tmax"])
    cfg, g, U, V = cases.make("ot", n=n)
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(U).tobytes())
    h.update(np.ascontiguousarray(V).tobytes())
    assert h.hexdigest() == str(fx["input_sha256"]), "initial condition differs from the fixture's"
    gpu = rt.Solver(cases.runtime_config(cfg))
    gpu.upload(U, V)
    gpu.cfl_computedt()
    dts_ref = fx["dts"]
    t, k = 0.0, 0
    while t < tmax:
        dt = rt.valid_time_step(gpu.cfl_computedt(), t, tmax)
        assert k < len(dts_ref) and dt == dts_ref[k], f"step {k + 1}: dt {dt!r}, reference {dts_ref[min(k, len(dts_ref) - 1)]!r}"
        gpu.advance_solution(dt)
        t += dt
        k += 1
    assert k == int(fx["steps"]) and t == float(fx["t"])
    Ug, Vg, Sg = gpu.download()
    ngc = cfg.R + 2
    inner = Ug[ngc:-ngc, ngc:-ngc, :]
    totals = np.array([math.fsum(inner[:, :, v].ravel()) for v in range(8)])
    sample = inner[::16, ::16, 0]
    ref_sample = fx["rho_sample"]
    print(f"OT {n}^2 to t = {t}: {k} steps; density sample rel L1 "
          f"{np.abs(sample - ref_sample).sum() / np.abs(ref_sample).sum():.3e}; "
          f"totals diff {np.abs(totals - fx['totals']).max():.3e}")
    assert sha(Ug) == str(fx["U_sha256"]), "final U differs from the reference's"
    assert sha(Vg) == str(fx["V_sha256"]), "final V differs from the reference's"
    assert sha(Sg.astype(np.int32)) == str(fx["scheme_sha256"]), "final scheme differs from the reference's"
    assert np.array_equal(totals, fx["totals"])
    gpu.close()
