"""Parity of the HIP step with the oracle (hmhd_oracle.c), through the C ABI.

Both run on the same seeded or closed-form inputs; the library is built
without FMA contraction and follows the reference's operation order, so the
bar is bit-for-bit equality of U, V, the per-stage MOOD scheme masks, dt and
the MOOD pass counts.  The oracle itself is bit-identical to the reference's
Fortran (tests/test_oracle_vs_reference.py).
"""
import math

import numpy as np
import pytest

import cases
import hosono_2dmhd_amd.runtime as rt
import oracle as O

pytestmark = pytest.mark.gpu


def _pair(name, **kw):
    cfg, g, U, V = cases.make(name, **kw)
    orc = O.Oracle(cfg)
    orc.set_state(U, V)
    gpu = rt.Solver(cases.runtime_config(cfg))
    gpu.upload(U, V)
    return cfg, g, orc, gpu


def _step_both(orc, gpu, record=True):
    d1 = orc.cfl_dt()
    d2 = gpu.cfl_computedt()
    assert d1 == d2, f"dt differs: oracle {d1!r} gpu {d2!r}"
    orc.advance(d1)
    if record:
        gpu.record_stage_schemes(True)
    gpu.advance_solution(d2)
    return d1


def _compare_state(orc, gpu, where):
    Uo, Vo = orc.get_state()
    Ug, Vg, Sg = gpu.download()
    so = orc.stage_schemes()
    if not np.array_equal(Uo, Ug):
        d = np.abs(Uo - Ug)
        j, i, v = np.unravel_index(np.argmax(d), d.shape)
        pytest.fail(f"{where}: U differs, max |diff| {d.max():.3e} at (j={j}, i={i}, var={v}), "
                    f"{int((d > 0).sum())} values differ")
    assert np.array_equal(Vo, Vg), f"{where}: V differs (max {np.abs(Vo - Vg).max():.3e})"
    assert np.array_equal(so[2], Sg), f"{where}: final scheme differs"
    assert np.array_equal(so, gpu.stage_schemes()), f"{where}: per-stage scheme masks differ"
    po, qo = orc.stats()
    pg, qg = gpu.stats()
    assert np.array_equal(po, pg) and np.array_equal(qo, qg), \
        f"{where}: MOOD passes/inner oracle {po}/{qo} gpu {pg}/{qg}"


def test_gp_tables_match_oracle():
    for name, R in (("ot", 2), ("ot_ref", 3), ("explosion_ref", 1), ("ot_rect", 2)):
        cfg, g, orc, gpu = _pair(name, R=R)
        pv1, qw1, st1 = orc.gp_tables()
        pv2, qw2, st2 = gpu.gp_tables()
        assert np.array_equal(pv1, pv2), name
        assert np.array_equal(qw1, qw2), name
        assert np.array_equal(st1, st2), name
        gpu.close()


def test_cfl_matches_oracle_with_and_without_uploaded_prims():
    cfg, g, U, V = cases.make("ot")
    orc = O.Oracle(cfg)
    gpu = rt.Solver(cases.runtime_config(cfg))
    orc.set_state(U, V)
    gpu.upload(U, V)
    assert orc.cfl_dt() == gpu.cfl_computedt()
    orc.set_state(U, None)
    gpu.upload(U, None)
    assert orc.cfl_dt() == gpu.cfl_computedt()


def test_download_before_any_step_returns_upload():
    cfg, g, orc, gpu = _pair("blast", n=32)
    Ug, Vg, Sg = gpu.download()
    Uo, Vo = orc.get_state()
    assert np.array_equal(Ug, Uo) and np.array_equal(Vg, Vo)
    assert not Sg.any()


@pytest.mark.parametrize("name,steps", [
    ("ot", 12),             # periodic MHD, shocks form, many MOOD passes
    ("blast", 10),          # outflow MHD, fallback to first order
    ("sedov", 10),          # strong point blast with B
    ("ot_ref", 6),          # R = 3, l = 6 dx, the reference's own IC
    ("explosion_ref", 8),   # R = 1, hydro (sim_forceHydro)
    ("sedov_ref", 6),       # R = 3 hydro
    ("ot_rect", 8),         # nx != ny, dx != dy, periodic
    ("blast_rect", 8),      # nx != ny, outflow
    ("blast_periodic", 8),  # inner-loop rounds across periodic guard bands
])
def test_steps_bitwise(name, steps):
    cfg, g, orc, gpu = _pair(name)
    for n in range(steps):
        _step_both(orc, gpu)
        _compare_state(orc, gpu, f"{name} step {n + 1}")


@pytest.mark.parametrize("name,n,R,steps", [
    ("blast", 48, 3, 10),  # R = 3 (eight lanes a face on pairs), fallback through radii 2, 1 and 0
    ("ot", 48, 1, 12),     # R = 1 (four lanes a face), periodic shocks
])
def test_steps_bitwise_other_radii(name, n, R, steps):
    """Later MOOD passes at every lower radius of R = 3 and R = 1 solves, on
    the later-pass face kernel's lane pairs and the four-block tile update."""
    cfg, g, orc, gpu = _pair(name, n=n, R=R)
    for k in range(steps):
        _step_both(orc, gpu)
        _compare_state(orc, gpu, f"{name} {n}^2 R={R} step {k + 1}")


@pytest.mark.parametrize("name,n,R", [
    ("ot", 6, 1), ("ot", 8, 2), ("ot", 10, 3),  # N = 2 NGC: guard and image bands meet
    ("blast", 6, 1), ("blast", 8, 2),           # one tile holds the whole padded grid
    ("sedov", 10, 3),                           # two ragged tiles a side, several passes
])
def test_smallest_grids_bitwise(name, n, R):
    cfg, g, orc, gpu = _pair(name, n=n, R=R)
    for k in range(6):
        _step_both(orc, gpu)
        _compare_state(orc, gpu, f"{name} N={n} R={R} step {k + 1}")


def _totals(U, ngc):
    """Interior sums of the eight conserved variables, summed exactly
    (math.fsum: a plain floating-point sum over 10^6 cells carries errors
    near 1e-11 of the total), and the sums of their moduli."""
    inner = U[ngc:-ngc, ngc:-ngc, :]
    exact = np.array([math.fsum(inner[:, :, v].ravel()) for v in range(inner.shape[2])])
    return exact, np.abs(inner).sum(axis=(0, 1))


def test_full_size_first_steps_bitwise():
    """The benchmarked workload itself, Orszag-Tang 1024^2 at R = 2: the
    first two steps, bit for bit."""
    cfg, g, orc, gpu = _pair("ot", n=1024)
    for k in range(2):
        _step_both(orc, gpu)
        _compare_state(orc, gpu, f"ot 1024^2 step {k + 1}")
    gpu.close()


def test_full_size_late_time_from_gpu_state():
    """Orszag-Tang 1024^2 on the GPU alone to t = 0.2 (shocks formed, many
    MOOD passes per stage).  The periodic totals of the conserved variables
    keep their initial values to rounding (the update is in flux form), and
    from the GPU's state the oracle and the GPU take two more steps, bit for
    bit."""
    cfg, g, U0, V0 = cases.make("ot", n=1024)
    gpu = rt.Solver(cases.runtime_config(cfg))
    gpu.upload(U0, V0)
    t = 0.0
    while t < 0.2:
        dt = rt.valid_time_step(gpu.cfl_computedt(), t, 0.2)
        gpu.advance_solution(dt)
        t += dt
    Ug, Vg, _ = gpu.download()
    ngc = cfg.R + 2
    s0, a0 = _totals(U0, ngc)
    s1, a1 = _totals(Ug, ngc)
    drift = np.abs(s1 - s0) / np.maximum(a0, 1.0)
    # the oracle's own run (OT 128^2 to t = 0.2) moves them by 1.5e-17 at most
    assert drift.max() < 1e-14, f"conserved totals drift by {drift}"
    orc = O.Oracle(cfg)
    orc.set_state(Ug, Vg)
    passes = 0
    for k in range(2):
        _step_both(orc, gpu)
        _compare_state(orc, gpu, f"ot 1024^2 from t=0.2, step {k + 1}")
        passes += int(gpu.stats()[0].sum())
    assert passes > 6, "expected MOOD passes beyond the first at t = 0.2"
    gpu.close()


def test_late_time_ot_with_heavy_mood_activity():
    """Advance the oracle to t = 0.3 (shocks fully formed), hand the same
    state to the GPU, and compare several steps with many MOOD passes."""
    cfg, g, U, V = cases.make("ot", n=64)
    orc = O.Oracle(cfg)
    orc.set_state(U, V)
    orc.run(10_000, 0.3)
    Ut, Vt = orc.get_state()
    gpu = rt.Solver(cases.runtime_config(cfg))
    gpu.upload(Ut, Vt)
    total_passes = 0
    for n in range(5):
        _step_both(orc, gpu)
        _compare_state(orc, gpu, f"ot@0.3 step {n + 1}")
        total_passes += int(gpu.stats()[0].sum())
    assert total_passes > 15, "expected MOOD passes beyond the first in most stages"


def test_calls_between_cfl_and_advance():
    """cfl_computedt starts the next step's first pass ahead of dt.  A
    download in between still returns the state and scheme of the last
    step, a second CFL the same dt, and the step that follows is the plain
    one."""
    cfg, g, orc, gpu = _pair("ot")
    for n in range(3):
        _step_both(orc, gpu)
    Uo, Vo = orc.get_state()
    last_scheme = orc.stage_schemes()[2]
    d = gpu.cfl_computedt()
    Ug, Vg, Sg = gpu.download()
    assert np.array_equal(Uo, Ug) and np.array_equal(Vo, Vg)
    assert np.array_equal(last_scheme, Sg), "scheme changed by cfl_computedt"
    assert gpu.cfl_computedt() == d == orc.cfl_dt()
    for n in range(2):
        _step_both(orc, gpu)
        _compare_state(orc, gpu, f"after cfl/download/cfl, step {n + 1}")


def test_upload_and_gp_change_after_cfl():
    """State or GP tables changed between cfl_computedt and the step: the
    first-pass work cfl_computedt started is dropped and redone."""
    cfg, g, orc, gpu = _pair("blast")
    for n in range(2):
        _step_both(orc, gpu)
    _, _, U0, V0 = cases.make("blast")  # back to the initial state
    gpu.cfl_computedt()
    orc.set_state(U0, V0)
    gpu.upload(U0, V0)
    for n in range(2):
        _step_both(orc, gpu)
        _compare_state(orc, gpu, f"after upload, step {n + 1}")
    pv, qw, _ = gpu.gp_tables()
    gpu.cfl_computedt()
    gpu.set_gp_tables(pv, qw)
    for n in range(2):
        _step_both(orc, gpu)
        _compare_state(orc, gpu, f"after set_gp, step {n + 1}")


def test_errors_are_reported():
    with pytest.raises(rt.HmhdError):
        rt.Solver(rt.Config(64, 64, R=4))
    with pytest.raises(rt.HmhdError):
        rt.Solver(rt.Config(6, 64, R=2))
    s = rt.Solver(rt.Config(32, 32))
    with pytest.raises(ValueError):
        s.upload(np.zeros((10, 10, 8)))


GOLDEN = __import__("pathlib").Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("name", sorted(p.stem[4:] for p in GOLDEN.glob("ref_*.npz")))
def test_gpu_reproduces_reference_fixture(name):
    """The reference's own Fortran step (tests/golden/make_golden.py) on the
    same inputs: every dt, the final U and V, the stage masks of the last
    step, bit for bit."""
    fx = np.load(GOLDEN / f"ref_{name}.npz")
    n = int(fx["n"])
    cfg, g, U, V = cases.make(name, n=n if n > 0 else None)
    gpu = rt.Solver(cases.runtime_config(cfg))
    gpu.upload(U, V)
    dts = fx["dts"]
    for k, dt_ref in enumerate(dts):
        dt = gpu.cfl_computedt()
        assert dt == dt_ref, f"dt of step {k + 1}"
        gpu.record_stage_schemes(k == len(dts) - 1)
        gpu.advance_solution(dt)
    Ug, Vg, Sg = gpu.download()
    assert np.array_equal(Ug, fx["U"]) and np.array_equal(Vg, fx["V"])
    assert np.array_equal(gpu.stage_schemes().astype(np.int8), fx["stage_schemes"])
    assert np.array_equal(gpu.stats()[0], fx["passes"])


@pytest.mark.parametrize("form", range(len(rt.FACE_FORMS)))
def test_every_face_kernel_form_is_bitwise(form):
    """Each tuned form of the first pass's face kernel (tile shape and wave
    target, forced) reproduces the oracle bit for bit: R = 2 Orszag-Tang on
    a grid that is not a multiple of any tile, outflow blast, R = 3 and the
    R = 1 hydro explosion."""
    try:
        rt.face_tuning(form)
        for name, n, R in (("ot", 50, 2), ("blast", 40, 2), ("ot_ref", 36, 3), ("explosion_ref", 40, 1)):
            cfg, g, orc, gpu = _pair(name, n=n, R=R)
            for k in range(3):
                _step_both(orc, gpu)
                _compare_state(orc, gpu, f"form {form}, {name} {n}^2 R={R}, step {k + 1}")
            gpu.close()
    finally:
        rt.face_tuning("tune")


def test_face_tuning_table_records_every_form():
    rt.face_tuning_clear()
    cfg, g, orc, gpu = _pair("ot", n=64)
    _step_both(orc, gpu)
    _compare_state(orc, gpu, "after tuning")
    table = rt.face_tuning_table()
    e = [x for x in table if x["R"] == 2]
    assert e and len(e[0]["ms_per_form"]) == len(rt.FACE_FORMS)
    gpu.close()


def test_fresh_handles_read_their_own_results():
    """Handles created and destroyed one after another, each with one CFL
    and one step on a state of its own: every dt and state is the oracle's.
    The host polls sequence words in pinned memory that the runtime
    recycles between handles; a word an earlier handle left there must not
    pass for the awaited one (it did when each handle counted from 1)."""
    for k in range(8):
        name, n = (("ot", 96), ("blast", 40), ("sedov", 32), ("ot", 64))[k % 4]
        cfg, g, U, V = cases.make(name, n=n)
        orc = O.Oracle(cfg)
        orc.set_state(U, V)
        with rt.Solver(cases.runtime_config(cfg)) as gpu:
            gpu.upload(U, V)
            dt = gpu.cfl_computedt()
            dt_ref = orc.cfl_dt()
            assert dt == dt_ref, f"handle {k} ({name} {n}^2): dt {dt!r}, oracle {dt_ref!r}"
            gpu.advance_solution(dt)
            orc.advance(dt_ref)
            Ug, _, _ = gpu.download(want_V=False, want_scheme=False)
            Uo, _ = orc.get_state()
            assert np.array_equal(Ug, Uo), f"handle {k} ({name} {n}^2): state after one step"
