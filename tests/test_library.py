"""The C-ABI library on the CPU: it loads, exports every entry point
include/hmhd.h declares, builds the same GP tables as the oracle without a
device, and rejects configurations the reference stops on.  Also the
driver-side helpers (guard fills, initial conditions)."""
import ctypes as C
import pathlib
import re

import numpy as np
import pytest

import cases
import hosono_2dmhd_amd.ics as ics
import hosono_2dmhd_amd.runtime as rt
import oracle as O

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_entry_points():
    text = (ROOT / "include" / "hmhd.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|uint64_t|const char\*)\s+(hmhd_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_entry_point():
    names = declared_entry_points()
    assert len(names) >= 15
    lib = C.CDLL(str(rt.LIB_PATH))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    assert set(names) == set(rt.EXPORTS)


@pytest.mark.parametrize("name", ["ot", "ot_ref", "explosion_ref", "sedov_ref", "ot_rect", "blast_rect"])
def test_gp_tables_computed_on_the_host_match_the_oracle(name):
    cfg, g, U, V = cases.make(name)
    a = O.Oracle(cfg).gp_tables()
    b = rt.gp_compute(cases.runtime_config(cfg))
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("bad,fragment", [
    (dict(R=4), "GP_maxRadius"),
    (dict(R=0), "GP_maxRadius"),
    (dict(nx=7), "at least"),
    (dict(xend=0.0), "domain"),
    (dict(bc=7), "boundary condition"),
    (dict(gamma=1.0), "sim_gamma"),
])
def test_invalid_configurations_are_rejected_before_any_device_work(bad, fragment):
    kw = dict(nx=32, ny=32)
    kw.update(bad)
    cfg = rt.Config(**kw)
    with pytest.raises(rt.HmhdError) as e:
        rt.gp_compute(cfg)
    assert fragment in str(e.value)
    with pytest.raises(rt.HmhdError) as e:
        rt.Solver(cfg)
    assert fragment in str(e.value) and e.value.code == 1


def test_valid_time_step_follows_the_driver():
    assert rt.valid_time_step(0.1, 0.45, 0.5) == pytest.approx(0.05)
    assert rt.valid_time_step(0.01, 0.0, 0.5) == 0.01
    for bad in (-1.0, float("nan")):
        with pytest.raises(rt.TimeStepError):
            rt.valid_time_step(bad, 0.0, 0.5)
    # the clamp comes first (src/mhd_driver.f90:124-126): an infinite dt is
    # clamped to tmax - t, never reported
    assert rt.valid_time_step(float("inf"), 0.2, 0.5) == pytest.approx(0.3)


@pytest.mark.parametrize("bc", [ics.BC_PERIODIC, ics.BC_OUTFLOW])
def test_guard_fill_matches_the_oracle(bc):
    g = ics.Grid(20, 12, R=2)
    o = O.Oracle(O.Config(20, 12, R=2, bc=bc))
    A = np.zeros((g.my, g.mx, 8))
    js, is_ = g.interior()
    A[js, is_] = np.random.default_rng(5).normal(size=(12, 20, 8))
    B = A.copy()
    ics.bc_fill(A, g, bc)
    o.lib.orc_bc_fill(o.h, B.ctypes.data_as(C.POINTER(C.c_double)), 8)
    assert np.array_equal(A, B)


def test_initial_conditions_are_well_formed():
    for name in ("ot", "blast", "sedov", "ot_ref", "explosion_ref", "sedov_ref"):
        cfg, g, U, V = cases.make(name)
        js, is_ = g.interior()
        assert np.all(U[js, is_, 0] > 0) and np.all(V[js, is_, 4] > 0)
        o = O.Oracle(cfg)
        for j, i in ((g.ngc, g.ngc), (g.my // 2, g.mx // 2)):
            assert np.allclose(o.prim2cons(V[j, i]), U[j, i], rtol=1e-15, atol=0)
    # Sedov: unit thermal energy in the central 3 x 3 cells
    g = ics.Grid(32, 32, -0.5, 0.5, -0.5, 0.5, R=2)
    U, V = ics.sedov(g, gamma=5.0 / 3.0, energy=1.0, p_bg=0.0, B=(0.0, 0.0, 0.0))
    js, is_ = g.interior()
    assert (U[js, is_, 4].sum() * g.dx * g.dy) == pytest.approx(1.0, rel=1e-14)
