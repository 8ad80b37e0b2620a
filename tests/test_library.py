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


def test_fortran_binding_compiles_and_links(tmp_path):
    """The iso_c_binding module a Fortran driver uses compiles with the
    image's Fortran compiler and a program using it links against the
    library (nothing is run: that needs a device)."""
    import shutil
    import subprocess
    fc = shutil.which("amdflang")
    if fc is None:
        pytest.skip("no Fortran compiler")
    mod = ROOT / "hosono-2dmhd_amd" / "fortran" / "hmhd_c.f90"
    prog = tmp_path / "use_it.f90"
    prog.write_text(
        "program use_it\n"
        "  use hmhd_c\n"
        "  use, intrinsic :: iso_c_binding\n"
        "  type(hmhd_config) :: cfg\n"
        "  type(c_ptr) :: h\n"
        "  real(c_double) :: dt\n"
        "  integer(c_int32_t) :: rc\n"
        "  cfg%nx = 64; cfg%ny = 64; cfg%R = 2\n"
        "  rc = hmhd_create(cfg, h)\n"
        "  if (rc /= HMHD_OK) print *, hmhd_error_message(c_null_ptr)\n"
        "  rc = hmhd_cfl_dt(h, dt)\n"
        "  rc = hmhd_advance(h, dt)\n"
        "  call hmhd_destroy(h)\n"
        "end program use_it\n")
    lib = rt.LIB_PATH
    r = subprocess.run([fc, "-J", str(tmp_path), str(mod), str(prog), "-o", str(tmp_path / "use_it"),
                        f"-L{lib.parent}", "-lhmhd", f"-Wl,-rpath,{lib.parent}"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert (tmp_path / "use_it").exists()
