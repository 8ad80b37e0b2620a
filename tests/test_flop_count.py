"""The instrumented FP64 operation count (hmhd_flop_count).

The counter compiles the kernels' formulas with a scalar type that counts
+ - * / sqrt.  Its totals are pinned here where the arithmetic fixes them
in closed form, for every GP radius:

  GP point values   2 faces x 2 sides x (R+1) points x n_s(R) cells x 8
                    variables x (1 mul + 1 add), n_s(R) = 1 + 2R(R+1)
  flux sums         2 faces x (R+1) points x 8 x (1 mul + 1 add)
  admissibility     2 faces x 2 sides x (R+1) x cons2pres (17: 6 for the
                    magnetic pressure, 11 for the rest)
  FLICD             4 corrections x (1 mul + 6 add/sub + 1 div)
  FV update         8 variables x (2 sub + 2 mul + 2 sub)
  RK combinations   stages 1 and 2, 8 variables, 3 operations each
  HLLC calls        2 per cell, (R+1) points each
and the per-zone total is 3 (face + update) + RK + CFL.  A counter that
dropped an operator overload (say sqrt or the unary path) or a face would
miss one of these.
"""
import json
import pathlib
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
EXE = ROOT / "hosono-2dmhd_amd" / "hmhd_flop_count"
GAMMA = 5.0 / 3.0


def _run(tmp_path, U, n, R, bc, hydro=0, windows=4, lo=0.0, hi=1.0):
    if not EXE.exists():
        pytest.skip("hmhd_flop_count is not built (make -C hosono-2dmhd_amd)")
    p = tmp_path / "U.bin"
    np.ascontiguousarray(U).tofile(p)
    r = subprocess.run([str(EXE), str(p), str(n), str(n), str(R), "12.0", repr(GAMMA), "1e-12", "1e-12",
                        str(hydro), str(bc), repr(lo), repr(hi), repr(lo), repr(hi), "0.8", str(windows), "7"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def _ot(n, R):
    import hosono_2dmhd_amd.ics as ics
    g = ics.Grid(n, n, 0.0, 1.0, 0.0, 1.0, R=R)
    U, _ = ics.orszag_tang(g, GAMMA)
    return U


@pytest.mark.parametrize("R", [1, 2, 3])
def test_closed_form_terms(tmp_path, R):
    n = 64
    out = _run(tmp_path, _ot(n, R), n, R, bc=0)
    z = out["per_zone"]
    nq, ns = R + 1, 1 + 2 * R * (R + 1)
    assert z["recon"] == 2 * 2 * nq * ns * 8 * 2
    assert z["flux_sums"] == 2 * nq * 16
    assert z["admissibility"] == 2 * 2 * nq * 17
    assert z["flicd"] == 4 * 8
    assert z["fv_update"] == 8 * 6
    assert z["rk_per_step"] == 2 * 8 * 3
    assert out["hllc_calls_per_zone"] == 2 * nq
    # the face total is its parts
    assert z["face"] == pytest.approx(z["recon"] + z["admissibility"] + z["hllc"] + z["flux_sums"], rel=1e-12)
    assert z["update"] == pytest.approx(z["flicd"] + z["fv_update"] + z["detect"], rel=1e-12)
    total = 3 * (z["face"] + z["update"]) + z["rk_per_step"] + z["cfl"]
    assert out["flops_per_zone_update"] == pytest.approx(total, rel=1e-6)


# HLLC, counted by hand from the formulas (src/riemannSolver.f90:68-221):
# two cons2prim 2 x 21, the averaged state 8 x 2, three fast speeds 3 x 21,
# sL and sR 4, two physical fluxes 2 x 33, the HLL state 8 x 7, its
# velocities and the three B.v 3 + 15, the total pressures 2 x 7, dL and dR
# 2, S* 13 and p* 8: 302 on every face; a star state adds 53 (its density
# 3, normal momentum 1, transverse momenta 2 x 7, energy 11, flux 8 x 3).
HLLC_COMMON, HLLC_STAR = 302, 53
# detection up to the shock indicators (src/detect.f90:83-130): cons2pres
# of the new state 17, five cons2prim 105, div v with its threshold 9, the
# Mach number 7, the pressure jump with its threshold 6
DETECT_SHOCK_TESTS = 144


def test_hllc_and_cfl_counts(tmp_path):
    """Every HLLC call counts between the supersonic case (302) and a star
    state (355); CFL (MHD): cons2prim 21 + BB 5 + eigen_cfs 16 + |v| + c
    twice 2 = 44."""
    n = 64
    out = _run(tmp_path, _ot(n, 2), n, 2, bc=0)
    assert HLLC_COMMON <= out["hllc_flops_per_call"] <= HLLC_COMMON + HLLC_STAR
    assert out["per_zone"]["cfl"] == 44.0
    assert out["per_zone"]["detect"] >= DETECT_SHOCK_TESTS


@pytest.mark.parametrize("bc", [0, 1])
def test_uniform_state_counts_exactly(tmp_path, bc):
    """A uniform state at rest: every face is a resting contact (S* = 0,
    the right star state), and detection stops after the shock indicators."""
    import hosono_2dmhd_amd.ics as ics
    n, R = 64, 2
    g = ics.Grid(n, n, 0.0, 1.0, 0.0, 1.0, R=R)
    V = np.zeros((n + 2 * (R + 2), n + 2 * (R + 2), 10))
    V[..., 0] = 1.0
    V[..., 4] = 1.0
    V[..., 5] = 0.3
    U = ics.prim2cons(V, GAMMA)
    out = _run(tmp_path, U, n, R, bc=bc)
    assert out["hllc_flops_per_call"] == HLLC_COMMON + HLLC_STAR
    assert out["per_zone"]["detect"] == DETECT_SHOCK_TESTS
