"""The oracle against the reference itself and against the reference's
fixtures (CPU only).

* golden fixtures (tests/golden/ref_*.npz, written by make_golden.py from the
  reference's Fortran step): the oracle must reproduce every dt, the final U
  and V, the per-stage scheme masks and MOOD pass counts bit for bit;
* the reference's own printed golden (test/test_reconstruct_faceValsWithGP/
  output.txt, tests/golden/reconstruct_R3_ot128.json): six significant digits;
* the reference build in oracle/_ref (when present): bit-for-bit equality
  of GP tables, dt, U, V and the per-stage masks over several steps.
"""
import json
import pathlib

import numpy as np
import pytest

import cases
import hosono_2dmhd_amd.ics as ics
import oracle as O

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
FIXTURES = sorted(p.stem[4:] for p in GOLDEN.glob("ref_*.npz"))


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_reproduces_reference_fixture(name):
    fx = np.load(GOLDEN / f"ref_{name}.npz")
    n = int(fx["n"])
    cfg, g, U, V = cases.make(name, n=n if n > 0 else None)
    import make_golden
    assert make_golden.input_digest(U, V) == str(fx["input_sha256"]), "case inputs drifted from the fixture"
    orc = O.Oracle(cfg)
    orc.set_state(U, V)
    for k, dt_ref in enumerate(fx["dts"]):
        dt = orc.cfl_dt()
        assert dt == dt_ref, f"dt of step {k + 1}: {dt!r} vs reference {dt_ref!r}"
        orc.advance(dt)
    Uo, Vo = orc.get_state()
    assert np.array_equal(Uo, fx["U"])
    assert np.array_equal(Vo, fx["V"])
    assert np.array_equal(orc.stage_schemes().astype(np.int8), fx["stage_schemes"])
    assert np.array_equal(orc.stats()[0], fx["passes"])


def test_oracle_reproduces_reference_printed_face_values():
    """R = 3, l = 6 dx, the reference's own OT IC at 128^2, every face at
    radius 3 (test/test_reconstruct_faceValsWithGP.f90:24-34)."""
    gold = json.loads((GOLDEN / "reconstruct_R3_ot128.json").read_text())
    gamma = 1.66666666666666666
    g = ics.Grid(128, 128, R=3)
    U, V = ics.orszag_tang_reference(g, gamma)
    orc = O.Oracle(O.Config(128, 128, R=3, ell_over_delta=6.0, gamma=gamma))
    fv = orc.reconstruct_all(U, 3)  # [j, i, face, q, var]
    c0 = g.ngc + 5
    faces = {"north": 0, "south": 1, "east": 2, "west": 3}
    checked = 0
    for key, cell in gold["cells"].items():
        di, dj = (int(x) for x in key.split(","))
        i, j = c0 + di, c0 + dj
        assert f"{U[j, i, 1]:.5E}" == f"{cell['average']:.5E}"
        for face, f in faces.items():
            for k in range(4):
                got = fv[j, i, f, k, 1]
                assert f"{got:.5E}" == f"{cell[face][k]:.5E}", (key, face, k, got, cell[face][k])
                checked += 1
    assert checked == 9 * 16


REF_CASES = [("ot", 32, None, 10), ("blast", 32, None, 8), ("sedov", 32, None, 6),
             ("ot_ref", 24, None, 5), ("explosion_ref", 24, None, 8), ("sedov_ref", 24, None, 5),
             ("ot_rect", None, None, 4), ("blast_rect", None, None, 4), ("ot", 24, 1, 8), ("blast", 24, 3, 4)]


@pytest.mark.skipif(not O.ref_available("exact"), reason="oracle/_ref not built (needs the reference sources)")
@pytest.mark.parametrize("name,n,R,steps", REF_CASES)
def test_oracle_bitwise_equal_to_reference_build(name, n, R, steps, capfd):
    cfg, g, U, V = cases.make(name, n=n, R=R)
    orc = O.Oracle(cfg)
    ref = O.RefHarness(cfg, "exact")
    pv1, qw1, st1 = orc.gp_tables()
    pv2, qw2, st2 = ref.gp_tables()
    assert np.array_equal(pv1, pv2) and np.array_equal(qw1, qw2) and np.array_equal(st1, st2)
    orc.set_state(U, V)
    ref.set_state(U, V)
    for k in range(steps):
        d1, d2 = orc.cfl_dt(), ref.cfl_dt()
        assert d1 == d2
        Um, masks, passes = ref.step_masks(d2)
        orc.advance(d1)
        ref.advance(d2)
        Uo, Vo = orc.get_state()
        Ur, Vr, Sr = ref.get_state()
        assert np.array_equal(Um, Ur), "re-assembled reference step differs from timeStep_advanceSolution"
        assert np.array_equal(Uo, Ur), f"U differs at step {k + 1}"
        assert np.array_equal(Vo, Vr), f"V differs at step {k + 1}"
        assert np.array_equal(orc.stage_schemes(), masks), f"stage masks differ at step {k + 1}"
        assert np.array_equal(orc.stats()[0], passes)
    ref.flush()
    capfd.readouterr()
