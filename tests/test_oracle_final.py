"""The oracle through the driver loop to the final time, against the
reference build's fixture (tests/golden/final_ot128.npz): every dt and the
sha256 of the final U and V."""
import hashlib
import pathlib

import numpy as np
import pytest

import cases
import oracle as O

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def test_oracle_ot128_final_time_equals_reference_build():
    path = GOLDEN / "final_ot128.npz"
    if not path.exists():
        pytest.skip("final_ot128.npz not generated")
    fx = np.load(path)
    tmax = float(fx["tmax"])
    cfg, g, U, V = cases.make("ot", n=int(fx["n"]))
    orc = O.Oracle(cfg)
    orc.set_state(U, V)
    orc.cfl_dt()
    t, dts = 0.0, []
    while t < tmax:
        dt = orc.cfl_dt()
        if t + dt > tmax:
            dt = tmax - t
        orc.advance(dt)
        t += dt
        dts.append(dt)
    assert np.array_equal(np.array(dts), fx["dts"])
    Uo, Vo = orc.get_state()
    assert hashlib.sha256(np.ascontiguousarray(Uo).tobytes()).hexdigest() == str(fx["U_sha256"])
    assert hashlib.sha256(np.ascontiguousarray(Vo).tobytes()).hexdigest() == str(fx["V_sha256"])
