"""hmhd_cfl_dt_noalloc, the entry point that allocates nothing, against the
entry point it varies, hmhd_cfl_dt: the same dt, bit for bit, on two calls
in a row (the second on a smaller grid, which reads fewer of the shared
scratch array's partials), from uploaded primitives and from U alone, MHD
and hydro, at the start of a run and after steps."""
import numpy as np
import pytest
import torch

import cases
import hosono_2dmhd_amd.runtime as rt

pytestmark = pytest.mark.gpu


def _noalloc_dt(rcfg, U, V):
    Ud = torch.from_numpy(np.ascontiguousarray(U)).cuda()
    Vd = torch.from_numpy(np.ascontiguousarray(V)).cuda() if V is not None else None
    dt = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream()
    torch.cuda.synchronize()
    rt.cfl_dt_noalloc(rcfg, Ud.data_ptr(), Vd.data_ptr() if Vd is not None else None, dt.data_ptr(),
                      st.cuda_stream)
    torch.cuda.synchronize()
    return float(dt.item())


@pytest.mark.parametrize("cases_seq", [
    (("ot", 96), ("blast", 40)),            # second call on a smaller grid
    (("explosion_ref", 48), ("sedov", 24)),  # hydro, then MHD, smaller
])
def test_noalloc_cfl_equals_hmhd_cfl_dt(cases_seq):
    for name, n in cases_seq:
        cfg, g, U, V = cases.make(name, n=n)
        rcfg = cases.runtime_config(cfg)
        with rt.Solver(rcfg) as s:
            s.upload(U, V)
            want_v = s.cfl_computedt()   # reads the uploaded primitives
            s.upload(U, None)
            want_u = s.cfl_computedt()   # derives them from U
            for _ in range(3):
                s.advance_solution(s.cfl_computedt())
            Us, _, _ = s.download(want_V=False, want_scheme=False)
            want_late = s.cfl_computedt()
        assert _noalloc_dt(rcfg, U, V) == want_v, f"{name} {n}: with V"
        assert _noalloc_dt(rcfg, U, None) == want_u, f"{name} {n}: from U"
        assert _noalloc_dt(rcfg, Us, None) == want_late, f"{name} {n}: after three steps"


def test_noalloc_rejects_oversized_grids():
    cfg = rt.Config(40000, 40000)
    dummy = torch.zeros(1, dtype=torch.float64, device="cuda")
    with pytest.raises(rt.HmhdError, match="HMHD_NOALLOC_MAX_CELLS"):
        rt.cfl_dt_noalloc(cfg, dummy.data_ptr(), None, dummy.data_ptr())
