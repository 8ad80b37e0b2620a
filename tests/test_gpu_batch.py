"""The batched and served forms: many runs of the driver loop
(src/mhd_driver.f90:76-95) advanced together in one engine.

Each run must come out exactly as it does on its own: the same dt
sequence, the same final U and V, bit for bit with the oracle's driver loop
(Oracle.run, which clamps the last dt to tmax as validTimeStep does), however
the runs are mixed, whenever they join or leave, and whatever else shares
the device.  Sizes are deliberately not multiples of the 16-cell tile.
"""
import functools
import time

import numpy as np
import pytest

import cases
import hosono_2dmhd_amd.runtime as rt
import oracle as O

pytestmark = pytest.mark.gpu

# (case, n, R, tmax, nstepmax): endings by tmax (the last dt clamped) and by
# step count, so runs leave the batch at different engine steps
RUNS = [
    ("ot", 20, 2, 0.02, 0),
    ("blast", 40, 2, 1.0, 6),
    ("sedov", 36, 3, 1.0, 8),
    ("explosion_ref", 30, 1, 1.0, 5),
    ("ot_rect", None, 2, 1.0, 4),
    ("blast_rect", None, 2, 1.0, 7),
    ("ot", 33, 2, 1.0, 9),
]


@functools.lru_cache(maxsize=None)
def expected(k):
    name, n, R, tmax, nmax = RUNS[k]
    cfg, g, U, V = cases.make(name, n=n, R=R)
    orc = O.Oracle(cfg)
    orc.set_state(U, V)
    t, steps = orc.run(nmax if nmax > 0 else 1_000_000, tmax)
    Uo, Vo = orc.get_state()
    return t, steps, Uo, Vo


def initial(k):
    name, n, R, tmax, nmax = RUNS[k]
    cfg, g, U, V = cases.make(name, n=n, R=R)
    return cases.runtime_config(cfg), U, V, tmax, nmax


def _check(k, t, steps, U, V, where):
    te, se, Ue, Ve = expected(k)
    assert steps == se, f"{where}: {steps} steps, oracle {se}"
    assert t == te, f"{where}: t {t!r}, oracle {te!r}"
    assert np.array_equal(U, Ue), f"{where}: U differs (max {np.abs(U - Ue).max():.3e})"
    assert np.array_equal(V, Ve), f"{where}: V differs"


@pytest.mark.parametrize("chunk", [0, 3])
def test_batch_runs_match_the_oracle_driver_loop(chunk):
    """All runs in one batch, to the end at once (chunk 0) or three engine
    steps per call; every run equals its oracle run."""
    with rt.Batch() as b:
        ids = []
        for k in range(len(RUNS)):
            cfg, U, V, tmax, nmax = initial(k)
            ids.append(b.add(cfg, U, V, tmax, nmax))
        if chunk:
            for _ in range(100):
                b.run(chunk)
                if all(b.status(i)[2] != rt.RUN_ACTIVE for i in ids):
                    break
        else:
            b.run(0)
        for k, i in enumerate(ids):
            t, steps, state = b.status(i)
            assert state == rt.RUN_DONE
            U, V, S = b.download(i)
            _check(k, t, steps, U, V, f"run {k} {RUNS[k][0]}")


def test_runs_added_while_others_are_midway():
    """A run that joins after others have stepped, and one added after
    others have finished, still match their solo results."""
    with rt.Batch() as b:
        a = b.add(*initial(1))
        b.run(2)
        c = b.add(*initial(2))
        b.run(0)
        d = b.add(*initial(0))
        b.run(0)
        for k, i in ((1, a), (2, c), (0, d)):
            t, steps, state = b.status(i)
            U, V, _ = b.download(i)
            _check(k, t, steps, U, V, f"run {k}")


def test_batch_equals_solo_solver():
    """The single-run handle and a batch of one give the same bits."""
    cfg, U, V, tmax, nmax = initial(6)
    s = rt.Solver(cfg)
    s.upload(U, V)
    t, n = rt.run(s, tmax, nmax)
    Us, Vs, Ss = s.download()
    with rt.Batch() as b:
        i = b.add(cfg, U, V, tmax, nmax)
        b.run(0)
        Ub, Vb, Sb = b.download(i)
        assert b.status(i)[:2] == (t, n)
    assert np.array_equal(Us, Ub) and np.array_equal(Vs, Vb) and np.array_equal(Ss, Sb)


def test_nan_state_aborts_only_that_run():
    """A NaN dt stops its run before any step (src/mhd_driver.f90:134-138);
    the other runs of the batch go on."""
    cfg, U, V, tmax, nmax = initial(3)
    Ubad = U.copy()
    Ubad[cfg.R + 2 + 3, cfg.R + 2 + 4, 0] = np.nan
    with rt.Batch() as b:
        bad = b.add(cfg, Ubad, None, tmax, nmax)
        good = b.add(cfg, U, V, tmax, nmax)
        b.run(0)
        assert b.status(bad)[1:] == (0, rt.RUN_ABORTED)
        t, steps, state = b.status(good)
        U2, V2, _ = b.download(good)
        _check(3, t, steps, U2, V2, "run beside the aborted one")


def test_server_admits_by_budget_and_results_match_alone():
    """Requests wait for HBM budget, run mixed with others, and give the
    same bits as the same request served alone; the budget is never
    exceeded and a request above it fails at once."""
    fp = [rt.footprint(initial(k)[0]) for k in range(len(RUNS))]
    budget = 2 * max(fp)
    with rt.Server(budget) as srv:
        reqs = {}
        for k in range(len(RUNS)):
            reqs[k] = srv.submit(*initial(k))
            time.sleep(0.002 * (k % 3))
        results = {k: srv.wait(r, timeout_s=120.0) for k, r in reqs.items()}
        st = srv.stats()
    assert st["completed"] == len(RUNS) and st["peak_bytes"] <= budget
    assert 1 <= st["peak_active"] < len(RUNS)
    for k, res in results.items():
        assert res.state == rt.REQ_DONE
        assert res.submit_s <= res.admit_s <= res.done_s
        _check(k, res.t, res.steps, res.U, res.V, f"request {k} in the mix")
    # the same request on a server of its own
    with rt.Server(budget) as srv:
        alone = srv.wait(srv.submit(*initial(2)), timeout_s=120.0)
    assert np.array_equal(alone.U, results[2].U) and np.array_equal(alone.V, results[2].V)
    assert np.array_equal(alone.scheme, results[2].scheme)
    with rt.Server(fp[0] // 2) as srv:
        res = srv.wait(srv.submit(*initial(0)), timeout_s=60.0)
        assert res.state == rt.REQ_FAILED


def test_server_with_one_slot_runs_requests_in_order():
    with rt.Server(1 << 34, max_active=1) as srv:
        r = [srv.submit(*initial(k)) for k in (3, 4)]
        res = [srv.wait(x, timeout_s=120.0) for x in r]
        assert srv.stats()["peak_active"] == 1
    assert res[0].done_s <= res[1].admit_s
    for k, x in zip((3, 4), res):
        _check(k, x.t, x.steps, x.U, x.V, f"request {k}")


def test_server_with_many_distinct_gp_table_sets():
    """Ten runs whose GP tables all differ (l / dx from 6 to 15, four
    aspect ratios dy / dx) through one server at once: each equals its
    oracle run bit for bit."""
    import hosono_2dmhd_amd.ics as ics
    runs = []
    for k in range(10):
        nx, ny, yend = 24, 20 + 2 * (k % 4), 1.0 + 0.1 * (k % 4)
        g = ics.Grid(nx, ny, 0.0, 1.0, 0.0, yend, R=2)
        U, V = ics.orszag_tang(g)
        cfg = O.Config(nx, ny, 0.0, 1.0, 0.0, yend, R=2, ell_over_delta=6.0 + k)
        runs.append((cfg, U, V))
    tables = {cases.runtime_config(c).ell_over_delta for c, _, _ in runs}
    assert len(tables) == 10
    with rt.Server(1 << 34) as srv:
        reqs = [srv.submit(cases.runtime_config(c), U, V, 1.0, 5) for c, U, V in runs]
        res = [srv.wait(r, timeout_s=120.0) for r in reqs]
    for k, ((c, U, V), x) in enumerate(zip(runs, res)):
        orc = O.Oracle(c)
        orc.set_state(U, V)
        t, steps = orc.run(5, 1.0)
        Uo, Vo = orc.get_state()
        assert x.state == rt.REQ_DONE, f"run {k}: state {x.state}"
        assert (x.t, x.steps) == (t, steps), f"run {k}"
        assert np.array_equal(x.U, Uo) and np.array_equal(x.V, Vo), f"run {k}: state differs from the oracle"
