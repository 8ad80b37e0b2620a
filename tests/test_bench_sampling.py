"""The bench's choice of timed steps (--steps K): one step from each K-th of
the solve, the one whose MOOD pass count is nearest that stretch's mean."""
import importlib.util
import pathlib
import random

ROOT = pathlib.Path(__file__).resolve().parents[1]


def _bench():
    spec = importlib.util.spec_from_file_location("hmhd_bench", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_one_step_per_stretch_nearest_its_mean():
    b = _bench()
    rng = random.Random(7)
    passes = [1] * 40 + [rng.randint(3, 40) for _ in range(960)]
    want = b.representative_steps(passes, 20)
    assert len(want) == 20 and want == sorted(want)
    for k, i in enumerate(want):
        a, e = k * 1000 // 20, (k + 1) * 1000 // 20
        assert a <= i < e
        m = sum(passes[a:e]) / (e - a)
        assert abs(passes[i] - m) == min(abs(p - m) for p in passes[a:e])
        # the first such step of its stretch
        assert all(abs(passes[j] - m) > abs(passes[i] - m) for j in range(a, i))


def test_sample_mean_tracks_the_solve():
    b = _bench()
    rng = random.Random(11)
    # a solve whose passes grow over time with large step-to-step swings
    passes = [max(1, int(1 + 0.04 * j + rng.gauss(0, 6))) for j in range(1375)]
    want = b.representative_steps(passes, 20)
    mean_all = sum(passes) / len(passes)
    mean_sample = sum(passes[i] for i in want) / len(want)
    assert abs(mean_sample - mean_all) / mean_all < 0.05


def test_fewer_steps_than_samples():
    b = _bench()
    assert b.representative_steps([3, 1, 2], 5) == [0, 1, 2]
    assert b.representative_steps([], 4) == []
