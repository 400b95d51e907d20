"""Host logic of the launch-shape autotuner: the candidate shapes stay inside
the kernels' launch limits, selection and the JSON cache (no device)."""

import json

import pytest

from blackhorizon_amd import autotune
from blackhorizon_amd.tracer import Tune


@pytest.mark.parametrize("cands", [autotune.plain_candidates(), autotune.batch_candidates(),
                                   autotune.batch_candidates(quad=True)])
def test_candidates_within_launch_limits(cands):
    assert cands[0] == Tune()  # the default wins ties
    assert len(set(cands)) == len(cands)
    for t in cands:
        assert t.block_size % 64 == 0 and 64 <= t.block_size <= 256
        assert 1 <= t.refill_min <= 64 // t.lanes_per_ray
        assert t.rays_per_lane >= 1
        assert t.lanes_per_ray in (1, 4)


def test_batch_candidates_have_one_ray_per_lane():
    assert all(t.rays_per_lane == 1 for t in autotune.batch_candidates())
    assert all(t.lanes_per_ray == 1 for t in autotune.batch_candidates())
    quads = [t for t in autotune.batch_candidates(quad=True) if t.lanes_per_ray == 4]
    assert quads and all(t.refill_min <= 16 for t in quads)
    assert all(t.lanes_per_ray == 1 for t in autotune.plain_candidates())


def test_select_fastest_earliest_on_ties():
    a, b, c = Tune(64, 1, 1), Tune(128, 1, 1), Tune(256, 2, 16)
    assert autotune.select([(a, 3.0), (b, 2.0), (c, 2.0)]) == (b, 2.0)
    assert autotune.select([(c, 1.0)]) == (c, 1.0)
    with pytest.raises(ValueError):
        autotune.select([])


def test_time_candidates_uses_the_timer_per_candidate():
    calls = []

    def timer(plan, tune, reps, warmup):
        calls.append((plan, tune, reps, warmup))
        return 10.0 / tune.block_size

    cands = autotune.plain_candidates()
    table = autotune.time_candidates("plan", cands, reps=4, warmup=2, timer=timer)
    assert [t for t, _ in table] == cands
    assert all(c == ("plan", t, 4, 2) for c, t in zip(calls, cands))
    best, ms = autotune.select(table)
    assert best.block_size == 256 and ms == pytest.approx(10.0 / 256)


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "sub" / "tune.json")
    assert autotune.load_cache(path) == {}
    r = autotune.TuneResult("k1", Tune(128, 2, 16), 1.5, [({"block_size": 128}, 1.5)])
    autotune.store_cache(path, "k1", r)
    autotune.store_cache(path, "k2", autotune.TuneResult("k2", Tune(), 2.0))
    data = autotune.load_cache(path)
    assert set(data) == {"k1", "k2"}
    assert Tune(**data["k1"]["best"]) == Tune(128, 2, 16)
    assert data["k1"]["best_ms"] == 1.5
    # a corrupt cache reads as empty and is replaced on the next store
    with open(path, "w") as f:
        f.write("{not json")
    assert autotune.load_cache(path) == {}
    autotune.store_cache(path, "k3", r)
    assert list(json.load(open(path))) == ["k3"]
