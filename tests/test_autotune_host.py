"""Autotuner host logic (no GPU): the candidate space and the size-class key."""
from pixelflow_amd import autotune


def test_candidates_cover_tiles_and_grids():
    c = autotune.candidates(60, False)
    assert {x["tile_x"] for x in c} == {32, 64, 128}
    assert {x["persistent_blocks"] for x in c} == {0, 512, 1024, 2048}
    assert all(x["inner_iters"] == 0 for x in c)


def test_mixed_schedules_have_at_least_min_cycles():
    for it in (30, 60, 100):
        ks = {x["inner_iters"] for x in autotune.candidates(it, True, min_cycles=2)}
        assert ks and all(-(-it // k) >= 2 for k in ks)


def test_size_class_key_buckets():
    a = autotune.size_class_key(2, "uniform", True, 60, [1000] * 500)
    b = autotune.size_class_key(2, "uniform", True, 60, [1010] * 505)
    c = autotune.size_class_key(2, "uniform", False, 60, [1000] * 500)
    assert a == b and a != c
