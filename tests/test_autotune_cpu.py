"""Round-aligned strip heights of the autotuner (host arithmetic only)."""

from __future__ import annotations

from hybrid_flood_amd.autotune import UNCAPPED_WAVES, round_rows


def test_round_rows_fill_whole_rounds():
    cus, waves = 256, 16
    segs = -(-4096 // 62)
    slots = cus * waves
    rows = round_rows((4096, 4096), waves, cus)
    assert 17 in rows and 34 in rows
    for r in (17, 23, 34):  # 4, 3 and 2 rounds
        n = segs * -(-4096 // r)  # wavefronts of the launch
        k = -(-n // slots)        # rounds it takes
        assert n > (k - 0.1) * slots, (r, n, k)  # the last round at least 90 % full
    for r in rows:  # never just over a round boundary
        n = segs * -(-4096 // r)
        assert n % slots == 0 or n % slots > slots // 4, (r, n)


def test_round_rows_bounds_and_uncapped():
    assert round_rows((4096, 4096), 0, 256) == round_rows((4096, 4096), UNCAPPED_WAVES, 256)
    assert all(4 <= r <= 64 for r in round_rows((300, 5000), 12, 256))
    assert round_rows((8, 8), 16, 256) == []
