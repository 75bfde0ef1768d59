"""Tile autotuner and the occupancy cap: tuning never changes a result.

Strip height and the resident-wavefront cap only change which wavefront
computes a cell and how many run at once, so every candidate must give the
same states bit for bit (the float64 volume sums may differ in order only).
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import random_case
from test_gpu_step import _params, _state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", ["fast", "reference"])
def test_waves_cap_and_rows_do_not_change_states(cuda, variant):
    from hybrid_flood_amd import numerics as nm

    case = random_case(31, (150, 200), open_sides=(0, 1, 1, 0), speed=0.8)
    params = _params(case, cuda)
    ctx = nm.context_for(params, variant)
    times = np.array([0.0, 0.3, 0.6], np.float32)
    results = []
    for rows, waves in ((16, 0), (5, 16), (32, 8), (16, 20)):
        ctx.set_tile_rows(rows)
        ctx.set_waves_per_cu(waves)
        results.append(nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=500,
                                               variant=variant))
    ctx.set_tile_rows(0)
    ctx.set_waves_per_cu(0)
    base = results[0]
    for r in results[1:]:
        assert r.steps_used == base.steps_used
        for g, w in zip(r.states, base.states):
            assert np.array_equal(g.cpu().numpy(), w.cpu().numpy())
        assert r.diagnostics["volume_final"] == pytest.approx(base.diagnostics["volume_final"],
                                                              rel=1e-12)


def test_waves_cap_rejects_bad_values(cuda):
    from hybrid_flood_amd import numerics as nm

    case = random_case(2, (20, 20))
    ctx = nm.context_for(_params(case, cuda))
    for bad in (-4, 3, 4, 10):
        with pytest.raises(Exception):
            ctx.set_waves_per_cu(bad)
    ctx.set_waves_per_cu(0)


def test_tune_rows_records_and_applies(cuda, tmp_path, monkeypatch):
    from hybrid_flood_amd import autotune
    from hybrid_flood_amd import numerics as nm

    monkeypatch.setenv("SWE_TUNE_CACHE", str(tmp_path / "tune.json"))
    case = random_case(5, (96, 130))
    params = _params(case, cuda)
    rec = autotune.tune_rows(params, _state(case, cuda), steps=8, candidates=(8, 16),
                             waves=(0, 16))
    assert rec["rows"] in (8, 16) and rec["waves_per_cu"] in (0, 16)
    assert len(rec["candidates"]) == 4
    # the winner comes from the longer re-timing of the fastest candidates
    assert len(rec["finalists"]) == 4
    assert rec["ms_per_step"] == min(rec["finalists"].values())
    assert rec["ms_per_step"] == rec["finalists"][f"{rec['rows']}x{rec['waves_per_cu']}"]
    assert (tmp_path / "tune.json").exists()
    again = autotune.tune_rows(params, _state(case, cuda), steps=8, candidates=(8, 16),
                               waves=(0, 16))
    assert again == rec  # served from the cache
    # the tuned context still matches an untuned one bit for bit
    times = np.array([0.0, 0.2], np.float32)
    tuned = nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=200)
    ctx = nm.context_for(params)
    ctx.set_tile_rows(0)
    ctx.set_waves_per_cu(0)
    plain = nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=200)
    for g, w in zip(tuned.states, plain.states):
        assert np.array_equal(g.cpu().numpy(), w.cpu().numpy())
