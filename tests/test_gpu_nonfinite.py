"""A non-finite value in the state stops the device loop and is reported.

The reference's validator rejects non-finite outputs and negative depths
(scripts/10_validate_gpu_results.py:139-146, scripts/07_run_hybrid_forecast.py:
71-79); here the loop itself detects them.  NaN is data in these tests, not a
GPU fault: every kernel still ends normally.

* one NaN cell in the initial depth: the loop start sees it, no step is taken,
  swe_run_to_outputs returns SWE_ENONFINITE with info.nonfinite = 1, and
  ShallowWaterSolver.run raises RuntimeError;
* a NaN Manning coefficient on one wet cell: the first step writes a NaN
  momentum there, the next step kernel sees it in the CFL slots and the loop
  stops after exactly one step (both kernel variants);
* one ensemble member with a NaN cell: that member stops at once with its own
  nonfinite flag; the others run to t_end and equal their solo runs.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from conftest import random_case
from test_gpu_step import _params, _state

pytestmark = pytest.mark.gpu


def _run_cabi(case, cuda, variant, times, max_steps=2000):
    from hybrid_flood_amd import _lib
    from hybrid_flood_amd import numerics as nm

    params = _params(case, cuda)
    ctx = nm.context_for(params, variant)
    st = _state(case, cuda)
    import torch

    T = len(times)
    cube = [torch.empty((T,) + tuple(case["h"].shape), dtype=torch.float32, device=cuda)
            for _ in range(3)]
    info = _lib.SweRunInfo()
    t64 = np.asarray(times, np.float64)
    rc = _lib.load().swe_run_to_outputs(
        ctx.handle, _lib.state_struct(*st), t64.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), T,
        max_steps, cube[0].data_ptr(), cube[1].data_ptr(), cube[2].data_ptr(),
        _lib.SweState(None, None, None), ctypes.byref(info), _lib.stream_handle(cuda))
    return rc, info, cube, ctx


@pytest.mark.parametrize("variant", ["fast", "reference"])
def test_nan_in_initial_depth_stops_at_once(cuda, variant):
    from hybrid_flood_amd import _lib

    case = random_case(61, (40, 70), open_sides=(0, 1, 0, 0), speed=0.5)
    case["h"][12, 33] = np.nan
    print(f"nan initial depth, variant {variant}", flush=True)
    rc, info, _, _ = _run_cabi(case, cuda, variant, [0.0, 1.0, 2.0])
    assert rc == _lib.SWE_ENONFINITE
    assert info.nonfinite == 1
    assert info.steps_used == 0 and info.outputs_written == 1
    assert "non-finite" in _lib.last_error()


def test_solver_run_raises_on_nan_state(cuda):
    from hybrid_flood_amd.shallow_water_2d import ShallowWaterSolver, StructuredGrid

    case = random_case(62, (30, 45), speed=0.5)
    grid = StructuredGrid(bed=case["bed"], manning_n=case["manning_n"],
                          domain_mask=case["domain_mask"], resolution_m=case["dx"])
    solver = ShallowWaterSolver(grid, rainfall_times_s=case["rain_t"],
                                rainfall_rates_m_s=case["rain_r"], device=cuda)
    st = solver.initial_state(depth_m=0.2)
    st.hu[5, 7] = float("nan")
    with pytest.raises(RuntimeError, match="non-finite"):
        solver.run(st, np.array([0.0, 1.0], np.float32), max_steps=500)


@pytest.mark.parametrize("variant,palette", [("fast", False), ("reference", False),
                                             ("fast", True)])
def test_nan_manning_stops_after_the_step_that_made_it(cuda, variant, palette):
    """palette: two Manning values and a uniform multiplier, so the fast loop
    stores Manning classes and the NaN is a class of its own."""
    from hybrid_flood_amd import _lib

    case = random_case(63, (48, 80), open_sides=(1, 0, 0, 0), speed=0.5)
    if palette:
        case["manning_n"] = np.where(case["manning_n"] > 0.03, 0.035, 0.015).astype(np.float32)
        case["rainfall_multiplier"] = case["domain_mask"].astype(np.float32)
    wet = np.argwhere(case["domain_mask"] & (case["h"] > 0.05))
    j, i = wet[len(wet) // 2]
    case["manning_n"] = case["manning_n"].copy()
    case["manning_n"][j, i] = np.nan
    print(f"nan manning at ({j}, {i}), variant {variant}, palette {palette}", flush=True)
    rc, info, cube, ctx = _run_cabi(case, cuda, variant, [0.0, 50.0])
    assert rc == _lib.SWE_ENONFINITE
    assert info.nonfinite == 1
    assert info.steps_used == 1
    assert info.final_time_s > 0.0
    if palette:
        assert ctx.static_layout() == 2


def test_ensemble_member_with_nan_stops_alone(cuda):
    import torch

    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd.ensemble import EnsembleRunner, Member

    cases = [random_case(70 + k, s, open_sides=(0, 1, 0, 0), speed=0.8)
             for k, s in enumerate([(40, 70), (33, 90), (64, 64)])]
    cases[1]["h"][10, 20] = np.nan
    members = []
    for c in cases:
        members.append(Member(_params(c, cuda), _state(c, cuda)))
    solo = [nm.integrate_to_outputs(_state(c, cuda), _params(c, cuda),
                                    np.array([0.0, 3.0], np.float32), max_steps=5000,
                                    variant="fast") for c in (cases[0], cases[2])]
    ens = EnsembleRunner(members, variant="fast")
    print("ensemble with one nan member", flush=True)
    ens.begin(3.0, max_steps=5000)
    for _ in range(400):
        ens.advance(16)
        if ens.running() == 0:
            break
        ens.compact()
    rc, infos = ens.end()
    from hybrid_flood_amd import _lib

    assert rc == _lib.SWE_ENONFINITE
    assert infos[1]["nonfinite"] == 1 and infos[1]["steps_used"] == 0
    for k, r in ((0, solo[0]), (2, solo[1])):
        assert infos[k]["nonfinite"] == 0
        assert infos[k]["final_time_s"] == pytest.approx(3.0, abs=1e-9)
        assert infos[k]["steps_used"] == r.steps_used
        for got, want in zip(members[k].state, r.states):
            assert torch.equal(got, want[-1])
