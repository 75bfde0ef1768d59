"""GPU parity of the fused step, the CFL reduction and the device loop.

Bars (stated once, used throughout):
  * reference-order kernel vs the FP32 oracle, one step from identical inputs:
    depth and wet mask bit-identical; hu, hv within 2 ulp (the only
    difference is float32 pow: ROCm's vs the host libm's);
  * either kernel vs the FP64 oracle, one step: |x - x64| <= 2e-7 + 2e-5 |x64|
    for h, hu, hv (the reference's own V1/V2 bar,
    tests/test_v1_v2_architecture.py:104);
  * V1 golden vectors: depth bit-identical, hu/hv at that same bar.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden, random_case

pytestmark = pytest.mark.gpu

RTOL, ATOL = 2.0e-5, 2.0e-7


def _torch():
    import torch

    return torch


def _ulp_diff(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


def _params(case, dev, **kw):
    from hybrid_flood_amd import numerics as nm

    return nm.make_params(
        case["bed"], case["manning_n"], case["domain_mask"], case.get("rainfall_multiplier"),
        boundary_types=case.get("boundary_types", (0, 0, 0, 0)), rainfall_times_s=case["rain_t"],
        rainfall_rates_m_s=case["rain_r"], dx=case["dx"], dry_tolerance=case.get("dry", 1e-4),
        cfl=case.get("cfl", 0.45), max_dt=case.get("max_dt", 10.0), device=dev, **kw)


def _oparams(case, dtype=np.float32):
    from oracle import swe_oracle as o

    return o.OracleParams(
        bed=case["bed"], manning_n=case["manning_n"], domain_mask=case["domain_mask"],
        rainfall_multiplier=case.get("rainfall_multiplier"),
        boundary_types=case.get("boundary_types", (0, 0, 0, 0)), rainfall_times_s=case["rain_t"],
        rainfall_rates_m_s=case["rain_r"], dx=case["dx"], dry_tolerance=case.get("dry", 1e-4),
        cfl=case.get("cfl", 0.45), max_dt=case.get("max_dt", 10.0), dtype=dtype)


def _state(case, dev):
    from hybrid_flood_amd import numerics as nm

    torch = _torch()
    return nm.SWEState(*(torch.as_tensor(case[k], device=dev).contiguous() for k in ("h", "hu", "hv")))


def _golden_case(z):
    return dict(bed=z["bed"], manning_n=z["manning_n"], domain_mask=z["domain_mask"],
                rain_t=z["rain_t"], rain_r=z["rain_r"], dx=float(z["dx"]), dry=float(z["dry"]),
                cfl=float(z["cfl"]), max_dt=float(z["max_dt"]), h=z["h"], hu=z["hu"], hv=z["hv"])


GOLDEN_STEPS = ["v1_step_parity_7x9.npz", "v1_step_wetdry_16x12.npz", "v1_step_walls_20x24.npz"]


@pytest.mark.parametrize("name", GOLDEN_STEPS)
@pytest.mark.parametrize("variant", ["reference", "fast"])
def test_step_against_v1_golden(cuda, name, variant):
    from hybrid_flood_amd import numerics as nm

    z = load_golden(name)
    case = _golden_case(z)
    params = _params(case, cuda)
    out = nm.finite_volume_step(_state(case, cuda), params, float(z["dt"]), float(z["t"]),
                                variant=variant)
    got = [x.cpu().numpy() for x in out]
    dry = np.float32(case["dry"])
    wet_gold = z["domain_mask"] & (z["out_h"] > dry)
    wet_got = z["domain_mask"] & (got[0] > dry)
    assert np.array_equal(wet_got, wet_gold)
    if variant == "reference":
        assert np.array_equal(got[0], z["out_h"])
    for g, key in zip(got, ("out_h", "out_hu", "out_hv")):
        np.testing.assert_allclose(g, z[key], rtol=RTOL, atol=ATOL)


@pytest.mark.parametrize("name", GOLDEN_STEPS)
def test_cfl_dt_against_v1_golden(cuda, name):
    from hybrid_flood_amd import numerics as nm

    z = load_golden(name)
    case = _golden_case(z)
    params = _params(case, cuda)
    st = _state(case, cuda)
    assert nm.compute_cfl_dt(st, params, variant="reference") == float(z["cfl_dt"])
    assert nm.compute_cfl_dt(st, params, variant="fast") == pytest.approx(float(z["cfl_dt"]), rel=1e-6)


SHAPES = [(1, 1), (1, 9), (9, 1), (3, 200), (130, 127), (64, 124), (257, 65), (17, 63), (40, 186)]
SIDES = [(0, 0, 0, 0), (1, 1, 1, 1), (0, 1, 0, 0), (1, 0, 1, 0)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("sides", SIDES)
def test_step_reference_variant_bit_exact_depth(cuda, shape, sides):
    from hybrid_flood_amd import numerics as nm
    from oracle import swe_oracle as o

    seed = hash((shape, sides)) % 10_000
    case = random_case(seed, shape, open_sides=sides, speed=2.0)
    params = _params(case, cuda)
    op = _oparams(case)
    st_o = o.OracleState(case["h"], case["hu"], case["hv"])
    dt = o.compute_cfl_dt(st_o, op)
    ref = o.finite_volume_step(st_o, op, dt, 47.5)
    out = nm.finite_volume_step(_state(case, cuda), params, dt, 47.5, variant="reference")
    got = [x.cpu().numpy() for x in out]
    assert np.array_equal(got[0], ref.h), f"depth differs at {np.argwhere(got[0] != ref.h)[:5]}"
    for g, r in zip(got[1:], ref[1:]):
        assert int(_ulp_diff(g, r).max()) <= 4
    assert nm.compute_cfl_dt(_state(case, cuda), params, variant="reference") == dt


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("variant", ["reference", "fast"])
def test_step_against_fp64_oracle(cuda, shape, variant):
    from hybrid_flood_amd import numerics as nm
    from oracle import swe_oracle as o

    seed = 1000 + hash(shape) % 10_000
    case = random_case(seed, shape, open_sides=(0, 1, 1, 0), speed=3.0)
    params = _params(case, cuda)
    op64 = _oparams(case, np.float64)
    op32 = _oparams(case)
    st_o = o.OracleState(case["h"], case["hu"], case["hv"])
    dt = o.compute_cfl_dt(st_o, op32)
    ref64 = o.finite_volume_step(st_o, op64, dt, 120.0)
    ref32 = o.finite_volume_step(st_o, op32, dt, 120.0)
    out = nm.finite_volume_step(_state(case, cuda), params, dt, 120.0, variant=variant)
    got = [x.cpu().numpy() for x in out]
    for g, r in zip(got, ref64):
        np.testing.assert_allclose(g.astype(np.float64), r, rtol=RTOL, atol=ATOL)
    dry = np.float32(case["dry"])
    assert np.array_equal(case["domain_mask"] & (got[0] > dry), case["domain_mask"] & (ref32.h > dry))


def test_cfl_field_matches_oracle(cuda):
    from hybrid_flood_amd import numerics as nm
    from oracle import swe_oracle as o

    case = random_case(5, (70, 90), speed=2.0)
    params = _params(case, cuda)
    inv = nm.cfl_inverse_field(_state(case, cuda), params, variant="reference").cpu().numpy()
    ref = o.cfl_inverse_dt(o.OracleState(case["h"], case["hu"], case["hv"]), _oparams(case))
    assert np.array_equal(inv, ref)


def test_run_matches_v1_golden_loop(cuda):
    """Adaptive loop of the V1 golden run; friction is on, so fields match at the
    V1/V2 bar while step count, final time and outputs_written match exactly."""
    from hybrid_flood_amd import numerics as nm

    z = load_golden("v1_run_dambreak_12x14.npz")
    case = _golden_case(z)
    params = _params(case, cuda)
    res = nm.integrate_to_outputs(_state(case, cuda), params, z["times"], max_steps=1000,
                                  variant="reference")
    assert res.outputs_written == int(z["outputs_written"])
    assert res.steps_used == int(z["steps_used"])
    assert res.final_time_s == pytest.approx(float(z["final_time_s"]), abs=1e-12)
    for g, key in zip(res.states, ("out_h", "out_hu", "out_hv")):
        np.testing.assert_allclose(g.cpu().numpy(), z[key], rtol=RTOL, atol=ATOL)


@pytest.mark.parametrize("shape", [(12, 14), (50, 130), (129, 64)])
def test_frictionless_loop_bit_exact(cuda, shape):
    """With n = 0 the friction divisor is exactly 1, so the reference-order loop
    must reproduce the FP32 oracle loop bit for bit, step count and clock too."""
    from hybrid_flood_amd import numerics as nm
    from oracle import swe_oracle as o

    case = random_case(77, shape, open_sides=(0, 1, 0, 1), speed=1.0)
    case["manning_n"] = np.zeros(shape, np.float32)
    params = _params(case, cuda)
    times = np.array([0.0, 0.5, 1.0, 1.7], np.float32)
    res = nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=500,
                                  variant="reference")
    ref = o.integrate_to_outputs(o.OracleState(case["h"], case["hu"], case["hv"]), _oparams(case),
                                 times.astype(np.float64), max_steps=500)
    assert res.steps_used == ref.steps_used and res.outputs_written == ref.outputs_written == 4
    assert res.final_time_s == ref.final_time_s
    for g, r in zip(res.states, ref.states):
        assert np.array_equal(g.cpu().numpy(), r)
    d = res.diagnostics
    assert d["volume_initial"] == pytest.approx(ref.volume_initial, rel=1e-12)
    assert d["volume_final"] == pytest.approx(ref.volume_final, rel=1e-12)
    assert d["boundary_volume"] == pytest.approx(ref.boundary_volume, rel=1e-9, abs=1e-9)
    assert abs(d["mass_balance_rel"]) < 1e-6


def test_max_steps_partial_outputs(cuda):
    from hybrid_flood_amd import numerics as nm
    from oracle import swe_oracle as o

    case = random_case(3, (30, 40), speed=1.0)
    params = _params(case, cuda)
    times = np.array([0.0, 1.0, 2.0, 3.0], np.float32)
    res = nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=5, variant="reference")
    ref = o.integrate_to_outputs(o.OracleState(case["h"], case["hu"], case["hv"]), _oparams(case),
                                 times.astype(np.float64), max_steps=5)
    assert res.steps_used == ref.steps_used == 5
    assert res.outputs_written == ref.outputs_written
    assert res.final_time_s == pytest.approx(ref.final_time_s, rel=1e-12)
    # Unreached slots hold the last state, as the reference's scan stacks it.
    for g in res.states:
        g = g.cpu().numpy()
        for k in range(res.outputs_written, 4):
            assert np.array_equal(g[k], g[-1])


def test_run_steps_matches_loop(cuda):
    from hybrid_flood_amd import numerics as nm
    from oracle import swe_oracle as o

    case = random_case(9, (45, 77), open_sides=(0, 1, 0, 0), speed=1.0)
    case["manning_n"] = np.zeros((45, 77), np.float32)
    params = _params(case, cuda)
    st = _state(case, cuda)
    info = nm.run_steps(st, params, 13, t0=3.0, variant="reference", diagnostics=True)
    ref, t, _ = o.run_steps(o.OracleState(case["h"], case["hu"], case["hv"]), _oparams(case), 13, t0=3.0)
    assert info["steps_used"] == 13
    assert info["final_time_s"] == t
    for g, r in zip(st, ref):
        assert np.array_equal(g.cpu().numpy(), r)
