"""The FP32 oracle against V1 golden vectors written from the reference's
NumPy twin (tests/golden/make_v1_golden.py).

With ``norm="hypot"`` the oracle uses V1's momentum norm and must equal V1 bit
for bit in every field; with the default ``sqrt`` norm (V2's,
numerics.py:355) depth must still be bit-identical and hu, hv stay within the
reference's V1/V2 bar (tests/test_v1_v2_architecture.py:104).
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden, oracle_params_from_golden

STEP_CASES = ["v1_step_parity_7x9.npz", "v1_step_wetdry_16x12.npz", "v1_step_walls_20x24.npz"]


@pytest.mark.parametrize("name", STEP_CASES)
def test_oracle_step_equals_v1_bitwise(name):
    from oracle import swe_oracle as o

    z = load_golden(name)
    p = oracle_params_from_golden(z)
    st = o.OracleState(z["h"], z["hu"], z["hv"])
    out = o.finite_volume_step(st, p, float(z["dt"]), float(z["t"]), norm="hypot")
    for got, key in zip(out, ("out_h", "out_hu", "out_hv")):
        assert got.dtype == np.float32
        assert np.array_equal(got, z[key]), key


@pytest.mark.parametrize("name", STEP_CASES)
def test_oracle_step_sqrt_norm_within_v1_v2_bar(name):
    from oracle import swe_oracle as o

    z = load_golden(name)
    p = oracle_params_from_golden(z)
    out = o.finite_volume_step(o.OracleState(z["h"], z["hu"], z["hv"]), p, float(z["dt"]),
                               float(z["t"]))
    assert np.array_equal(out.h, z["out_h"])
    np.testing.assert_allclose(out.hu, z["out_hu"], rtol=2e-5, atol=2e-7)
    np.testing.assert_allclose(out.hv, z["out_hv"], rtol=2e-5, atol=2e-7)


@pytest.mark.parametrize("name", STEP_CASES)
def test_oracle_cfl_dt_equals_v1(name):
    from oracle import swe_oracle as o

    z = load_golden(name)
    p = oracle_params_from_golden(z)
    assert o.compute_cfl_dt(o.OracleState(z["h"], z["hu"], z["hv"]), p) == float(z["cfl_dt"])


def test_oracle_loop_equals_v1_run():
    from oracle import swe_oracle as o

    z = load_golden("v1_run_dambreak_12x14.npz")
    p = oracle_params_from_golden(z)
    r = o.integrate_to_outputs(o.OracleState(z["h"], z["hu"], z["hv"]), p,
                               z["times"].astype(np.float64), max_steps=1000, norm="hypot")
    assert r.outputs_written == int(z["outputs_written"])
    assert r.steps_used == int(z["steps_used"])
    assert r.final_time_s == float(z["final_time_s"])
    for got, key in zip(r.states, ("out_h", "out_hu", "out_hv")):
        assert np.array_equal(got, z[key]), key
