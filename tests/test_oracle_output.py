"""Closed forms of the output-schema oracle (jax_solver/shallow_water_2d.py:309-350)."""

from __future__ import annotations

import numpy as np


def test_primitive_closed_forms():
    from oracle import output_oracle as oo

    h = np.array([[[2.0, 1e-5, 0.5], [1.0, 4.0, 0.0]]], np.float32)
    hu = np.array([[[6.0, 1.0, 0.0], [-3.0, 0.0, 0.0]]], np.float32)
    hv = np.array([[[8.0, 1.0, 1.0], [4.0, 4.0, 0.0]]], np.float32)
    mask = np.array([[True, True, True], [True, False, True]])
    bed = np.arange(6, dtype=np.float32).reshape(2, 3)
    out = oo.conserved_to_primitive(h, hu, hv, bed, mask, 1e-4)
    np.testing.assert_array_equal(out["x_velocity"][0, 0], [3.0, 0.0, 0.0])  # dry cell -> 0
    np.testing.assert_array_equal(out["y_velocity"][0, 0], [4.0, 0.0, 2.0])
    np.testing.assert_array_equal(out["velocity"][0, 0], [5.0, 0.0, 2.0])      # 3-4-5
    assert out["velocity"][0, 1, 0] == 5.0                                     # (-3, 4)
    for name in ("depth", "x_velocity", "y_velocity", "velocity"):
        assert np.isnan(out[name][0, 1, 1])
        assert np.isfinite(out[name][0][mask]).all()
    assert np.isnan(out["elevation"][1, 1]) and out["elevation"][0, 2] == 2.0
    np.testing.assert_array_equal(out["depth"][0][mask], h[0][mask])
