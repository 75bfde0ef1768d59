"""Accuracy record of a device run against the oracle.

TEST INFRASTRUCTURE ONLY: the checker of ``tests/`` and of the
``cpu_baseline`` leg of ``bench.py``; the product path never imports it.

A device run of the adaptive loop records its own float64 dt schedule.  The
oracle is replayed with that schedule (``integrate_to_outputs(...,
dt_schedule=...)``), once in float32 and once in float64, so the comparison
isolates the per-step arithmetic: the oracle's own CFL dt would differ in the
last bits and shift every later step.  Reported per field (h, hu, hv), over
active cells and every output slot:

* max |device - FP64 oracle| and the RMSE against it (the reference's
  accuracy gate is an RMSE, ``config/comparison/v1_v2_t4.yaml:41-46``);
* the wet-mask mismatch count against the FP32 oracle, wet meaning
  ``mask and h > dry_tolerance`` (``jax_solver/numerics.py:365-369``).
"""

from __future__ import annotations

import numpy as np

from . import swe_oracle as o


def field_errors(got: np.ndarray, want: np.ndarray, mask: np.ndarray) -> dict:
    """max-abs and RMSE of got - want over the active cells of every slot."""
    m = np.broadcast_to(mask, got.shape)
    d = (got.astype(np.float64) - want.astype(np.float64))[m]
    if d.size == 0:
        return {"max_abs": 0.0, "rmse": 0.0}
    return {"max_abs": float(np.abs(d).max()), "rmse": float(np.sqrt(np.mean(d * d)))}


def wet_mask_mismatch(got_h: np.ndarray, want_h: np.ndarray, mask: np.ndarray, dry: float) -> dict:
    """Cells (over every slot) whose wet flag differs, and the wet count."""
    dry32 = np.float32(dry)
    wet_got = mask & (got_h > dry32)
    wet_want = mask & (want_h > dry32)
    return {"mismatch": int(np.count_nonzero(wet_got != wet_want)),
            "wet_cells": int(np.count_nonzero(wet_want))}


def replay(params64: o.OracleParams, params32: o.OracleParams, h0: np.ndarray, times,
           dts: np.ndarray):
    """The oracle in float64 and float32 along the device's dt schedule."""
    z = np.zeros_like(h0)
    init = o.OracleState(h0, z, z)
    r64 = o.integrate_to_outputs(init, params64, times, max_steps=len(dts), dt_schedule=dts)
    r32 = o.integrate_to_outputs(init, params32, times, max_steps=len(dts), dt_schedule=dts)
    return r64, r32


def record(device_states, params32: o.OracleParams, h0: np.ndarray, times, dts: np.ndarray,
           replays=None) -> dict:
    """Accuracy record of ``device_states`` (h, hu, hv stacks of shape
    (T, ny, nx)) run from ``h0`` (still water start) with schedule ``dts``."""
    params64 = o.OracleParams(**{**params32.__dict__, "dtype": np.float64})
    r64, r32 = replays if replays is not None else replay(params64, params32, h0, times, dts)
    mask = np.asarray(params32.domain_mask, bool)
    out = {"steps": int(len(dts)), "outputs": int(np.asarray(times).shape[0]),
           "vs_fp64": {}, "vs_fp32": {}}
    for k, f in enumerate(("h", "hu", "hv")):
        got = np.asarray(device_states[k])
        out["vs_fp64"][f] = field_errors(got, getattr(r64.states, f), mask)
        out["vs_fp32"][f] = field_errors(got, getattr(r32.states, f), mask)
    out["wet_mask_vs_fp32"] = wet_mask_mismatch(np.asarray(device_states[0]), r32.states.h, mask,
                                                params32.dry_tolerance)
    out["oracle_fp64_mass_balance_rel"] = float(r64.mass_balance_rel)
    out["oracle_fp32_mass_balance_rel"] = float(r32.mass_balance_rel)
    return out
