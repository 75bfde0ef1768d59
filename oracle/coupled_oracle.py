"""CPU oracle for the hybrid coupled forecast around the step.

TEST INFRASTRUCTURE ONLY: the checker for ``hybrid_flood_amd.coupled_forecast``.
Only ``tests/`` may import it; the product path never calls it.

Restates ``src/hybrid_flood/hybrid/coupled_forecast.py`` (and the staging of
``src/hybrid_flood/comparison/v2_staging.py``) in NumPy float32:

* ``network_features``          -- coupled_forecast.py:53-79;
* ``apply_residual_correction`` -- coupled_forecast.py:82-120, with the two
  volume sums in float64 (the reference sums in float32) and the rescaled
  depth rounded once from float64, as the HIP kernel does;
* ``run_coupled_forecast``      -- coupled_forecast.py:164-231, with the
  physics advance of ``swe_oracle.advance_physics_to_time`` (float64 clock)
  and any NumPy callable as the network (NHWC in, NHWC out, batch 1);
* ``stage_features`` / ``correct_depth`` -- the batched staging of
  comparison/v2_staging.py:17-52.

Pinned in ``tests/test_oracle_coupled.py`` by closed forms: zero residual
keeps the masked depth, the corrected volume equals the physics volume for
any residual, positivity, clipping equals pre-clipped input, uniform-state
features, the piecewise-linear rain channel, staged channels of known
inputs.
"""

from __future__ import annotations

import numpy as np

from . import swe_oracle as so

M_S_TO_MM_HR = 3.6e6
F32_EPS = float(np.finfo(np.float32).eps)


def _velocity(q, h, dry):
    """coupled_forecast.py:49-50."""
    h = np.asarray(h, np.float32)
    q = np.asarray(q, np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(h > dry, q / np.maximum(h, dry), np.float32(0)).astype(np.float32)


def network_features(state, params, terrain, roughness, state_mean, state_std,
                     rainfall_scale_mm_hr, time_s):
    """Six channels (1, ny, nx, 6), float32 (coupled_forecast.py:53-79).

    The rain rate is the float64 interpolation of the oracle (V1 semantics)
    rounded to float32 before the float32 products of the reference.
    """
    dry = np.float32(params.dry_tolerance)
    h = np.asarray(state.h, np.float32)
    u = _velocity(state.hu, h, dry)
    v = _velocity(state.hv, h, dry)
    mean = np.asarray(state_mean, np.float32)
    std = np.asarray(state_std, np.float32)
    dyn = np.stack((h, u, v), axis=-1)
    dyn = ((dyn - mean) / std).astype(np.float32)
    dyn = np.where(np.asarray(params.domain_mask)[..., None], dyn, np.float32(0))
    mm_hr = np.float32(so.rainfall_rate(params, float(time_s))) * np.float32(M_S_TO_MM_HR)
    rain = np.float32(mm_hr / np.float32(rainfall_scale_mm_hr))
    feats = np.concatenate((dyn, np.full(h.shape + (1,), rain, np.float32),
                            np.asarray(terrain, np.float32)[..., None],
                            np.asarray(roughness, np.float32)[..., None]), axis=-1)
    return feats[None].astype(np.float32)


def apply_residual_correction(predicted, residual, params, *, correction_relaxation,
                              clip_limits=None):
    """Volume-preserving correction (coupled_forecast.py:82-120), clip of :199-201 first.

    Returns ``(state, (physics volume, provisional volume, scale, rescaled))``.
    """
    dry = np.float32(params.dry_tolerance)
    mask = np.asarray(params.domain_mask, bool)
    r = np.asarray(residual, np.float32)
    if clip_limits is not None:
        lim = np.asarray(clip_limits, np.float32)
        r = np.clip(r, -lim, lim).astype(np.float32)
    relax = np.float32(correction_relaxation)
    h = np.asarray(predicted.h, np.float32)
    u = _velocity(predicted.hu, h, dry)
    v = _velocity(predicted.hv, h, dry)
    prov = np.where(mask, np.maximum(h + relax * r[..., 0], np.float32(0)), np.float32(0))
    prov = prov.astype(np.float32)
    target = float(np.sum(np.where(mask, h, 0).astype(np.float64)))
    prov_vol = float(np.sum(prov.astype(np.float64)))
    scale = target / prov_vol if prov_vol > F32_EPS else 1.0
    rescaled = target > F32_EPS
    ch = (prov.astype(np.float64) * scale).astype(np.float32) if rescaled else h.copy()
    cu = (u + relax * r[..., 1]).astype(np.float32)
    cv = (v + relax * r[..., 2]).astype(np.float32)
    wet = mask & (ch > dry)
    chu = np.where(wet, ch * cu, np.float32(0)).astype(np.float32)
    chv = np.where(wet, ch * cv, np.float32(0)).astype(np.float32)
    return so.OracleState(ch, chu, chv), (target, prov_vol, scale, float(rescaled))


def run_coupled_forecast(initial, params, model, terrain, roughness, state_mean, state_std,
                         rainfall_scale_mm_hr, residual_std, output_times_s, *,
                         max_substeps_per_interval, correction_relaxation=0.05,
                         correction_clip_sigma=5.0):
    """Autoregressive loop (coupled_forecast.py:164-231); ``model(features_nhwc) -> (1, ny, nx, 3)``."""
    if max_substeps_per_interval < 1:
        raise ValueError("At least one physics substep per correction interval is required.")
    times = np.asarray(output_times_s, np.float32)
    if times.shape[0] < 2:
        raise ValueError("At least two output times are required.")
    limits = (np.float32(correction_clip_sigma) * np.asarray(residual_std, np.float32))
    state = so.OracleState(*(np.asarray(x, np.float32) for x in initial))
    outs = [state]
    total = 0
    for k in range(1, times.shape[0]):
        current, target = float(times[k - 1]), float(times[k])
        feats = network_features(state, params, terrain, roughness, state_mean, state_std,
                                 rainfall_scale_mm_hr, current)
        advanced, steps, _ = so.advance_physics_to_time(state, params, current, target,
                                                        max_substeps=max_substeps_per_interval)
        total += steps
        residual = np.asarray(model(feats), np.float32)[0]
        state, _ = apply_residual_correction(advanced, residual, params,
                                             correction_relaxation=correction_relaxation,
                                             clip_limits=limits)
        outs.append(state)
    stack = so.OracleState(*(np.stack([getattr(s, f) for s in outs]) for f in ("h", "hu", "hv")))
    return stack, times, total


def stage_features(state_t, rainfall_t_mm_hr, terrain, roughness, domain_mask, state_mean,
                   state_std, rainfall_scale_mm_hr):
    """(count, ny, nx, 6) float32 features of primitive states (comparison/v2_staging.py:17-44)."""
    st = np.asarray(state_t, np.float32)
    mean = np.asarray(state_mean, np.float32)
    std = np.asarray(state_std, np.float32)
    dyn = ((st - mean) / std).astype(np.float32)
    dyn = np.where(np.asarray(domain_mask, bool)[None, ..., None], dyn, np.float32(0))
    count, ny, nx, _ = st.shape
    rain = (np.asarray(rainfall_t_mm_hr, np.float32) / np.float32(rainfall_scale_mm_hr)).astype(np.float32)
    return np.concatenate((
        dyn, np.broadcast_to(rain[:, None, None, None], (count, ny, nx, 1)),
        np.broadcast_to(np.asarray(terrain, np.float32)[None, ..., None], (count, ny, nx, 1)),
        np.broadcast_to(np.asarray(roughness, np.float32)[None, ..., None], (count, ny, nx, 1))),
        axis=-1).astype(np.float32)


def correct_depth(raw_depth, residual_nhwc):
    """max(depth + residual[..., 0], 0) per sample (comparison/v2_staging.py:47-52)."""
    d = np.asarray(raw_depth, np.float32)
    return np.maximum(d + np.asarray(residual_nhwc, np.float32)[..., 0], np.float32(0))
