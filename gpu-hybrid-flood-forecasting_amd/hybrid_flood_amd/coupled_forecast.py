"""Hybrid coupled forecast: physics, features and learned correction
(hybrid/coupled_forecast.py).

``advance_physics_to_time`` is the fixed-length adaptive advance the coupled
loop calls once per correction interval (coupled_forecast.py:123-161): at
most ``max_substeps`` CFL steps toward one target time, and an all-NaN state
when the target is not reached.  It runs as one device-resident loop
(``swe_advance_to_time``): the step kernel and its CFL/finish kernel are
replayed from captured HIP graphs, and the host only polls the stop flag.
The clock is float64 (the reference's scan carries it in float32).

``network_features`` and ``apply_residual_correction`` are one HIP kernel
and three small kernels respectively (swe_coupled.hip); the CNN is any
PyTorch module and reads the feature buffer in place.  The correction's two
volume sums are float64 (the reference sums in float32).
"""

from __future__ import annotations

import ctypes
from typing import Any, Callable, NamedTuple

import numpy as np
import torch

from . import _lib
from .numerics import SWEParams, SWEState, _check_state, context_for


def advance_physics_to_time(state: SWEState, params: SWEParams, start_time_s, target_time_s, *,
                            max_substeps: int, variant: str | None = None,
                            out: SWEState | None = None) -> tuple[SWEState, int]:
    """Advance adaptively from ``start_time_s`` to ``target_time_s``.

    Returns ``(final_state, steps)`` like the reference (coupled_forecast.py:
    123-161); ``final_state`` is NaN everywhere if ``max_substeps`` steps did
    not reach the target within 1e-6 s.  ``out`` may be ``state`` itself.
    """
    if int(max_substeps) < 0:
        raise ValueError("max_substeps must be non-negative.")
    ctx = context_for(params, variant)
    _check_state(state, ctx.shape)
    if out is None:
        out = SWEState(torch.empty_like(state.h), torch.empty_like(state.hu), torch.empty_like(state.hv))
    else:
        _check_state(out, ctx.shape)
    steps = ctypes.c_int32(0)
    t_final = ctypes.c_double(0.0)
    _lib.check(
        _lib.load().swe_advance_to_time(
            ctx.handle, _lib.state_struct(*state), float(start_time_s), float(target_time_s),
            int(max_substeps), _lib.state_struct(*out), ctypes.byref(steps), ctypes.byref(t_final),
            _lib.stream_handle(state.h.device)),
        "swe_advance_to_time")
    return out, int(steps.value)


# ---------------------------------------------------------------------------
# Coupled loop: features -> physics advance -> CNN residual -> correction
# ---------------------------------------------------------------------------
M_S_TO_MM_HR = 3.6e6


class HybridInputs(NamedTuple):
    """Static normalised CNN inputs and training statistics (coupled_forecast.py:30-38).

    ``terrain_standardized`` / ``roughness_standardized``: device float32
    ``(ny, nx)``; ``state_mean``, ``state_std``, ``residual_std``: three
    per-channel values; ``rainfall_scale_mm_hr``: scalar.
    """

    terrain_standardized: torch.Tensor
    roughness_standardized: torch.Tensor
    state_mean: Any
    state_std: Any
    rainfall_scale_mm_hr: Any
    residual_std: Any


class CoupledRolloutResult(NamedTuple):
    """Corrected conserved states at the output times (coupled_forecast.py:41-46)."""

    states: SWEState
    output_times_s: torch.Tensor
    physics_steps_used: int


def _f32x3(v) -> ctypes.Array:
    a = np.asarray(torch.as_tensor(v).detach().cpu().numpy() if isinstance(v, torch.Tensor) else v,
                   dtype=np.float32).reshape(-1)
    if a.shape[0] != 3:
        raise ValueError("expected three per-channel values (h, u, v)")
    return (ctypes.c_float * 3)(*[float(x) for x in a])


def _scalar(v) -> float:
    return float(v.item() if isinstance(v, torch.Tensor) else v)


def network_features(state: SWEState, params: SWEParams, hybrid_inputs: HybridInputs, time_s, *,
                     layout: str = "nhwc", out: torch.Tensor | None = None) -> torch.Tensor:
    """The six normalised channels (coupled_forecast.py:53-79) in one kernel.

    ``layout="nhwc"`` returns ``(1, ny, nx, 6)`` as the reference does;
    ``"nchw"`` returns ``(1, 6, ny, nx)`` for a PyTorch convolution.
    """
    ctx = context_for(params)
    _check_state(state, ctx.shape)
    ny, nx = ctx.shape
    dev = state.h.device
    for f in (hybrid_inputs.terrain_standardized, hybrid_inputs.roughness_standardized):
        if tuple(f.shape) != (ny, nx) or f.dtype != torch.float32 or not f.is_contiguous():
            raise ValueError("static channels must be contiguous float32 (ny, nx) tensors")
    if layout == "nhwc":
        shape, cs, ps = (1, ny, nx, 6), 1, 6
    elif layout == "nchw":
        shape, cs, ps = (1, 6, ny, nx), ny * nx, 1
    else:
        raise ValueError("layout must be 'nhwc' or 'nchw'")
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=dev)
    elif tuple(out.shape) != shape or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous {shape} tensor")
    _lib.check(
        _lib.load().swe_network_features(
            ctx.handle, _lib.state_struct(*state),
            hybrid_inputs.terrain_standardized.data_ptr(),
            hybrid_inputs.roughness_standardized.data_ptr(), _f32x3(hybrid_inputs.state_mean),
            _f32x3(hybrid_inputs.state_std), _scalar(hybrid_inputs.rainfall_scale_mm_hr),
            float(time_s), out.data_ptr(), cs, ps, _lib.stream_handle(dev)),
        "swe_network_features")
    return out


def _residual_strides(residual: torch.Tensor, ny: int, nx: int) -> tuple[int, int]:
    """(channel stride, cell stride) of a (ny, nx, 3) view, e.g. NHWC or a permuted NCHW."""
    if residual.dim() == 4 and residual.shape[0] == 1:
        residual = residual[0]
    if tuple(residual.shape) != (ny, nx, 3) or residual.dtype != torch.float32:
        raise ValueError("residual must be a float32 (ny, nx, 3) view")
    s0, s1, s2 = residual.stride()
    if s0 != nx * s1:
        raise ValueError("residual rows must be evenly strided")
    return s2, s1


def apply_residual_correction(predicted: SWEState, residual: torch.Tensor, params: SWEParams, *,
                              correction_relaxation, clip_limits=None,
                              out: SWEState | None = None, volumes: bool = False):
    """Volume-preserving residual correction (coupled_forecast.py:82-120).

    ``residual`` is a float32 ``(ny, nx, 3)`` view (NHWC, or an NCHW network
    output permuted to (H, W, C)); ``clip_limits`` clips each channel first
    (coupled_forecast.py:199-201).  ``out`` may be ``predicted``.  With
    ``volumes=True`` also returns (physics volume, provisional volume, scale,
    rescaled) in cell-depth units.
    """
    ctx = context_for(params)
    _check_state(predicted, ctx.shape)
    ny, nx = ctx.shape
    cs, ps = _residual_strides(residual, ny, nx)
    if out is None:
        out = SWEState(torch.empty_like(predicted.h), torch.empty_like(predicted.hu),
                       torch.empty_like(predicted.hv))
    else:
        _check_state(out, ctx.shape)
    vols = (ctypes.c_double * 4)() if volumes else None
    lim = _f32x3(clip_limits) if clip_limits is not None else None
    _lib.check(
        _lib.load().swe_apply_residual_correction(
            ctx.handle, _lib.state_struct(*predicted), residual.data_ptr(), cs, ps,
            float(_scalar(correction_relaxation)), lim, _lib.state_struct(*out), vols,
            _lib.stream_handle(predicted.h.device)),
        "swe_apply_residual_correction")
    if volumes:
        return out, tuple(vols)
    return out


def run_coupled_forecast(initial_state: SWEState, params: SWEParams,
                         model: Callable[[torch.Tensor], torch.Tensor], hybrid_inputs: HybridInputs,
                         output_times_s, *, max_substeps_per_interval: int,
                         correction_relaxation: float = 0.05, correction_clip_sigma: float = 5.0,
                         variant: str | None = None) -> CoupledRolloutResult:
    """Autoregressive physics + learned-correction forecast (coupled_forecast.py:164-231).

    Per output interval: features of the current state at the current time,
    physics advance to the next output time, ``model`` on the features,
    clipped volume-preserving correction of the advanced state.  ``model``
    takes the features as a ``(1, 6, ny, nx)`` NCHW tensor (a channels-last
    view of the NHWC buffer, no copy) and returns ``(1, 3, ny, nx)``; the
    reference's flax ``apply_fn`` does the same in NHWC.
    """
    if int(max_substeps_per_interval) < 1:
        raise ValueError("At least one physics substep per correction interval is required.")
    times = np.asarray(output_times_s.detach().cpu().numpy() if isinstance(output_times_s, torch.Tensor)
                       else output_times_s, dtype=np.float32)
    if times.ndim != 1 or times.shape[0] < 2:
        raise ValueError("At least two output times are required.")
    ctx = context_for(params, variant)
    _check_state(initial_state, ctx.shape)
    ny, nx = ctx.shape
    dev = initial_state.h.device
    T = times.shape[0]
    cube = SWEState(*(torch.empty((T, ny, nx), dtype=torch.float32, device=dev) for _ in range(3)))
    for c, s in zip(cube, initial_state):
        c[0].copy_(s)
    rstd = np.asarray(torch.as_tensor(hybrid_inputs.residual_std).detach().cpu().numpy(), np.float32)
    limits = (np.float32(correction_clip_sigma) * rstd.reshape(-1)).astype(np.float32)
    feats = torch.empty((1, ny, nx, 6), dtype=torch.float32, device=dev)
    advanced = SWEState(*(torch.empty((ny, nx), dtype=torch.float32, device=dev) for _ in range(3)))
    total = 0
    for k in range(1, T):
        state = SWEState(cube.h[k - 1], cube.hu[k - 1], cube.hv[k - 1])
        current, target = float(times[k - 1]), float(times[k])
        network_features(state, params, hybrid_inputs, current, out=feats)
        _, steps = advance_physics_to_time(state, params, current, target,
                                           max_substeps=max_substeps_per_interval, variant=variant,
                                           out=advanced)
        total += steps
        pred = model(feats.permute(0, 3, 1, 2))
        if tuple(pred.shape) != (1, 3, ny, nx):
            raise ValueError(f"model must return (1, 3, {ny}, {nx}), got {tuple(pred.shape)}")
        residual = pred[0].permute(1, 2, 0)
        if residual.dtype != torch.float32:
            residual = residual.float()
        apply_residual_correction(advanced, residual, params,
                                  correction_relaxation=correction_relaxation, clip_limits=limits,
                                  out=SWEState(cube.h[k], cube.hu[k], cube.hv[k]))
    return CoupledRolloutResult(cube, torch.as_tensor(times), total)
