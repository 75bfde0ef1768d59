"""Drop-in for ``hybrid_flood.jax_solver.numerics`` backed by libswe_hip.

Same names, argument meaning and pytrees as the reference
(jax_solver/numerics.py:21-434); arrays are torch tensors on the GPU and every
step runs in the fused HIP kernel.  Scalars follow the reference's NumPy twin:
``dt``, ``time_s``, the rain rate and the CFL quotient are float64 and are
rounded to float32 where they meet the float32 fields
(numpy_solver/shallow_water_2d.py:254-327); the clock of
``integrate_to_outputs`` is float64.

``variant`` selects the kernel: ``"reference"`` (reference evaluation order,
correctly rounded quotients; bit-identical depth with the FP32 oracle) or
``"fast"`` (default; reciprocal-based, within the stated tolerance).
"""

from __future__ import annotations

import ctypes
from collections import OrderedDict
from typing import Any, NamedTuple

import numpy as np
import torch

from . import _lib

DEFAULT_VARIANT = "fast"


class SWEState(NamedTuple):
    """Cell-average conserved variables (numerics.py:21-26)."""

    h: torch.Tensor
    hu: torch.Tensor
    hv: torch.Tensor


class SWEParams(NamedTuple):
    """Static fields and scalar controls (numerics.py:29-45)."""

    bed: torch.Tensor
    manning_n: torch.Tensor
    domain_mask: torch.Tensor
    rainfall_multiplier: torch.Tensor
    boundary_types: Any
    rainfall_times_s: Any
    rainfall_rates_m_s: Any
    default_rainfall_m_s: Any
    dx: Any
    dy: Any
    gravity: Any
    cfl: Any
    dry_tolerance: Any
    max_dt: Any


class ScanResult(NamedTuple):
    """Fixed-output result (numerics.py:48-55) plus the FP64 run diagnostics."""

    states: SWEState
    output_times_s: torch.Tensor
    final_time_s: float
    outputs_written: int
    steps_used: int
    diagnostics: dict | None = None


def _scalar(x) -> float:
    if isinstance(x, torch.Tensor):
        return float(x.detach().cpu().double())
    return float(np.asarray(x, dtype=np.float64))


def _host_f32(x) -> np.ndarray:
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    return np.asarray(x, dtype=np.float32).reshape(-1)


def _device_field(x, dtype, device) -> torch.Tensor:
    t = torch.as_tensor(x)
    if t.dtype != dtype or t.device != device or not t.is_contiguous():
        t = t.to(device=device, dtype=dtype).contiguous()
    return t


def make_params(bed, manning_n, domain_mask, rainfall_multiplier=None, *, boundary_types=(0, 0, 0, 0),
                rainfall_times_s=(0.0,), rainfall_rates_m_s=(0.0,), default_rainfall_m_s=0.0,
                dx=1.0, dy=None, gravity=9.81, cfl=0.45, dry_tolerance=1.0e-4, max_dt=10.0,
                device=None) -> SWEParams:
    """Build ``SWEParams`` with device-resident float32 fields and a uint8 mask."""
    device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    bed_t = _device_field(np.asarray(bed, np.float32) if not isinstance(bed, torch.Tensor) else bed,
                          torch.float32, device)
    mask_t = _device_field(domain_mask, torch.bool, device).view(torch.uint8)
    if rainfall_multiplier is None:  # shallow_water_2d.py:229-233
        mult_t = mask_t.to(torch.float32)
    else:
        mult_t = _device_field(rainfall_multiplier, torch.float32, device)
    n_t = _device_field(manning_n, torch.float32, device)
    for name, t in (("manning_n", n_t), ("domain_mask", mask_t), ("rainfall_multiplier", mult_t)):
        if tuple(t.shape) != tuple(bed_t.shape):
            raise ValueError(f"{name} must match the bed shape {tuple(bed_t.shape)}.")
    return SWEParams(
        bed=bed_t, manning_n=n_t, domain_mask=mask_t, rainfall_multiplier=mult_t,
        boundary_types=tuple(int(b) for b in np.asarray(boundary_types).reshape(-1)),
        rainfall_times_s=_host_f32(rainfall_times_s), rainfall_rates_m_s=_host_f32(rainfall_rates_m_s),
        default_rainfall_m_s=float(default_rainfall_m_s), dx=float(dx),
        dy=float(dx if dy is None else dy), gravity=float(gravity), cfl=float(cfl),
        dry_tolerance=float(dry_tolerance), max_dt=float(max_dt),
    )


# Contexts are cached per (params identity, variant): creating one uploads the
# rainfall series and sizes the reduction buffers.
_CONTEXTS: "OrderedDict[tuple, _lib.Context]" = OrderedDict()
_MAX_CONTEXTS = 8


def _params_key(params: SWEParams, variant: str) -> tuple:
    return (
        variant,
        params.bed.data_ptr(), tuple(params.bed.shape), params.manning_n.data_ptr(),
        params.domain_mask.data_ptr(), params.rainfall_multiplier.data_ptr(),
        tuple(int(b) for b in np.asarray(torch.as_tensor(params.boundary_types).cpu()).reshape(-1)),
        _host_f32(params.rainfall_times_s).tobytes(), _host_f32(params.rainfall_rates_m_s).tobytes(),
        _scalar(params.default_rainfall_m_s), _scalar(params.dx), _scalar(params.dy),
        _scalar(params.gravity), _scalar(params.cfl), _scalar(params.dry_tolerance),
        _scalar(params.max_dt),
    )


def context_for(params: SWEParams, variant: str | None = None) -> _lib.Context:
    variant = variant or DEFAULT_VARIANT
    for name in ("bed", "manning_n", "domain_mask", "rainfall_multiplier"):
        t = getattr(params, name)
        if not isinstance(t, torch.Tensor) or not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"SWEParams.{name} must be a contiguous GPU tensor (see make_params).")
    if params.domain_mask.element_size() != 1:
        raise ValueError("SWEParams.domain_mask must be a bool or uint8 tensor.")
    key = _params_key(params, variant)
    ctx = _CONTEXTS.get(key)
    if ctx is not None:
        _CONTEXTS.move_to_end(key)
        return ctx
    ctx = _lib.Context(
        bed=params.bed, manning_n=params.manning_n, domain_mask=params.domain_mask,
        rainfall_multiplier=params.rainfall_multiplier,
        boundary_types=key[6], rainfall_times_s=_host_f32(params.rainfall_times_s),
        rainfall_rates_m_s=_host_f32(params.rainfall_rates_m_s),
        default_rainfall_m_s=key[9], dx=key[10], dy=key[11], gravity=key[12], cfl=key[13],
        dry_tolerance=key[14], max_dt=key[15], variant=variant,
    )
    _CONTEXTS[key] = ctx
    while len(_CONTEXTS) > _MAX_CONTEXTS:
        _CONTEXTS.popitem(last=False)
    return ctx


def _check_state(state: SWEState, shape) -> None:
    for name, t in zip(("h", "hu", "hv"), state):
        if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32:
            raise ValueError(f"state.{name} must be a float32 GPU tensor.")
        if tuple(t.shape) != tuple(shape) or not t.is_contiguous():
            raise ValueError(f"state.{name} must be contiguous with shape {tuple(shape)}.")


def rainfall_rate(params: SWEParams, time_s) -> float:
    """Uniform rain (m/s) at ``time_s`` (numerics.py:286-294), float64 interpolation."""
    times = _host_f32(params.rainfall_times_s)
    rates = _host_f32(params.rainfall_rates_m_s)
    out = ctypes.c_double()
    _lib.check(
        _lib.load().swe_interp_rate(
            times.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            rates.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(times),
            _scalar(params.default_rainfall_m_s), float(time_s), ctypes.byref(out)),
        "swe_interp_rate")
    return out.value


def compute_cfl_dt(state: SWEState, params: SWEParams, *, variant: str | None = None) -> float:
    """Unsplit CFL step of ``state`` (numerics.py:297-309); synchronises."""
    ctx = context_for(params, variant)
    _check_state(state, ctx.shape)
    out = ctypes.c_double()
    _lib.check(_lib.load().swe_cfl_dt(ctx.handle, _lib.state_struct(*state), ctypes.byref(out),
                                      _lib.stream_handle()), "swe_cfl_dt")
    return out.value


def cfl_inverse_field(state: SWEState, params: SWEParams, *, variant: str | None = None) -> torch.Tensor:
    """Per-cell ``(|u|+c)/dx + (|v|+c)/dy`` (numerics.py:299-303), 0 outside the mask."""
    ctx = context_for(params, variant)
    _check_state(state, ctx.shape)
    out = torch.empty_like(state.h)
    _lib.check(_lib.load().swe_cfl_field(ctx.handle, _lib.state_struct(*state), out.data_ptr(),
                                         _lib.stream_handle()), "swe_cfl_field")
    return out


def finite_volume_step(state: SWEState, params: SWEParams, dt, time_s, *,
                       variant: str | None = None, out: SWEState | None = None) -> SWEState:
    """One conservative, well-balanced HLL step (numerics.py:312-369)."""
    ctx = context_for(params, variant)
    _check_state(state, ctx.shape)
    if out is None:
        out = SWEState(torch.empty_like(state.h), torch.empty_like(state.hu), torch.empty_like(state.hv))
    else:
        _check_state(out, ctx.shape)
    _lib.check(
        _lib.load().swe_step(ctx.handle, _lib.state_struct(*state), _lib.state_struct(*out),
                             float(dt), float(time_s), _lib.stream_handle()),
        "swe_step")
    return out


def _output_times_f64(output_times_s) -> np.ndarray:
    # The reference holds output times in the float32 state dtype
    # (jax_solver/shallow_water_2d.py:286); the clock compares against their
    # float64 values as the NumPy twin does (numpy_solver/shallow_water_2d.py:390,403).
    if isinstance(output_times_s, torch.Tensor):
        output_times_s = output_times_s.detach().cpu().numpy()
    return np.asarray(np.asarray(output_times_s, dtype=np.float32), dtype=np.float64)


def integrate_to_outputs(initial_state: SWEState, params: SWEParams, output_times_s, *,
                         max_steps: int, variant: str | None = None,
                         rain_rounding: str | None = None) -> ScanResult:
    """Integrate adaptively to fixed output times (numerics.py:372-434).

    The whole loop runs on the device; the host launches chunks of steps and
    polls the stop flag.  Like the reference's scan it does not raise when
    ``max_steps`` ends the loop early: ``outputs_written`` tells.
    ``rain_rounding`` ("nearest", the reference's, or "stochastic"; None keeps
    the context's setting) selects how the rain source is rounded for this
    run (``swe_set_rain_rounding``).
    """
    times = _output_times_f64(output_times_s)
    if times.ndim != 1 or times.shape[0] < 1:
        raise ValueError("At least one one-dimensional output time is required.")
    ctx = context_for(params, variant)
    _check_state(initial_state, ctx.shape)
    if rain_rounding is not None and rain_rounding != ctx.rain_rounding:
        previous = ctx.rain_rounding
        ctx.set_rain_rounding(rain_rounding)
        try:
            return integrate_to_outputs(initial_state, params, output_times_s,
                                        max_steps=max_steps, variant=variant)
        finally:
            ctx.set_rain_rounding(previous)
    ny, nx = ctx.shape
    T = times.shape[0]
    dev = initial_state.h.device
    cube = [torch.empty((T, ny, nx), dtype=torch.float32, device=dev) for _ in range(3)]
    info = _lib.SweRunInfo()
    rc = _lib.load().swe_run_to_outputs(
        ctx.handle, _lib.state_struct(*initial_state),
        times.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), T, int(max_steps),
        cube[0].data_ptr(), cube[1].data_ptr(), cube[2].data_ptr(), _lib.SweState(None, None, None),
        ctypes.byref(info), _lib.stream_handle(dev))
    if rc not in (_lib.SWE_OK, _lib.SWE_EMAXSTEPS, _lib.SWE_ENONFINITE):
        _lib.check(rc, "swe_run_to_outputs")
    return ScanResult(
        states=SWEState(*cube),
        output_times_s=torch.as_tensor(times.astype(np.float32)),
        final_time_s=info.final_time_s,
        outputs_written=info.outputs_written,
        steps_used=info.steps_used,
        diagnostics=info.as_dict(),
    )


def run_steps(state: SWEState, params: SWEParams, n_steps: int, *, t0: float = 0.0,
              variant: str | None = None, diagnostics: bool = False):
    """Advance ``state`` in place by exactly ``n_steps`` adaptive (CFL) steps on the device."""
    ctx = context_for(params, variant)
    _check_state(state, ctx.shape)
    info = _lib.SweRunInfo() if diagnostics else None
    rc = _lib.load().swe_run_steps(ctx.handle, _lib.state_struct(*state), int(n_steps), float(t0),
                                   ctypes.byref(info) if info is not None else None,
                                   _lib.stream_handle(state.h.device))
    if rc not in (_lib.SWE_OK, _lib.SWE_ENONFINITE):
        _lib.check(rc, "swe_run_steps")
    return info.as_dict() if info is not None else None
