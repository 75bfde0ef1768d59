"""CPU oracle for the explicit 2-D shallow-water finite-volume step.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP path is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product path never
calls it and never falls back to it.

It restates, in NumPy, the reference solver's arithmetic for one time step and
for the adaptive output loop, parameterised by the array dtype (float32 for the
FP32 parity oracle, float64 for the FP64 accuracy oracle):

* array arithmetic and evaluation order follow
  ``src/hybrid_flood/jax_solver/numerics.py:58-369`` (V2) and
  ``src/hybrid_flood/jax_solver/boundary_conditions.py:29-52``;
* scalar handling follows the NumPy twin
  ``src/hybrid_flood/numpy_solver/shallow_water_2d.py:254-327`` (V1): ``dt``,
  ``t``, the rainfall rate and the CFL quotient are float64 and are rounded to
  the array dtype where they meet an array (``dt * rate``, ``dt * g``);
* the momentum norm is ``sqrt(hu*hu + hv*hv)`` as in V2 (numerics.py:355);
  ``norm="hypot"`` selects V1's ``np.hypot`` (shallow_water_2d.py:312) so the
  oracle can be compared bit-for-bit with V1 golden vectors;
* the output loop follows ``integrate_to_outputs`` (numerics.py:372-434) with a
  float64 clock, as V1 keeps it (shallow_water_2d.py:399-412);
* the mass balance is new: FP64 ``V - V0 - R - B - C`` in the form of the
  reference's ANUGA bookkeeping (``anuga_solver/run_simulation.py:100-107``),
  with rain ``R``, open-boundary flux ``B`` and positivity-clip volume ``C``.

Pinning: every function here is checked in ``tests/test_oracle_*.py`` against
closed forms (HLL on equal / supersonic / mirrored states, ghost-cell known
answers, lake at rest, flat-basin conservation, rain and friction closed forms,
CFL of a uniform state, Ritter dam break, x/y transposition symmetry) and
against V1 golden vectors committed under ``tests/golden/``.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import NamedTuple

import numpy as np

REFLECTIVE = 0
OPEN = 1
OUTPUT_TIME_SLACK_S = 1.0e-7  # numerics.py:401 / shallow_water_2d.py:408


class OracleState(NamedTuple):
    h: np.ndarray
    hu: np.ndarray
    hv: np.ndarray


@dataclass
class OracleParams:
    """Static fields and controls (numerics.py:29-45)."""

    bed: np.ndarray
    manning_n: np.ndarray
    domain_mask: np.ndarray
    rainfall_multiplier: np.ndarray | None = None
    boundary_types: tuple[int, int, int, int] = (REFLECTIVE,) * 4
    rainfall_times_s: np.ndarray = field(default_factory=lambda: np.zeros(1, np.float32))
    rainfall_rates_m_s: np.ndarray = field(default_factory=lambda: np.zeros(1, np.float32))
    default_rainfall_m_s: float = 0.0
    dx: float = 1.0
    dy: float | None = None
    gravity: float = 9.81
    cfl: float = 0.45
    dry_tolerance: float = 1.0e-4
    max_dt: float = 10.0
    dtype: type = np.float32
    # Fields are stored in float32 first (shallow_water_2d.py:237-251), then
    # widened for the float64 oracle, so both oracles see identical data.
    # Finite-difference checks pass np.float64 here to keep tiny perturbations.
    field_dtype: type = np.float32

    def __post_init__(self) -> None:
        dt = self.dtype
        fd = self.field_dtype
        self.bed = np.asarray(np.asarray(self.bed, fd), dt)
        self.manning_n = np.asarray(np.asarray(self.manning_n, fd), dt)
        self.domain_mask = np.asarray(self.domain_mask, bool)
        if self.rainfall_multiplier is None:  # shallow_water_2d.py:229-233
            self.rainfall_multiplier = self.domain_mask.astype(np.float32)
        self.rainfall_multiplier = np.asarray(np.asarray(self.rainfall_multiplier, fd), dt)
        self.rainfall_times_s = np.asarray(self.rainfall_times_s, np.float32)
        self.rainfall_rates_m_s = np.asarray(self.rainfall_rates_m_s, np.float32)
        if self.dy is None:
            self.dy = self.dx
        shape = self.bed.shape
        for name in ("manning_n", "domain_mask", "rainfall_multiplier"):
            if getattr(self, name).shape != shape:
                raise ValueError(f"{name} must match the bed shape {shape}.")

    @property
    def shape(self) -> tuple[int, int]:
        return self.bed.shape


# ----------------------------------------------------------------------------
# Pieces of the step
# ----------------------------------------------------------------------------


def apply_ghost_cells(h, hu, hv, boundary_types):
    """One edge-padded ghost layer; reflective sides negate the normal momentum.

    boundary_conditions.py:29-52 (side order left, right, bottom, top).
    """
    hg = np.pad(h, 1, mode="edge")
    hug = np.pad(hu, 1, mode="edge")
    hvg = np.pad(hv, 1, mode="edge")
    left, right, bottom, top = (int(b) for b in boundary_types)
    hug[:, 0] = -hug[:, 1] if left == REFLECTIVE else hug[:, 1]
    hug[:, -1] = -hug[:, -2] if right == REFLECTIVE else hug[:, -2]
    hvg[0, :] = -hvg[1, :] if bottom == REFLECTIVE else hvg[1, :]
    hvg[-1, :] = -hvg[-2, :] if top == REFLECTIVE else hvg[-2, :]
    return hg, hug, hvg


def safe_velocity(q, h, dry):
    """``h > dry ? q / max(h, dry) : 0`` (numerics.py:58-59)."""
    out = np.zeros_like(q)
    np.divide(q, np.maximum(h, dry), out=out, where=h > dry)
    return out


def hll(hA, qnA, qtA, hB, qnB, qtB, g, half_g, dry, tangential_from_normal):
    """HLL flux in face-normal coordinates (numerics.py:62-159).

    Side A is left (x) or bottom (y).  Returns ``(F_h, F_normal, F_tangential)``.
    The x-face tangential flux is ``q_n * v_t'`` (numerics.py:85) and the
    y-face one is ``q_t * v_n'`` (numerics.py:135); ``tangential_from_normal``
    selects the latter.
    """
    unA = safe_velocity(qnA, hA, dry)
    unB = safe_velocity(qnB, hB, dry)
    cA = np.sqrt(g * np.maximum(hA, 0.0))
    cB = np.sqrt(g * np.maximum(hB, 0.0))
    sA = np.minimum(unA - cA, unB - cB)
    sB = np.maximum(unA + cA, unB + cB)
    if tangential_from_normal:
        ftA = qtA * unA
        ftB = qtB * unB
    else:
        ftA = qnA * safe_velocity(qtA, hA, dry)
        ftB = qnB * safe_velocity(qtB, hB, dry)
    fA = (qnA, qnA * unA + half_g * (hA * hA), ftA)
    fB = (qnB, qnB * unB + half_g * (hB * hB), ftB)
    uA = (hA, qnA, qtA)
    uB = (hB, qnB, qtB)
    spread = sB - sA
    den = np.where(spread > 1.0e-12, spread, 1.0).astype(hA.dtype)
    sAsB = sA * sB
    out = []
    for k in range(3):
        mid = ((sB * fA[k] - sA * fB[k]) + sAsB * (uB[k] - uA[k])) / den
        out.append(np.where(sA >= 0.0, fA[k], np.where(sB <= 0.0, fB[k], mid)))
    return out


def face_fluxes(hA, qnA, qtA, bA, aA, hB, qnB, qtB, bB, aB, g, half_g, dry, axis):
    """Interface fluxes with wall mirroring and hydrostatic reconstruction.

    numerics.py:162-224 (x) and 227-283 (y), written once in face-normal
    coordinates.  Returns ``(F_h, F_n, F_t, corrA, corrB)``: the face flux
    (zero on faces with no active side) and the pressure corrections that the
    A- and B-side cells add to the normal component.
    """
    wallA = aA & ~aB  # A active, B inactive: B becomes A's mirror image
    wallB = ~aA & aB
    hB = np.where(wallA, hA, hB)
    qnB = np.where(wallA, -qnA, qnB)
    qtB = np.where(wallA, qtA, qtB)
    bB = np.where(wallA, bA, bB)
    hA = np.where(wallB, hB, hA)
    qnA = np.where(wallB, -qnB, qnA)
    qtA = np.where(wallB, qtB, qtA)
    bA = np.where(wallB, bB, bA)

    bf = np.maximum(bA, bB)
    hrA = np.maximum((hA + bA) - bf, 0.0)
    hrB = np.maximum((hB + bB) - bf, 0.0)
    unA = safe_velocity(qnA, hA, dry)
    utA = safe_velocity(qtA, hA, dry)
    unB = safe_velocity(qnB, hB, dry)
    utB = safe_velocity(qtB, hB, dry)
    fh, fn, ft = hll(
        hrA, hrA * unA, hrA * utA, hrB, hrB * unB, hrB * utB, g, half_g, dry,
        tangential_from_normal=(axis == "y"),
    )
    relevant = aA | aB
    zero = np.zeros((), hA.dtype)
    fh = np.where(relevant, fh, zero)
    fn = np.where(relevant, fn, zero)
    ft = np.where(relevant, ft, zero)
    corrA = np.where(relevant, half_g * (hA * hA - hrA * hrA), zero)
    corrB = np.where(relevant, half_g * (hB * hB - hrB * hrB), zero)
    return fh, fn, ft, corrA, corrB


def rainfall_rate(params: OracleParams, time_s: float) -> float:
    """Uniform rain in m/s at ``time_s``: float64 ``np.interp`` with the default
    outside the series (shallow_water_2d.py:254-264 of the NumPy twin)."""
    return float(
        np.interp(
            time_s,
            params.rainfall_times_s,
            params.rainfall_rates_m_s,
            left=params.default_rainfall_m_s,
            right=params.default_rainfall_m_s,
        )
    )


def cfl_inverse_dt(state: OracleState, params: OracleParams) -> np.ndarray:
    """Per-cell ``(|u|+c)/dx + (|v|+c)/dy``, zero outside the mask (numerics.py:297-303)."""
    dt_ = params.dtype
    dry = dt_(params.dry_tolerance)
    u = safe_velocity(state.hu, state.h, dry)
    v = safe_velocity(state.hv, state.h, dry)
    wave = np.sqrt(dt_(params.gravity) * np.maximum(state.h, 0.0))
    inv = (np.abs(u) + wave) / dt_(params.dx) + (np.abs(v) + wave) / dt_(params.dy)
    return np.where(params.domain_mask, inv, 0.0).astype(dt_)


def compute_cfl_dt(state: OracleState, params: OracleParams) -> float:
    """Unsplit CFL step; the quotient is float64 as in V1 (shallow_water_2d.py:267-274)."""
    maximum = float(np.max(cfl_inverse_dt(state, params)))
    return min(params.cfl / maximum, params.max_dt) if maximum > 0.0 else float(params.max_dt)


class StepDiagnostics(NamedTuple):
    boundary_inflow: float  # signed volume entering through the outer faces this step (m^3)
    clip_volume: float      # volume created by max(h, 0) this step (m^3)
    rain_volume: float      # rain volume applied to active cells this step (m^3)


# ----------------------------------------------------------------------------
# Stochastic rounding of the rain source (an option of this build; the
# reference rounds to nearest).  Restates the device's sr_hash / sr_seed /
# sr_add bit for bit: "lowbias32" (C. Wellons, "Prospecting for hash
# functions", 2018) on uint32, TwoSum for the exact error of the nearest sum,
# and the float neighbour in the error's direction taken when u * gap < |e|.
# ----------------------------------------------------------------------------
_M32 = np.uint64(0xFFFFFFFF)


def sr_hash(x):
    """lowbias32 of uint32 values (array or scalar), as uint32."""
    x = np.asarray(x, dtype=np.uint64) & _M32
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & _M32
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & _M32
    x ^= x >> np.uint64(16)
    return x.astype(np.uint32)


def sr_seed(step: int) -> np.uint32:
    return sr_hash((np.uint64(step) * np.uint64(0x9E3779B9) + np.uint64(0x632BE5AB)) & _M32)


def sr_add(a, r, bits):
    """float32 a + r rounded stochastically with the uint32 random ``bits``."""
    a = np.asarray(a, np.float32)
    r = np.asarray(r, np.float32)
    s = (a + r).astype(np.float32)
    bb = (s - a).astype(np.float32)
    e = ((a - (s - bb)).astype(np.float32) + (r - bb).astype(np.float32)).astype(np.float32)
    d = np.where((e > 0) == (s >= 0), 1, -1).astype(np.int64)
    t = ((s.view(np.int32).astype(np.int64) + d) & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    with np.errstate(invalid="ignore"):
        gap = np.abs((t - s).astype(np.float32))
        u = ((np.asarray(bits, np.uint32) >> np.uint32(8)).astype(np.float32)
             * np.float32(2.0 ** -24))
        take = (e != 0) & ((u * gap).astype(np.float32) < np.abs(e))
    return np.where(take, t, s).astype(np.float32)


def finite_volume_step(state, params, dt, time_s, *, norm="sqrt", diagnostics=False,
                       sr_step=None):
    """One conservative, well-balanced HLL step (numerics.py:312-369).

    ``sr_step`` (this build's option): add the rain with :func:`sr_add`, drawn
    for (cell index, step ``sr_step``), as the device loops do with stochastic
    rain rounding.
    """
    T = params.dtype
    g = T(params.gravity)
    half_g = T(0.5 * params.gravity)
    dry = T(params.dry_tolerance)
    h, hu, hv = (np.asarray(x, T) for x in state)
    ny, nx = h.shape
    hg, hug, hvg = apply_ghost_cells(h, hu, hv, params.boundary_types)
    bg = np.pad(params.bed, 1, mode="edge")
    ag = np.pad(params.domain_mask, 1, mode="edge")

    # x faces: rows of the interior, columns k | k+1 of the ghosted grid, k = 0..nx
    r = slice(1, -1)
    fxh, fxn, fxt, cxA, cxB = face_fluxes(
        hg[r, :-1], hug[r, :-1], hvg[r, :-1], bg[r, :-1], ag[r, :-1],
        hg[r, 1:], hug[r, 1:], hvg[r, 1:], bg[r, 1:], ag[r, 1:],
        g, half_g, dry, "x",
    )
    # y faces: columns of the interior, rows k | k+1, k = 0..ny (normal = hv)
    fyh, fyn, fyt, cyA, cyB = face_fluxes(
        hg[:-1, r], hvg[:-1, r], hug[:-1, r], bg[:-1, r], ag[:-1, r],
        hg[1:, r], hvg[1:, r], hug[1:, r], bg[1:, r], ag[1:, r],
        g, half_g, dry, "y",
    )
    zero = T(0.0)
    # Per-side outputs in (h, hu, hv) order: the side's own correction is added
    # to the normal component (numerics.py:216-224, 275-283).
    x_left = (fxh + zero, fxn + cxA, fxt + zero)
    x_right = (fxh + zero, fxn + cxB, fxt + zero)
    y_bottom = (fyh + zero, fyt + zero, fyn + cyA)
    y_top = (fyh + zero, fyt + zero, fyn + cyB)

    dx_, dy_ = T(params.dx), T(params.dy)
    dt_a = T(dt)
    upd = []
    for k, u0 in enumerate((h, hu, hv)):
        div_x = (x_left[k][:, 1:] - x_right[k][:, :-1]) / dx_
        div_y = (y_bottom[k][1:, :] - y_top[k][:-1, :]) / dy_
        upd.append(u0 - dt_a * (div_x + div_y))
    rate = rainfall_rate(params, time_s)
    dtrate = T(dt * rate)
    if sr_step is None:
        h_star = upd[0] + dtrate * params.rainfall_multiplier
    else:
        cells = np.arange(ny * nx, dtype=np.uint64).reshape(ny, nx)
        h_star = sr_add(upd[0], dtrate * params.rainfall_multiplier,
                        sr_hash(cells ^ np.uint64(sr_seed(sr_step))))
    h_new = np.maximum(h_star, zero)
    hu_s, hv_s = upd[1], upd[2]
    if norm == "hypot":
        qnorm = np.hypot(hu_s, hv_s)
    else:
        qnorm = np.sqrt(hu_s * hu_s + hv_s * hv_s)
    dtg = T(dt * params.gravity)
    n2 = params.manning_n * params.manning_n
    denom = T(1.0) + ((dtg * n2) * qnorm) / np.power(np.maximum(h_new, dry), T(7.0 / 3.0))
    hu_new = hu_s / denom
    hv_new = hv_s / denom
    mask = params.domain_mask
    wet = mask & (h_new > dry)
    out = OracleState(
        np.where(mask, h_new, zero).astype(T),
        np.where(wet, hu_new, zero).astype(T),
        np.where(wet, hv_new, zero).astype(T),
    )
    if not diagnostics:
        return out
    cell = float(params.dx) * float(params.dy)
    inflow = (
        float(np.sum(fxh[:, 0], dtype=np.float64)) - float(np.sum(fxh[:, nx], dtype=np.float64))
    ) * float(params.dy) + (
        float(np.sum(fyh[0, :], dtype=np.float64)) - float(np.sum(fyh[ny, :], dtype=np.float64))
    ) * float(params.dx)
    clip = float(np.sum(np.where(mask, np.maximum(-h_star, zero), zero), dtype=np.float64)) * cell
    rain = float(dtrate) * float(np.sum(np.where(mask, params.rainfall_multiplier, 0), dtype=np.float64)) * cell
    return out, StepDiagnostics(dt * inflow, clip, rain)


# ----------------------------------------------------------------------------
# Output loop
# ----------------------------------------------------------------------------


class OracleResult(NamedTuple):
    states: OracleState            # (T, ny, nx) each
    output_times_s: np.ndarray
    final_time_s: float
    outputs_written: int
    steps_used: int
    dts: np.ndarray                # the dt of every step taken (float64)
    volume_initial: float
    volume_final: float
    rain_volume: float
    boundary_volume: float
    clip_volume: float

    @property
    def mass_balance_error(self) -> float:
        return (self.volume_final - self.volume_initial
                - self.rain_volume - self.boundary_volume - self.clip_volume)

    @property
    def mass_balance_rel(self) -> float:
        scale = max(abs(self.volume_initial) + abs(self.rain_volume)
                    + abs(self.boundary_volume) + abs(self.clip_volume), 1.0)
        return self.mass_balance_error / scale


def volume(state: OracleState, params: OracleParams) -> float:
    return float(np.sum(np.where(params.domain_mask, state.h, 0), dtype=np.float64)) * (
        float(params.dx) * float(params.dy)
    )


def integrate_to_outputs(initial, params, output_times_s, *, max_steps, norm="sqrt",
                         dt_schedule=None, rain_rounding="nearest"):
    """Adaptive loop to fixed output times (numerics.py:372-434), float64 clock.

    ``dt_schedule`` replays a recorded list of step sizes instead of the CFL
    rule (used to compare runs with a frozen schedule).
    """
    times = np.asarray(output_times_s, np.float64)
    if times.ndim != 1 or times.shape[0] < 1:
        raise ValueError("At least one one-dimensional output time is required.")
    if rain_rounding not in ("nearest", "stochastic"):
        raise ValueError("rain_rounding must be 'nearest' or 'stochastic'")
    if rain_rounding == "stochastic" and params.dtype != np.float32:
        raise ValueError("stochastic rain rounding is a float32 option")
    T = params.dtype
    state = OracleState(*(np.asarray(x, T) for x in initial))
    outs = [state]
    t = float(times[0])
    steps = 0
    written = 1
    dts = []
    v0 = volume(state, params)
    rain = bnd = clip = 0.0
    for target in times[1:]:
        while t < target - OUTPUT_TIME_SLACK_S and steps < max_steps:
            if dt_schedule is not None:
                dt = float(dt_schedule[steps])
            else:
                dt = min(compute_cfl_dt(state, params), float(target) - t)
            state, diag = finite_volume_step(
                state, params, dt, t, norm=norm, diagnostics=True,
                sr_step=steps if rain_rounding == "stochastic" else None)
            rain += diag.rain_volume
            bnd += diag.boundary_inflow
            clip += diag.clip_volume
            dts.append(dt)
            t = t + dt
            steps += 1
        written += int(t >= target - OUTPUT_TIME_SLACK_S)
        outs.append(state)
    stack = OracleState(*(np.stack([getattr(s, f) for s in outs]) for f in ("h", "hu", "hv")))
    return OracleResult(stack, times, t, written, steps, np.asarray(dts, np.float64),
                        v0, volume(state, params), rain, bnd, clip)


def run_steps(initial, params, n_steps, *, t0=0.0, norm="sqrt"):
    """``n_steps`` adaptive steps with no output target (dt = CFL dt)."""
    T = params.dtype
    state = OracleState(*(np.asarray(x, T) for x in initial))
    t = float(t0)
    dts = []
    for _ in range(n_steps):
        dt = compute_cfl_dt(state, params)
        state = finite_volume_step(state, params, dt, t, norm=norm)
        dts.append(dt)
        t += dt
    return state, t, np.asarray(dts)


def advance_physics_to_time(initial, params, start_time_s, target_time_s, *, max_substeps,
                            norm="sqrt"):
    """Fixed-length adaptive advance to one target (hybrid/coupled_forecast.py:123-161).

    ``max_substeps`` iterations of: if ``t < target - 1e-7`` take
    ``dt = min(cfl_dt, target - t)`` and step, else hold.  A clock that ends
    more than 1e-6 s short of the target gives an all-NaN state
    (coupled_forecast.py:152-156).  float64 clock, as in the rest of this
    oracle.  Returns ``(state, steps, t)``.
    """
    T = params.dtype
    state = OracleState(*(np.asarray(x, T) for x in initial))
    t = float(start_time_s)
    target = float(target_time_s)
    steps = 0
    for _ in range(int(max_substeps)):
        if not t < target - OUTPUT_TIME_SLACK_S:
            continue
        dt = min(compute_cfl_dt(state, params), target - t)
        state = finite_volume_step(state, params, dt, t, norm=norm)
        t = t + dt
        steps += 1
    if not t >= target - 1.0e-6:
        state = OracleState(*(np.full_like(x, np.nan) for x in state))
    return state, steps, t
