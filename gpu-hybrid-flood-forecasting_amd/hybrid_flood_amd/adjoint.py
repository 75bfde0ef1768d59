"""Checkpointed adjoint of the device time loop behind ``torch.autograd.Function``.

The forward is the product path: ``integrate_to_outputs`` on the device,
recording its float64 step sizes and, every ``C = ceil(sqrt(max_steps))``
steps, a checkpoint of the state (``swe_set_checkpoints``; the loop's own
stores write it).  The schedule is then data: the reverse sweep walks the
segments backwards, recomputes each from its checkpoint into a C-state
buffer with ``swe_step`` (bit-identical to the loop's steps, same kernels and
the same float64 scalars), then reverses it step by step with the
hand-written adjoint kernel (``swe_adjoint_sweep``, or
``swe_adjoint_sweep_f64`` for float64 cotangents).  Memory is ~2 sqrt(max_steps)
states; compute is ~2 forward steps + 1 reverse step per step.  Without
in-loop checkpoints the forward rebuilds them by one more replay, every
ceil(sqrt(K)) steps.

``rollout_depths(manning_n, rain_multiplier, model)`` returns the depths at
the model's output times and differentiates with respect to the per-cell
Manning field and the scalar rain multiplier (the rain source is
``multiplier * base_multiplier``).  The reference has no adjoint; its one
reverse-differentiable driver form is the fixed-length scan of
hybrid/coupled_forecast.py:123-161.  jax.grad of the reference is not
available here, so the gradients are checked against central finite
differences of the FP64 oracle with the same frozen schedule.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .numerics import SWEParams, SWEState, _output_times_f64, context_for, make_params

OUTPUT_SLACK_S = 1.0e-7


@dataclass
class RolloutModel:
    """Everything of a rollout except the two differentiated parameters."""

    bed: torch.Tensor
    domain_mask: torch.Tensor
    base_multiplier: torch.Tensor
    h0: SWEState
    output_times_s: np.ndarray
    boundary_types: tuple = (0, 0, 0, 0)
    rainfall_times_s: np.ndarray = field(default_factory=lambda: np.zeros(1, np.float32))
    rainfall_rates_m_s: np.ndarray = field(default_factory=lambda: np.zeros(1, np.float32))
    dx: float = 1.0
    cfl: float = 0.45
    dry_tolerance: float = 1.0e-4
    max_dt: float = 10.0
    max_steps: int = 20_000
    variant: str = "fast"
    checkpoint_every: int = 0  # 0 = ceil(sqrt(max_steps)) recorded in the loop, else ceil(sqrt(K))
    record_checkpoints: bool = True  # store checkpoints during the forward loop (no replay pass)
    adjoint_precision: str = "f32"  # cotangent arithmetic of the reverse sweep: "f32" or "f64"
    adjoint_pipeline: bool = True  # replay the next segment on a second stream during a reversal

    def params_for(self, manning_n: torch.Tensor, rain_multiplier) -> SWEParams:
        mult = (float(rain_multiplier) * self.base_multiplier.double()).float().contiguous()
        return make_params(self.bed, manning_n.detach().float().contiguous(), self.domain_mask, mult,
                           boundary_types=self.boundary_types,
                           rainfall_times_s=self.rainfall_times_s,
                           rainfall_rates_m_s=self.rainfall_rates_m_s, dx=self.dx, cfl=self.cfl,
                           dry_tolerance=self.dry_tolerance, max_dt=self.max_dt,
                           device=self.bed.device)


@dataclass
class Tape:
    params: SWEParams
    dts: np.ndarray          # float64 step sizes
    ts: np.ndarray           # float64 start time of every step
    out_steps: np.ndarray    # step count at which each output was taken
    checkpoints: dict        # step index -> SWEState
    every: int
    steps: int


def _clone(s: SWEState) -> SWEState:
    return SWEState(*(x.clone() for x in s))


def rollout_forward(params: SWEParams, state0: SWEState, times, *, max_steps: int,
                    variant: str, every: int = 0, record: bool = True):
    """Device-loop forward; returns the depth stack (T, ny, nx) and the tape.

    With ``record`` the loop itself stores a checkpoint every ``every`` steps
    (default ceil(sqrt(max_steps)), since the step count is known only after
    the run); a checkpoint that falls on an output step is taken from the
    output cube.  Without it (or when the checkpoint storage does not fit in
    device memory) the checkpoints are rebuilt afterwards by replaying the
    recorded schedule, every ceil(sqrt(K)) steps.
    """
    from .numerics import integrate_to_outputs

    ctx = context_for(params, variant)
    dev = state0.h.device
    lib = _lib.load()
    log = torch.zeros(max(max_steps, 1), dtype=torch.float64, device=dev)
    C = every if every > 0 else max(1, int(math.ceil(math.sqrt(max(max_steps, 1)))))
    ck = None
    if record:
        try:
            ck = torch.empty((3, max_steps // C + 1) + tuple(state0.h.shape), dtype=torch.float32,
                             device=dev)
        except torch.cuda.OutOfMemoryError:
            ck = None
    _lib.check(lib.swe_set_dt_log(ctx.handle, log.data_ptr(), int(log.numel())), "swe_set_dt_log")
    if ck is not None:
        _lib.check(lib.swe_set_checkpoints(ctx.handle, ck[0].data_ptr(), ck[1].data_ptr(),
                                           ck[2].data_ptr(), C, int(ck.shape[1])),
                   "swe_set_checkpoints")
    try:
        res = integrate_to_outputs(state0, params, times, max_steps=max_steps, variant=variant)
    finally:
        lib.swe_set_dt_log(ctx.handle, None, 0)
        lib.swe_set_checkpoints(ctx.handle, None, None, None, 0, 0)
    K = int(res.steps_used)
    dts = log[:K].cpu().numpy().astype(np.float64)
    t64 = _output_times_f64(times)
    ts = np.empty(K, np.float64)
    t = float(t64[0])
    for k in range(K):  # same float64 additions as the device clock
        ts[k] = t
        t = t + float(dts[k])
    out_steps = np.full(t64.shape[0], K, np.int64)
    out_steps[0] = 0
    clock = np.empty(K + 1, np.float64)  # clock[k] = time after k steps
    clock[0] = float(t64[0])
    for k in range(K):
        clock[k + 1] = clock[k] + float(dts[k])
    m = 1
    for k in range(K + 1):
        while m < t64.shape[0] and clock[k] >= t64[m] - OUTPUT_SLACK_S:
            out_steps[m] = k
            m += 1
    cps = {0: _clone(state0)}
    if ck is not None:
        # Slot s holds the state after s*C steps, unless an output was taken at
        # that step: the output cube holds it then.
        recorded = {int(k): i for i, k in enumerate(out_steps) if i >= 1}
        for start in range(C, K, C):
            i = recorded.get(start)
            src = (res.states.h[i], res.states.hu[i], res.states.hv[i]) if i is not None else \
                (ck[0, start // C], ck[1, start // C], ck[2, start // C])
            cps[start] = SWEState(*(x.clone() for x in src))
        del ck
    else:
        # Checkpoints by replay of the recorded schedule, one segment per call.
        C = every if every > 0 else max(1, int(math.ceil(math.sqrt(max(K, 1)))))
        s = _lib.stream_handle(dev)
        for start in range(0, K - C, C) if K > C else ():
            nxt = SWEState(*(torch.empty_like(x) for x in state0))
            _lib.check(lib.swe_replay(ctx.handle, _lib.state_struct(*cps[start]), _dptr(dts, start),
                                      _dptr(ts, start), C, _lib.state_struct(*nxt), s), "swe_replay")
            cps[start + C] = nxt
    tape = Tape(params, dts, ts, out_steps, cps, C, K)
    return res.states.h, res, tape


def replay_depths(params: SWEParams, state0: SWEState, tape: Tape, *, variant: str) -> torch.Tensor:
    """Depths (T, ny, nx) at the tape's output steps, replaying its frozen dt
    schedule under ``params`` (e.g. perturbed parameters for a finite-difference
    check of the adjoint, which differentiates the same frozen-schedule map)."""
    ctx = context_for(params, variant)
    lib = _lib.load()
    s = _lib.stream_handle(state0.h.device)
    st = _clone(state0)
    out = [st.h.clone()]
    pos = 0
    for m in range(1, len(tape.out_steps)):
        k = int(tape.out_steps[m])
        if k > pos:
            _lib.check(lib.swe_replay(ctx.handle, _lib.state_struct(*st), _dptr(tape.dts, pos),
                                      _dptr(tape.ts, pos), k - pos, _lib.state_struct(*st), s),
                       "swe_replay")
            pos = k
        out.append(st.h.clone())
    return torch.stack(out)


def _dptr(arr: np.ndarray, offset: int):
    """ctypes double* to arr[offset:] (arr must be a contiguous float64 array)."""
    return ctypes.cast(arr.ctypes.data + 8 * int(offset), ctypes.POINTER(ctypes.c_double))


def rollout_backward(tape: Tape, grad_depths: torch.Tensor, *, variant: str,
                     precision: str = "f32", pipeline: bool = True):
    """Reverse sweep; returns (grad_manning, grad_multiplier_field, lambda_0).

    ``precision="f32"`` runs the cotangents, the reverse arithmetic and the
    sensitivities in float32; ``"f64"`` in float64 (gfx950 runs float64
    add/mul/FMA at the float32 rate; the forward recompute stays the float32
    kernel's either way).  On the tested rollouts the two agree to 1e-6.

    ``pipeline=True`` replays the next (earlier) checkpoint segment on a
    second stream into a second segment cube while the current segment is
    reversed, so the replay's kernels fill the GPU beside the reverse step's;
    the arithmetic and its order are unchanged (``False`` runs them one
    after the other on the caller's stream, with the same results).
    """
    if precision not in ("f64", "f32"):
        raise ValueError("precision must be 'f64' or 'f32'")
    f64 = precision == "f64"
    ctx = context_for(tape.params, variant)
    lib = _lib.load()
    h0 = tape.checkpoints[0].h
    dev = h0.device
    s = _lib.stream_handle(dev)
    N = h0.numel()
    cdt = torch.float64 if f64 else torch.float32
    gn = torch.zeros(h0.shape, dtype=cdt, device=dev)
    gm = torch.zeros(h0.shape, dtype=cdt, device=dev)
    lam = SWEState(*(torch.zeros(h0.shape, dtype=cdt, device=dev) for _ in range(3)))
    tmp = SWEState(*(torch.empty(h0.shape, dtype=cdt, device=dev) for _ in range(3)))
    lam_struct = _lib.state_struct_f64 if f64 else _lib.state_struct
    sweep = lib.swe_adjoint_sweep_f64 if f64 else lib.swe_adjoint_sweep
    outs_at: dict[int, list[int]] = {}
    for m, k in enumerate(tape.out_steps):
        outs_at.setdefault(int(k), []).append(m)
    K = tape.steps
    C = tape.every
    dts = np.ascontiguousarray(tape.dts, np.float64)
    ts = np.ascontiguousarray(tape.ts, np.float64)
    # Segments (start, end) in the order they are reversed.
    segments = []
    end = K
    for start in sorted(tape.checkpoints, reverse=True):
        if start < end:
            segments.append((start, end))
            end = start
    # Segment cubes: slot j holds the state entering step start + j.  Two of
    # them when pipelining (the one being reversed and the one being filled).
    cubes = [torch.empty((3, max(C, 1)) + tuple(h0.shape), dtype=h0.dtype, device=dev)
             for _ in range(2 if pipeline and len(segments) > 1 else 1)]
    main = torch.cuda.current_stream(dev)
    side = torch.cuda.Stream(device=dev) if len(cubes) > 1 else main
    side_handle = ctypes.c_void_p(side.cuda_stream) if len(cubes) > 1 else s
    if side is not main:
        # The side stream reads the checkpoints (filled on the main stream by
        # the forward) and writes the cubes (allocated on the main stream,
        # possibly over memory main-stream kernels still read): order all of
        # that after the main stream's queued work, and tell the allocator the
        # cubes are used on the side stream.
        side.wait_stream(main)
        for cube in cubes:
            cube.record_stream(side)
    freed = [None] * len(cubes)  # event: the last reverse sweep that read each cube ended

    def replay(i):
        """Fill cube i % len(cubes) with segment i's states (on the side stream)."""
        start, end = segments[i]
        cube = cubes[i % len(cubes)]
        with torch.cuda.stream(side):
            if freed[i % len(cubes)] is not None:
                side.wait_event(freed[i % len(cubes)])
            for f in range(3):
                cube[f, 0].copy_(tape.checkpoints[start][f])
            if end - start > 1:
                _lib.check(lib.swe_replay_store(
                    ctx.handle, _lib.state_struct(*tape.checkpoints[start]), _dptr(dts, start),
                    _dptr(ts, start), end - start - 1, cube[0, 1].data_ptr(), cube[1, 1].data_ptr(),
                    cube[2, 1].data_ptr(), side_handle), "swe_replay_store")
            ready = torch.cuda.Event()
            ready.record(side)
        return ready

    def inject(k):
        for m in outs_at.get(k, ()):
            lam.h.add_(grad_depths[m])

    inject(K)
    pos = K  # lam is the cotangent of the state after `pos` steps
    ready = replay(0) if segments else None
    for i, (start, end) in enumerate(segments):
        main.wait_event(ready)
        nxt = replay(i + 1) if len(cubes) > 1 and i + 1 < len(segments) else None
        cube = cubes[i % len(cubes)]
        # reverse down to `start`, stopping at every output step to add its loss cotangent
        stops = sorted((k for k in outs_at if start <= k < pos), reverse=True) + [start]
        for k in stops:
            if k < pos:
                j = k - start
                _lib.check(sweep(
                    ctx.handle, cube[0, j].data_ptr(), cube[1, j].data_ptr(), cube[2, j].data_ptr(),
                    pos - k, _dptr(dts, k), _dptr(ts, k), lam_struct(*lam), lam_struct(*tmp),
                    gn.data_ptr(), gm.data_ptr(), s), "swe_adjoint_sweep")
                pos = k
                inject(k)
        done = torch.cuda.Event()
        done.record(main)
        freed[i % len(cubes)] = done
        if nxt is None and i + 1 < len(segments):
            nxt = replay(i + 1)
        ready = nxt
    return gn, gm, lam


class _RolloutFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, manning_n, rain_multiplier, model: RolloutModel):
        params = model.params_for(manning_n, rain_multiplier)
        depths, res, tape = rollout_forward(params, model.h0, model.output_times_s,
                                            max_steps=model.max_steps, variant=model.variant,
                                            every=model.checkpoint_every,
                                            record=model.record_checkpoints)
        ctx.tape = tape
        ctx.model = model
        ctx.res = res
        ctx.mult_dtype = rain_multiplier.dtype if isinstance(rain_multiplier, torch.Tensor) else torch.float32
        return depths

    @staticmethod
    def backward(ctx, grad_depths):
        model: RolloutModel = ctx.model
        gn, gm, _ = rollout_backward(ctx.tape, grad_depths.contiguous().float(), variant=model.variant,
                                     precision=model.adjoint_precision,
                                     pipeline=model.adjoint_pipeline)
        g_alpha = (gm.double() * model.base_multiplier.double()).sum()
        return gn.float(), g_alpha.to(ctx.mult_dtype), None


def rollout_depths(manning_n: torch.Tensor, rain_multiplier: torch.Tensor, model: RolloutModel):
    """Depths (T, ny, nx) at the model's output times, differentiable in both parameters."""
    return _RolloutFunction.apply(manning_n, rain_multiplier, model)


def depth_rmse(depths: torch.Tensor, target: torch.Tensor, mask: torch.Tensor | None = None,
               skip_first: bool = True) -> torch.Tensor:
    """Depth RMSE over the output times (excluding the initial state) and cells."""
    d = depths[1:] if skip_first else depths
    t = target[1:] if skip_first else target
    err = (d - t) ** 2
    if mask is not None:
        m = mask.bool().expand_as(err)
        return torch.sqrt(err[m].mean())
    return torch.sqrt(err.mean())
