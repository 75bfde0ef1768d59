"""Tile-shape autotuner: rows per wavefront strip of the fused step, and the
number of resident wavefronts per compute unit.

A strip is 62 columns wide (one wavefront) and ``rows`` tall.  Taller strips
amortise the halo row and the strip prologue; shorter strips give more
wavefronts to hide memory latency and a smaller tail.  The best value depends
on the grid shape and the kernel variant, so it is measured, together with a
cap on resident wavefronts per compute unit (fewer concurrent row streams can
use HBM better): each candidate pair runs a short device loop on a copy of
the caller's state, timed with HIP
events, and the winner is cached in a JSON file keyed by device
architecture, grid shape and variant (``$SWE_TUNE_CACHE`` or
``~/.cache/hybrid_flood_amd/tuning.json``).  Every candidate gives
bit-identical states (the tiling does not change any cell's arithmetic).
"""

from __future__ import annotations

import json
import os
from pathlib import Path

import torch

from . import _lib
from .numerics import SWEParams, SWEState, context_for

# Heights between the round numbers let the last round of wavefronts fill the
# device better (the wavefront count is segs_x * ceil(ny / rows)).
CANDIDATE_ROWS = (4, 6, 8, 12, 14, 16, 17, 20, 21, 24, 32, 48, 64)
# Resident wavefronts per compute unit: 0 = as many as registers allow.
CANDIDATE_WAVES = (0, 20, 16, 12)
# Resident wavefronts per compute unit of the fused loop step when registers
# alone limit it (79 VGPRs: 6 per SIMD).
UNCAPPED_WAVES = 24
STRIP_COLS = 62


def round_rows(shape, waves, cus: int, rounds=range(1, 9), lo: int = 4, hi: int = 64) -> list[int]:
    """Strip heights whose wavefront count falls just under a whole number of
    rounds of the resident capacity (cus x waves per CU), so that the last
    round of a launch is (nearly) full: for k rounds, the strips per column
    are floor(k * slots / segs_x) and the height is ceil(ny / strips).  For
    many rounds the integer heights get coarse relative to a round, so the
    fill is only approximate there; the tuner measures every candidate."""
    ny, nx = int(shape[0]), int(shape[1])
    segs = (nx + STRIP_COLS - 1) // STRIP_COLS
    slots = cus * (waves if waves > 0 else UNCAPPED_WAVES)
    out = set()
    for k in rounds:
        strips = (k * slots) // segs
        if strips <= 0:
            continue
        r = -(-ny // strips)
        if lo <= r <= hi:
            out.add(r)
    return sorted(out)


def cache_path() -> Path:
    env = os.environ.get("SWE_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "hybrid_flood_amd" / "tuning.json"


def _load_cache() -> dict:
    try:
        return json.loads(cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _store_cache(data: dict) -> None:
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
        tmp.replace(p)
    except OSError:
        pass  # a read-only home only loses the cache


def _key(kind: str, shape, variant: str) -> str:
    arch = torch.cuda.get_device_properties(torch.cuda.current_device()).gcnArchName.split(":")[0]
    return f"{arch}/{kind}/{shape[0]}x{shape[1]}/{variant}"


def time_loop_rows(ctx, state: SWEState, rows: int, steps: int) -> float:
    """ms per step of `steps` device-loop steps with the given strip height."""
    lib = _lib.load()
    ctx.set_tile_rows(rows)
    work = SWEState(*(x.clone() for x in state))
    s = _lib.stream_handle(state.h.device)
    st = _lib.state_struct(*work)
    _lib.check(lib.swe_loop_begin(ctx.handle, st, 0.0, 2 * steps, s), "swe_loop_begin")
    _lib.check(lib.swe_loop_advance(ctx.handle, steps, s), "swe_loop_advance")  # warm
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    _lib.check(lib.swe_loop_advance(ctx.handle, steps, s), "swe_loop_advance")
    e1.record()
    torch.cuda.synchronize(state.h.device)
    _lib.check(lib.swe_loop_end(ctx.handle, st, None, s), "swe_loop_end")
    return e0.elapsed_time(e1) / steps


def tune_rows(params: SWEParams, state: SWEState, *, variant: str | None = None, steps: int = 32,
              candidates=None, waves=CANDIDATE_WAVES, use_cache: bool = True, finalists: int = 4,
              final_steps: int = 256) -> dict:
    """Pick and apply the fastest (strip height, waves per CU) for this grid; returns the record.

    ``candidates=None`` searches CANDIDATE_ROWS plus, for each cap, the
    heights of round_rows; explicit candidates are searched as given.  Every
    pair is timed over ``steps`` loop steps; the ``finalists`` fastest are
    then timed again over ``final_steps``, twice, and the best of those wins
    (short timings alone confuse tiles within a percent of each other)."""
    from .numerics import DEFAULT_VARIANT

    variant = variant or DEFAULT_VARIANT
    ctx = context_for(params, variant)
    key = _key("step2", tuple(state.h.shape), variant)
    cache = _load_cache() if use_cache else {}
    rec = cache.get(key)
    if rec is None:
        chunk_steps = max(2, steps - steps % 2)
        ctx.set_graph_chunk(chunk_steps)
        timings = {}
        cus = torch.cuda.get_device_properties(state.h.device).multi_processor_count
        for w in waves:
            ctx.set_waves_per_cu(int(w))
            rows = (sorted(set(CANDIDATE_ROWS) | set(round_rows(state.h.shape, int(w), cus)))
                    if candidates is None else [int(r) for r in candidates])
            for r in rows:
                timings[(int(r), int(w))] = time_loop_rows(ctx, state, int(r), chunk_steps)
        final = {}
        if finalists > 1 and len(timings) > 1:
            long_steps = max(2, final_steps - final_steps % 2)
            ctx.set_graph_chunk(min(long_steps, 64))
            for tile in sorted(timings, key=timings.get)[:finalists]:
                ctx.set_waves_per_cu(tile[1])
                final[tile] = min(time_loop_rows(ctx, state, tile[0], long_steps) for _ in range(2))
        pool = final or timings
        best = min(pool, key=pool.get)
        rec = {"rows": best[0], "waves_per_cu": best[1], "ms_per_step": pool[best],
               "candidates": {f"{r}x{w}": v for (r, w), v in timings.items()},
               "finalists": {f"{r}x{w}": v for (r, w), v in final.items()}}
        if use_cache:
            cache[key] = rec
            _store_cache(cache)
    ctx.set_tile_rows(int(rec["rows"]))
    ctx.set_waves_per_cu(int(rec.get("waves_per_cu", 0)))
    return rec


def tune_ensemble_rows(runner, *, t_end: float = 1.0e9, launches: int = 16,
                       candidates=(8, 12, 16, 20, 24, 32, 40, 48), use_cache: bool = True) -> dict:
    """Pick the strip height for an ensemble by timing whole ensemble launches."""
    shapes = sorted(tuple(m.params.bed.shape) for m in runner.members)
    key = _key("ensemble", (len(shapes), sum(a * b for a, b in shapes)), runner.variant)
    cache = _load_cache() if use_cache else {}
    rec = cache.get(key)
    if rec is None:
        timings = {}
        saved = [SWEState(*(x.clone() for x in m.state)) for m in runner.members]
        for r in candidates:
            runner.set_tile_rows(int(r))
            runner.set_graph_chunk(launches)
            runner.begin(t_end, max_steps=2 * launches)
            runner.advance(launches)
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            runner.advance(launches)
            e1.record()
            torch.cuda.synchronize()
            runner.end()
            timings[int(r)] = e0.elapsed_time(e1) / launches
            for m, s in zip(runner.members, saved):
                for dst, src in zip(m.state, s):
                    dst.copy_(src)
        best = min(timings, key=timings.get)
        rec = {"rows": best, "ms_per_launch": timings[best],
               "candidates": {str(k): v for k, v in timings.items()}}
        if use_cache:
            cache[key] = rec
            _store_cache(cache)
    runner.set_tile_rows(int(rec["rows"]))
    return rec
