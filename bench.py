"""Throughput of the fused shallow-water step on one MI355X.

Workload: the 4096 x 4096 FP32 urban dam-break case (dx = 1 m, wet start
h = 0.05 m with a 0.5 m square in the centre, no rain, cfl 0.45, one open edge),
advanced by the device-resident adaptive loop: one fused kernel per step
(ghosts, reconstruction, HLL, rain, clip, friction, wet/dry mask; its
epilogue folds the CFL maximum of the new state into atomic slots and stores
FP64 mass partials; every workgroup of the next step derives dt from the
slots).  Inputs are resident in HBM before timing starts.

Timing protocol: the tile autotune, then ``--settle`` untimed steps of the
same loop, ``--warmup`` untimed steps, then exactly ``--steps`` timed steps in
``--segments`` segments, all queued back to back with no host synchronisation
in between.  The settle steps are there because the GPU's power management
slows steps 20-150 after an idle gap by 10-25 % (measured: constant cycles and
instructions per dispatch, so not the kernel); without them the timed region
of a short run would fall inside that transient.  HIP events on the stream the
kernels run on bracket every segment; ``value`` = cell-updates/s over all
timed steps, ``segments`` gives each segment's ms per step and their median.

Also reported: the in-house streaming probes (copy, read, and the packed
step's streams without its arithmetic) as the achievable-bandwidth ceiling;
the batched ensemble (64 uneven urban members, 3600 s each) and the
checkpointed adjoint (1024^2 urban, dry start with rain, ~2000 steps, with a
frozen-schedule finite-difference check) against the plain step; the FP64
mass balance of each benchmark configuration (configurations 2 and 4 also
with the rain source rounded stochastically, ``swe_set_rain_rounding``); and,
in the ``cpu_baseline``
leg, the FP32 oracle's step timed on the host: its C/OpenMP restatement on
every host thread (the reported value; ``oracle/swe_step_omp.c``) and the
NumPy oracle on one thread, each over a few steps of the 4096^2 workload and
the 256^2 sloped-plane configuration in full (median of reps), together with the accuracy record of the GPU's run of that configuration
against the oracle (per-field max-abs and RMSE against the FP64 oracle,
wet-mask mismatches against the FP32 oracle).

Prints ONE JSON line.  ``roofline`` prices the step at 37 algorithmic bytes
per cell-update against 8 TB/s.

    python bench.py [--steps 2000] [--warmup 20] [--variant fast|reference]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-hybrid-flood-forecasting_amd"))

METRIC = "cell-updates/s & %HBM roofline at 4096²; ensemble & adjoint vs plain step"
BYTES_PER_CELL = 37  # h,hu,hv,bed,n,mult read (24) + mask (1) + h,hu,hv written (12)
HBM_PEAK_GBS = 8000.0


def _progress(msg: str) -> None:
    """One line per finished phase on stderr (the JSON line alone goes to stdout)."""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def _even_divisor(k: int, cap: int = 64) -> int:
    for c in range(min(cap, k), 1, -1):
        if k % c == 0 and c % 2 == 0:
            return c
    return 0


def _segment_sizes(steps: int, segments: int) -> list[int]:
    n = max(1, min(segments, steps))
    base, extra = divmod(steps, n)
    return [base + (1 if k < extra else 0) for k in range(n)]


def time_loop(params, h0, steps, warmup, settle, segments, variant, rows=None):
    """Advance the device loop: ``settle`` + ``warmup`` untimed steps, then
    ``steps`` timed steps in segments, queued without host synchronisation."""
    import ctypes

    import numpy as np
    import torch

    from hybrid_flood_amd import _lib
    from hybrid_flood_amd import numerics as nm

    dev = params.bed.device
    lib = _lib.load()
    ctx = nm.context_for(params, variant)
    tuned = None
    if rows:
        ctx.set_tile_rows(rows)
    else:
        from hybrid_flood_amd.autotune import tune_rows

        h = torch.as_tensor(h0, device=dev).contiguous()
        tuned = tune_rows(params, nm.SWEState(h, torch.zeros_like(h), torch.zeros_like(h)),
                          variant=variant)
    sizes = _segment_sizes(steps, segments)
    chunk = _even_divisor(sizes[0])
    ctx.set_graph_chunk(chunk)
    h = torch.as_tensor(h0, device=dev).contiguous()
    state = nm.SWEState(h, torch.zeros_like(h), torch.zeros_like(h))
    s = _lib.stream_handle(dev)
    st = _lib.state_struct(*state)
    total = settle + warmup + steps
    _lib.check(lib.swe_loop_begin(ctx.handle, st, 0.0, total, s), "swe_loop_begin")
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    if settle:
        _lib.check(lib.swe_loop_advance(ctx.handle, settle, s), "swe_loop_advance")
    if warmup:
        _lib.check(lib.swe_loop_advance(ctx.handle, warmup, s), "swe_loop_advance")
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(len(sizes) + 1)]
    evs[0].record()
    for k, n in enumerate(sizes):
        _lib.check(lib.swe_loop_advance(ctx.handle, n, s), "swe_loop_advance")
        evs[k + 1].record()
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t0
    ms = evs[0].elapsed_time(evs[-1])
    seg = [evs[k].elapsed_time(evs[k + 1]) / n for k, n in enumerate(sizes)]
    info = _lib.SweRunInfo()
    _lib.check(lib.swe_loop_end(ctx.handle, st, ctypes.byref(info), s), "swe_loop_end")
    if info.steps_used != total:
        raise RuntimeError(f"device loop took {info.steps_used} steps, expected {total}")
    return {"ms": ms, "wall_s": wall, "info": info.as_dict(), "chunk": chunk,
            "segments": {"count": len(sizes), "steps": sizes,
                         "ms_per_step": seg, "median_ms_per_step": float(np.median(seg))},
            "tile": ({"rows": tuned["rows"], "waves_per_cu": tuned["waves_per_cu"], "autotuned": True}
                     if tuned else {"rows": rows, "waves_per_cu": 0, "autotuned": False})}


def window_probe(params, h0, variant: str, rows, idle_s: float = 0.5, seg: int = 2,
                 segments: int = 200, per_window: int = 10) -> dict:
    """Step time and average shader clock in windows of ``seg * per_window``
    steps: after ``idle_s`` seconds of an idle GPU, and then again under
    continuous load.  Between segments of the device loop a one-thread probe
    kernel samples the shader clock counter and the real-time counter on the
    same stream, so each window's clock is (cycles elapsed) / (time elapsed)
    on the device itself, without a profiler."""
    import ctypes

    import numpy as np
    import torch

    from hybrid_flood_amd import _lib
    from hybrid_flood_amd import numerics as nm

    dev = params.bed.device
    lib = _lib.load()
    ctx = nm.context_for(params, variant)
    if rows:
        ctx.set_tile_rows(rows)
    ctx.set_graph_chunk(_even_divisor(seg))
    h = torch.as_tensor(h0, device=dev).contiguous()
    state = nm.SWEState(h, torch.zeros_like(h), torch.zeros_like(h))
    s = _lib.stream_handle(dev)
    st = _lib.state_struct(*state)
    total = 400 + 2 * seg * segments
    _lib.check(lib.swe_loop_begin(ctx.handle, st, 0.0, total, s), "swe_loop_begin")
    _lib.check(lib.swe_loop_advance(ctx.handle, 400, s), "swe_loop_advance")
    clk = torch.zeros(2 * (segments + 1), dtype=torch.int64, device=dev)
    vp = ctypes.c_void_p
    out = {}
    for phase, idle in (("after_idle", idle_s), ("continuous", 0.0)):
        torch.cuda.synchronize(dev)
        if idle:
            time.sleep(idle)
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(segments + 1)]
        _lib.check(lib.swe_stream_probe(4, None, None, vp(clk.data_ptr()), 1, s), "clock probe")
        evs[0].record()
        for k in range(segments):
            _lib.check(lib.swe_loop_advance(ctx.handle, seg, s), "swe_loop_advance")
            evs[k + 1].record()
            _lib.check(lib.swe_stream_probe(4, None, None, vp(clk.data_ptr()), k + 2, s),
                       "clock probe")
        torch.cuda.synchronize(dev)
        c = clk.cpu().numpy().reshape(-1, 2).astype(np.float64)
        ms = np.array([evs[k].elapsed_time(evs[k + 1]) for k in range(segments)])
        wall_rate = (c[-1, 1] - c[0, 1]) / (ms.sum() * 1e-3)  # real-time counter ticks per s
        wins = []
        for w in range(segments // per_window):
            a, b = w * per_window, (w + 1) * per_window
            dt_s = (c[b, 1] - c[a, 1]) / wall_rate
            wins.append({"steps": seg * per_window,
                         "us_per_step": float(ms[a:b].sum() / (seg * per_window) * 1e3),
                         "shader_mhz": float((c[b, 0] - c[a, 0]) / dt_s / 1e6)})
        out[phase] = {"idle_s": idle, "realtime_counter_hz": float(wall_rate), "windows": wins}
        _progress(f"window probe: {phase} done")
    info = _lib.SweRunInfo()
    _lib.check(lib.swe_loop_end(ctx.handle, st, ctypes.byref(info), s), "swe_loop_end")
    return out


def stream_probes(cells: int, reps: int = 10) -> dict:
    """Achievable bandwidth of this device, measured with the library's own
    streaming kernels: a copy, a read, and the packed step's streams (12 B of
    state and 8 B, or 5 B with Manning classes, of static record read, 12 B
    written per cell) over the step's cell count."""
    import ctypes

    import torch

    from hybrid_flood_amd import _lib

    lib = _lib.load()
    s = _lib.stream_handle()
    out = {}
    n = 128 * 1024 * 1024  # floats: 512 MiB per buffer
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)

    def run(mode, src, src2, dst, count, nbytes):
        for _ in range(3):
            _lib.check(lib.swe_stream_probe(mode, src, src2, dst, count, s), "swe_stream_probe")
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            _lib.check(lib.swe_stream_probe(mode, src, src2, dst, count, s), "swe_stream_probe")
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        return {"us_per_launch": ms * 1e3, "GB_s": nbytes / (ms * 1e-3) / 1e9,
                "frac_of_8TBs": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}

    vp = ctypes.c_void_p
    out["copy"] = run(0, vp(a.data_ptr()), None, vp(b.data_ptr()), n, 8 * n)
    out["read"] = run(1, vp(a.data_ptr()), None, vp(b.data_ptr()), n, 4 * n)
    del a, b
    st = torch.ones(3 * cells, device="cuda")
    sx = torch.ones(2 * cells, device="cuda")
    dst = torch.empty(3 * cells, device="cuda")
    out["step_streams"] = run(2, vp(st.data_ptr()), vp(sx.data_ptr()), vp(dst.data_ptr()), cells,
                              32 * cells)
    out["step_streams"]["note"] = ("the packed step's streams with the (bed, n^2) record "
                                   "(32 B per cell) without its arithmetic")
    del sx
    sc = torch.ones(((cells + 63) // 64) * 80, device="cuda")  # 320-byte blocks
    out["step_streams_class"] = run(3, vp(st.data_ptr()), vp(sc.data_ptr()), vp(dst.data_ptr()),
                                    cells, 29 * cells)
    out["step_streams_class"]["note"] = ("the packed step's streams with the class-indexed record "
                                         "(29 B per cell) without its arithmetic: the memory-only "
                                         "time of one step of this workload")
    return out


def config1_gpu_run(variant: str) -> dict:
    """The 256^2 sloped-plane configuration (dx 2 m, dry start, trapezoid
    rain, cfl 0.9, 600 s, outputs every 60 s) on the GPU: output states, the
    loop's own dt schedule and its FP64 mass balance."""
    import torch

    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd import synthetic as sy
    from hybrid_flood_amd.adjoint import rollout_forward

    dev = torch.device("cuda", 0)
    c = sy.sloped_bumps_case(256)
    params = nm.make_params(c.bed, c.manning_n, c.domain_mask, boundary_types=c.boundary_types,
                            rainfall_times_s=c.rain_times_s, rainfall_rates_m_s=c.rain_rates_m_s,
                            dx=c.dx, cfl=c.cfl, dry_tolerance=c.dry_tolerance, max_dt=c.max_dt,
                            device=dev)
    h = torch.as_tensor(c.h0, device=dev)
    times = c.output_times()
    _, res, tape = rollout_forward(params, nm.SWEState(h, torch.zeros_like(h), torch.zeros_like(h)),
                                   times, max_steps=20000, variant=variant, record=False)
    return {"case": c, "times": times, "states": [x.cpu().numpy() for x in res.states],
            "dts": tape.dts, "steps": int(tape.steps), "diagnostics": res.diagnostics,
            "variant": variant}


def cpu_baseline(case, steps: int, cfg1: dict, reps: int, omp_steps: int) -> tuple[dict, dict]:
    """The oracle's float32 step on the host: the reported value is the C/OpenMP
    restatement (oracle/swe_step_omp.c, every host thread) over ``omp_steps``
    adaptive steps of the 4096^2 workload, beside the NumPy oracle on one
    thread over ``steps`` steps; the 256^2 configuration runs in full with both
    (median of ``reps``).  The NumPy oracle is also the checker of the GPU's run
    of that configuration: replayed in float32 and float64 along the GPU's own
    dt schedule, it gives the accuracy record."""
    import subprocess

    import numpy as np

    sys.path.insert(0, ROOT)
    from oracle import swe_oracle as o
    from oracle import swe_step_omp as comp

    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)

    p = o.OracleParams(bed=case.bed, manning_n=case.manning_n, domain_mask=case.domain_mask,
                       boundary_types=case.boundary_types, rainfall_times_s=case.rain_times_s,
                       rainfall_rates_m_s=case.rain_rates_m_s, dx=case.dx, cfl=case.cfl,
                       dry_tolerance=case.dry_tolerance, max_dt=case.max_dt)
    z = np.zeros_like(case.h0)
    t0 = time.perf_counter()
    o.run_steps(o.OracleState(case.h0, z, z), p, steps)
    el = time.perf_counter() - t0
    _progress(f"cpu baseline: {steps} oracle steps at {case.bed.shape[0]}^2, {el:.1f} s")
    n = case.bed.size
    c1 = cfg1["case"]
    op1 = o.OracleParams(bed=c1.bed, manning_n=c1.manning_n, domain_mask=c1.domain_mask,
                         boundary_types=c1.boundary_types, rainfall_times_s=c1.rain_times_s,
                         rainfall_rates_m_s=c1.rain_rates_m_s, dx=c1.dx, cfl=c1.cfl,
                         dry_tolerance=c1.dry_tolerance, max_dt=c1.max_dt)
    z1 = np.zeros_like(c1.h0)
    secs = []
    r1 = None
    for _ in range(reps):
        t1 = time.perf_counter()
        r1 = o.integrate_to_outputs(o.OracleState(c1.h0, z1, z1), op1, cfg1["times"],
                                    max_steps=20000)
        secs.append(time.perf_counter() - t1)
        _progress(f"cpu baseline: configuration 1 rep {len(secs)}/{reps}, {secs[-1]:.1f} s")
    med = float(np.median(secs))
    # C/OpenMP restatement: one untimed step, then the timed ones.
    threads = comp.threads()
    st0 = o.OracleState(case.h0, z, z)
    comp.run_steps(st0, p, 1)
    t0 = time.perf_counter()
    comp.run_steps(st0, p, omp_steps)
    el_c = time.perf_counter() - t0
    _progress(f"cpu baseline: {omp_steps} C/OpenMP steps at {case.bed.shape[0]}^2 on {threads} "
              f"threads, {el_c:.1f} s")
    secs_c = []
    rc1 = None
    for _ in range(reps):
        t1 = time.perf_counter()
        rc1 = comp.integrate_to_outputs(o.OracleState(c1.h0, z1, z1), op1, cfg1["times"],
                                        max_steps=20000)
        secs_c.append(time.perf_counter() - t1)
    med_c = float(np.median(secs_c))
    cfg1_desc = ("sloped plane with bumps and a channel, 256x256, dx 2 m, dry start, "
                 "trapezoid rain to 50 mm/h, cfl 0.9, 600 s, 11 outputs")
    base = {
        "value": n * omp_steps / el_c, "unit": "cell-updates/s", "cores": threads, "kind": "port",
        "sample": f"{omp_steps} adaptive steps (CFL dt + step) of the same {case.bed.shape[0]}x"
                  f"{case.bed.shape[1]} workload, C/OpenMP restatement of the FP32 oracle "
                  f"(oracle/swe_step_omp.c), {threads} threads, {el_c:.1f} s",
        "host_cpu_count": os.cpu_count(),
        "config1_full": {
            "workload": cfg1_desc, "steps": int(rc1[1]), "reps": reps, "seconds": secs_c,
            "median_s": med_c, "cell_updates_per_s": c1.bed.size * rc1[1] / med_c,
        },
        "numpy_single_thread": {
            "value": n * steps / el, "unit": "cell-updates/s", "cores": 1,
            "sample": f"{steps} adaptive steps of the same workload, NumPy FP32 oracle, {el:.1f} s",
            "config1_full": {
                "workload": cfg1_desc, "steps": int(r1.steps_used), "reps": reps, "seconds": secs,
                "median_s": med, "cell_updates_per_s": c1.bed.size * r1.steps_used / med,
            },
        },
    }
    rec = accuracy_record(cfg1)
    rec["oracle_own_schedule_steps"] = int(r1.steps_used)
    return base, rec


def accuracy_record(cfg1: dict) -> dict:
    """Accuracy record of a GPU run of configuration 1 against the oracle
    replayed along that run's dt schedule in float32 and float64."""
    sys.path.insert(0, ROOT)
    from oracle import accuracy as acc
    from oracle import swe_oracle as o

    c1 = cfg1["case"]
    op1 = o.OracleParams(bed=c1.bed, manning_n=c1.manning_n, domain_mask=c1.domain_mask,
                         boundary_types=c1.boundary_types, rainfall_times_s=c1.rain_times_s,
                         rainfall_rates_m_s=c1.rain_rates_m_s, dx=c1.dx, cfl=c1.cfl,
                         dry_tolerance=c1.dry_tolerance, max_dt=c1.max_dt)
    rec = acc.record(cfg1["states"], op1, c1.h0, cfg1["times"], cfg1["dts"])
    rec["workload"] = ("256x256 sloped plane (config 1), 600 s, 11 outputs, GPU "
                       f"{cfg1['variant']} variant")
    rec["gpu_steps"] = cfg1["steps"]
    rec["gpu_mass_balance_rel"] = cfg1["diagnostics"]["mass_balance_rel"]
    return rec


def config2_mass_balance(variant: str, rain_rounding: str = "nearest") -> dict:
    """The 1024^2 urban configuration (one open edge, 1-min rain series over
    10800 s, 181 outputs) on the GPU: steps, wall time and FP64 mass balance,
    with the rain rounded to nearest (the reference's) or stochastically."""
    import numpy as np
    import torch

    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd import synthetic as sy

    dev = torch.device("cuda", 0)
    u = sy.urban_case(1024)
    params = nm.make_params(u.bed, u.manning_n, u.domain_mask, boundary_types=u.boundary_types,
                            rainfall_times_s=u.rain_times_s, rainfall_rates_m_s=u.rain_rates_m_s,
                            dx=u.dx, cfl=u.cfl, dry_tolerance=u.dry_tolerance, max_dt=u.max_dt,
                            device=dev)
    h = torch.as_tensor(u.h0, device=dev)
    times = np.arange(0.0, 10801.0, 60.0).astype(np.float32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = nm.integrate_to_outputs(nm.SWEState(h, torch.zeros_like(h), torch.zeros_like(h)), params,
                                times, max_steps=200000, variant=variant,
                                rain_rounding=rain_rounding)
    wall = time.perf_counter() - t0
    d = r.diagnostics
    return {"rain_rounding": rain_rounding, "outputs_written": int(r.outputs_written),
            "steps": int(r.steps_used), "wall_s": wall,
            "mass_balance_rel": d["mass_balance_rel"], "nonfinite": d["nonfinite"]}


def cpu_config2_full() -> dict:
    """The 1024^2 urban configuration in full (10800 s, 181 outputs) with the
    C/OpenMP restatement of the FP32 oracle on every host thread (about a
    minute on 16 EPYC threads; opt-in with --cpu-config2)."""
    import subprocess

    import numpy as np

    sys.path.insert(0, ROOT)
    from hybrid_flood_amd import synthetic as sy
    from oracle import swe_oracle as o
    from oracle import swe_step_omp as comp

    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    u = sy.urban_case(1024)
    p = o.OracleParams(bed=u.bed, manning_n=u.manning_n, domain_mask=u.domain_mask,
                       boundary_types=u.boundary_types, rainfall_times_s=u.rain_times_s,
                       rainfall_rates_m_s=u.rain_rates_m_s, dx=u.dx, cfl=u.cfl,
                       dry_tolerance=u.dry_tolerance, max_dt=u.max_dt)
    z = np.zeros_like(u.h0)
    times = np.arange(0.0, 10801.0, 60.0).astype(np.float32)
    t0 = time.perf_counter()
    _, steps, t, written = comp.integrate_to_outputs(o.OracleState(u.h0, z, z), p, times,
                                                     max_steps=200000)
    el = time.perf_counter() - t0
    _progress(f"cpu baseline: configuration 2 in full, {steps} steps, {el:.1f} s")
    return {"workload": "urban 1024x1024, dx 2 m, one open edge, 1-min rain over 10800 s, "
                        "181 outputs", "threads": comp.threads(), "steps": steps,
            "outputs_written": written, "final_time_s": t, "seconds": el,
            "cell_updates_per_s": u.bed.size * steps / el}


def bench_ensemble(variant: str, count: int, seed: int, t_end: float, plain_cups: float):
    """Config 4: `count` urban members of uneven size, n x U[0.7,1.3], rain x U[0.5,2]."""
    import torch

    from hybrid_flood_amd import synthetic as sy
    from hybrid_flood_amd.autotune import tune_ensemble_rows
    from hybrid_flood_amd.ensemble import EnsembleRunner, member_from_case

    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    members = []
    sides = []
    for side, mseed, n_scale, rain_scale in sy.ensemble_members(count, seed):
        case = sy.urban_case(side, dx=2.0, seed=mseed, duration_s=t_end)
        members.append(member_from_case(case, device=dev, n_scale=n_scale, rain_scale=rain_scale))
        sides.append(side)
    gen_s = time.perf_counter() - t0
    start = [m.state.h.clone() for m in members]  # hu = hv = 0 at the start
    runner = EnsembleRunner(members, variant=variant)
    tune = tune_ensemble_rows(runner)
    chunk = 32
    runner.set_graph_chunk(chunk)
    total_cells = int(sum(m.params.bed.numel() for m in members))

    def timed_run():
        for m, h0 in zip(members, start):
            m.state.h.copy_(h0)
            m.state.hu.zero_()
            m.state.hv.zero_()
        torch.cuda.synchronize(dev)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        evf = torch.cuda.Event(enable_timing=True)
        w0 = time.perf_counter()
        ev0.record()
        runner.begin(t_end, max_steps=200_000)
        launches = 0
        full_launches = 0  # launches while every member was still running
        last = count
        full_open = True
        while True:
            runner.advance(chunk)
            launches += chunk
            running = runner.running()
            if full_open and running == count:
                full_launches = launches
                evf.record()
            else:
                full_open = False
            if running == 0:
                break
            if running < last:
                runner.compact()
                last = running
        ev1.record()
        torch.cuda.synchronize(dev)
        wall = time.perf_counter() - w0
        rc, infos = runner.end()
        ms = ev0.elapsed_time(ev1)
        ms_full = ev0.elapsed_time(evf) if full_launches else float("nan")
        return rc, infos, ms, ms_full, wall, launches, full_launches

    rc, infos, ms, ms_full, wall, launches, full_launches = timed_run()
    cell_steps = sum(m.params.bed.numel() * i["steps_used"] for m, i in zip(members, infos))
    cups = cell_steps / (ms * 1e-3)
    cups_full = total_cells * full_launches / (ms_full * 1e-3) if full_launches else None
    # the same job with every member's rain rounded stochastically
    runner.set_rain_rounding("stochastic")
    rc_s, infos_s, ms_s, _, _, _, _ = timed_run()
    runner.set_rain_rounding("nearest")
    cell_steps_s = sum(m.params.bed.numel() * i["steps_used"] for m, i in zip(members, infos_s))
    stochastic = {"status": rc_s, "ms": ms_s, "cell_updates_per_s": cell_steps_s / (ms_s * 1e-3),
                  "vs_plain_step": cell_steps_s / (ms_s * 1e-3) / plain_cups,
                  "max_abs_mass_balance_rel": max(abs(i["mass_balance_rel"]) for i in infos_s)}
    return {
        "workload": f"{count} urban members, sides from {{256,512,1024,2048}} (seed {seed}), "
                    f"dx 2 m, {t_end:.0f} s each, per-member CFL dt, graph-captured launches of one "
                    "wavefront per strip over all running members",
        "members": count, "cells": int(sum(m.params.bed.numel() for m in members)),
        "member_sides": {str(k): sides.count(k) for k in sorted(set(sides))},
        "launches": launches, "member_steps_total": int(sum(i["steps_used"] for i in infos)),
        "ms": ms, "wall_s": wall, "generation_s": gen_s,
        "cell_updates_per_s": cups, "vs_plain_step": cups / plain_cups,
        "all_members_running": {"launches": full_launches,
                                "cell_updates_per_s": cups_full,
                                "vs_plain_step": (cups_full / plain_cups) if cups_full else None},
        "max_member_steps": int(max(i["steps_used"] for i in infos)),
        "rows": tune["rows"], "tune_ms_per_launch": tune.get("candidates"), "status": rc,
        "max_abs_mass_balance_rel": max(abs(i["mass_balance_rel"]) for i in infos),
        "stochastic_rain": stochastic,
    }


def bench_adjoint(variant: str, n: int, steps: int, plain_cups: float):
    """Checkpointed adjoint of a depth-RMSE loss at 10 output times (urban, dry start, rain),
    with a frozen-schedule central finite-difference check of both gradients."""
    import numpy as np
    import torch

    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd import synthetic as sy
    from hybrid_flood_amd.adjoint import RolloutModel, replay_depths, rollout_depths, rollout_forward

    dev = torch.device("cuda", 0)
    case = sy.urban_case(n, dx=2.0, seed=7)
    mask_t = torch.ones((n, n), dtype=torch.bool, device=dev)
    h = torch.zeros((n, n), dtype=torch.float32, device=dev)
    state0 = nm.SWEState(h, torch.zeros_like(h), torch.zeros_like(h))
    # The rollout horizon is the simulated time `steps` CFL steps reach.
    p0 = nm.make_params(case.bed, case.manning_n, case.domain_mask, boundary_types=case.boundary_types,
                        rainfall_times_s=case.rain_times_s, rainfall_rates_m_s=case.rain_rates_m_s,
                        dx=case.dx, device=dev)
    probe = nm.SWEState(*(x.clone() for x in state0))
    horizon = nm.run_steps(probe, p0, steps, variant=variant, diagnostics=True)["final_time_s"]
    times = np.linspace(0.0, horizon, 11).astype(np.float32)
    model = RolloutModel(bed=p0.bed, domain_mask=mask_t, base_multiplier=torch.ones_like(h),
                         h0=state0, output_times_s=times, boundary_types=case.boundary_types,
                         rainfall_times_s=case.rain_times_s, rainfall_rates_m_s=case.rain_rates_m_s,
                         dx=case.dx, max_steps=4 * steps, variant=variant)
    hidden = torch.as_tensor(sy.smooth_field(n, np.random.default_rng(11), 0.01, 0.06), device=dev)
    with torch.no_grad():
        target = rollout_depths(hidden, torch.tensor(1.3), model).clone()
    n0 = torch.as_tensor(case.manning_n, device=dev)

    def rmse(d):
        return torch.sqrt(((d[1:].double() - target[1:].double()) ** 2).mean())

    # Plain forward of the same rollout and forward + gradient, each after one
    # untimed warm-up call (first-use allocations and graph capture excluded,
    # as in the step benchmark), median of `reps` timed calls.
    reps = 3
    pl = model.params_for(n0, 1.0)

    def plain():
        return nm.integrate_to_outputs(state0, pl, times, max_steps=4 * steps, variant=variant)

    def gradient():
        n_var = n0.clone().requires_grad_(True)
        alpha = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
        loss = rmse(rollout_depths(n_var, alpha, model))
        loss.backward()
        return loss, n_var, alpha

    def timed(fn):
        out, ts = None, []
        fn()
        for _ in range(reps):
            torch.cuda.synchronize(dev)
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize(dev)
            ts.append(time.perf_counter() - t0)
        return out, float(np.median(ts))

    res, fwd_s = timed(plain)
    (loss, n_var, alpha), grad_s = timed(gradient)
    K = res.steps_used
    cells = n * n
    # Frozen-schedule central differences (the map the adjoint differentiates).
    _, _, tape = rollout_forward(pl, state0, times, max_steps=4 * steps, variant=variant)
    eps_a = 1e-2
    fd_a = (float(rmse(replay_depths(model.params_for(n0, 1 + eps_a), state0, tape, variant=variant)))
            - float(rmse(replay_depths(model.params_for(n0, 1 - eps_a), state0, tape, variant=variant)))) / (2 * eps_a)
    dn = torch.as_tensor(sy.smooth_field(n, np.random.default_rng(3), -0.01, 0.01), device=dev).double()
    eps_n = 1e-1
    fd_dir = (float(rmse(replay_depths(model.params_for((n0.double() + eps_n * dn).float(), 1.0), state0,
                                       tape, variant=variant)))
              - float(rmse(replay_depths(model.params_for((n0.double() - eps_n * dn).float(), 1.0), state0,
                                         tape, variant=variant)))) / (2 * eps_n)
    g_a = float(alpha.grad)
    adj_dir = float((n_var.grad.double() * dn).sum())
    return {
        "workload": f"urban {n}x{n}, dx 2 m, dry start, 1-min rain series, {K} CFL steps "
                    f"({horizon:.0f} s), loss = depth RMSE at 10 output times vs a run with a "
                    "hidden smooth n-field in [0.01, 0.06] and rain x1.3, checkpoints every "
                    "ceil(sqrt(max_steps)) steps stored by the forward loop",
        "timing": f"median of {reps} calls after one untimed warm-up call, each",
        "steps": K, "forward_s": fwd_s, "forward_plus_gradient_s": grad_s,
        "gradient_vs_forward_time": grad_s / fwd_s,
        "gradient_cell_steps_per_s": cells * K / grad_s,
        "vs_plain_step": (cells * K / grad_s) / plain_cups,
        "loss": float(loss.detach()),
        "forward_mass_balance_rel": res.diagnostics["mass_balance_rel"],
        "fd_check": {
            "d_loss_d_rain_multiplier": {"adjoint": g_a, "central_fd": fd_a, "eps": eps_a,
                                         "rel_err": abs(g_a - fd_a) / max(abs(fd_a), 1e-300)},
            "d_loss_along_smooth_n_direction": {"adjoint": adj_dir, "central_fd": fd_dir, "eps": eps_n,
                                                "rel_err": abs(adj_dir - fd_dir) / max(abs(fd_dir), 1e-300)},
            "schedule": "frozen (the dt sequence of the unperturbed run)",
        },
        "grad_manning_l2": float(n_var.grad.norm()),
    }


# Algorithmic bytes per cell of the path's callers (what each kernel must read
# and write once): features 12 state + 1 mask + 8 static in, 24 out;
# correction = sums pass (h, mask, the NHWC residual's line) 17 + apply pass
# (state, mask, residual) 25 in, 12 out; staging 12 + 1 + 8 in, 24 out;
# depth correction 4 + 4 in, 4 out; output schema 12 + 1 in, 16 out.
CALLER_BYTES = {"network_features": 45, "apply_residual_correction": 54, "stage_features": 45,
                "correct_depth": 12, "conserved_to_primitive": 29}


def bench_callers(params, case, reps: int = 20) -> dict:
    """The step's callers at the roofline grid: the coupled forecast's
    features and residual correction, batched staging, depth correction and
    the output schema.  Each op runs `reps` times back to back between HIP
    events on torch's current stream (the stream the library launches on);
    the time includes the Python call overhead between launches."""
    import numpy as np
    import torch

    from hybrid_flood_amd import coupled_forecast as cf
    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd import staging as stg
    from hybrid_flood_amd import structured_io as sio

    dev = params.bed.device
    ny, nx = params.bed.shape
    cells = ny * nx
    h = torch.as_tensor(case.h0, device=dev).contiguous()
    state = nm.SWEState(h, (0.1 * h).contiguous(), (-0.05 * h).contiguous())
    bed = params.bed
    terrain = ((bed - bed.mean()) / bed.std()).contiguous()
    rough = ((params.manning_n - params.manning_n.mean()) / params.manning_n.std()).contiguous()
    hi = cf.HybridInputs(terrain_standardized=terrain, roughness_standardized=rough,
                         state_mean=(0.1, 0.0, 0.0), state_std=(0.2, 0.05, 0.05),
                         rainfall_scale_mm_hr=50.0, residual_std=(0.01, 0.01, 0.01))
    feats = torch.empty((1, ny, nx, 6), dtype=torch.float32, device=dev)
    g = torch.Generator(device=dev).manual_seed(7)
    resid = 1e-3 * torch.randn((1, ny, nx, 3), generator=g, device=dev)
    out = nm.SWEState(*(torch.empty_like(h) for _ in range(3)))
    prim = torch.stack([h, 0.1 * torch.ones_like(h), torch.zeros_like(h)], dim=-1)[None].contiguous()
    rain = np.array([5.0], np.float32)
    pred = 1e-3 * torch.randn((1, 3, ny, nx), generator=g, device=dev)
    depth = h[None].contiguous()
    ops = {
        "network_features": lambda: cf.network_features(state, params, hi, 600.0, out=feats),
        "apply_residual_correction": lambda: cf.apply_residual_correction(
            state, resid, params, correction_relaxation=0.5, out=out),
        "stage_features": lambda: stg.stage_features(prim, rain, terrain, rough,
                                                     params.domain_mask, (0.1, 0.0, 0.0),
                                                     (0.2, 0.05, 0.05), 50.0),
        "correct_depth": lambda: stg.correct_depth(depth, pred),
        "conserved_to_primitive": lambda: sio._primitive(params, h, state.hu, state.hv),
    }
    res = {"grid": f"{ny}x{nx}", "reps": reps}
    for name, op in ops.items():
        op()
        op()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            op()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        gbs = CALLER_BYTES[name] * cells / (ms * 1e-3) / 1e9
        res[name] = {"api_ms": ms, "bytes_per_cell": CALLER_BYTES[name], "api_GB_s": gbs,
                     "api_frac_of_8TBs": gbs / HBM_PEAK_GBS}
    res["note"] = ("api_* = per call of the Python API, launch overhead included; the kernels' own "
                   "durations are in profiles/callers_4096_kernel_stats.md (rocprofv3)")
    return res


def copy_calibration() -> int:
    """Four device-to-device copies of 256 MiB each: 256 MiB read and 256 MiB
    written per dispatch, the reference for the per-dispatch HBM byte counters."""
    import torch

    n = 256 * 1024 * 1024 // 4
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(4):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"dispatches": 4, "bytes_read_per_dispatch": n * 4,
                                           "bytes_written_per_dispatch": n * 4}}), flush=True)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--settle", type=int, default=400,
                    help="untimed steps before the warm-up, so the timed region does not start "
                         "in the GPU's power-management transient after an idle gap")
    ap.add_argument("--segments", type=int, default=10, help="event-bracketed timed segments")
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--variant", default="fast", choices=["fast", "reference"])
    ap.add_argument("--rows", type=int, default=0, help="rows per wavefront strip (0 = autotune)")
    ap.add_argument("--cpu-steps", type=int, default=3, help="oracle steps for cpu_baseline (0 = skip)")
    ap.add_argument("--cpu-omp-steps", type=int, default=50,
                    help="steps of the C/OpenMP restatement for cpu_baseline's value")
    ap.add_argument("--cpu-config2", action="store_true",
                    help="also run the 1024^2 configuration in full with the C/OpenMP oracle")
    ap.add_argument("--cpu-reps", type=int, default=5,
                    help="timed repetitions of the oracle on the 256^2 configuration")
    ap.add_argument("--extra", action="store_true", help="also time the other kernel variant")
    ap.add_argument("--ensemble", type=int, default=64, help="ensemble members (0 = skip)")
    ap.add_argument("--ensemble-t", type=float, default=3600.0, help="simulated seconds per member")
    ap.add_argument("--adjoint-n", type=int, default=1024, help="adjoint grid side (0 = skip)")
    ap.add_argument("--adjoint-steps", type=int, default=2000)
    ap.add_argument("--callers", type=int, default=1,
                    help="time the coupled-forecast, staging and output kernels (0 = skip)")
    ap.add_argument("--probes", type=int, default=1, help="streaming bandwidth probes (0 = skip)")
    ap.add_argument("--configs", type=int, default=1,
                    help="run configurations 1 and 2 for their mass balance (0 = skip)")
    ap.add_argument("--window-probe", action="store_true",
                    help="only print the step time and shader clock per 20-step window after an "
                         "idle gap and under continuous load (the 4096^2 loop) and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 4 device copies of 256 MiB (the known-size kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args(argv)
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build runs on one GPU "
              "(the step is measured on one MI355X; multi-GPU slab sharding is design only).",
              file=sys.stderr)
        return 2
    if args.steps < 1 or args.warmup < 0 or args.settle < 0 or args.segments < 1:
        print("bench.py: --steps >= 1, --warmup >= 0, --settle >= 0, --segments >= 1",
              file=sys.stderr)
        return 2

    import torch

    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd import synthetic as sy

    dev = torch.device("cuda", 0)
    case = sy.dam_break_urban_case(args.size, n_steps=args.steps)
    params = nm.make_params(case.bed, case.manning_n, case.domain_mask,
                            boundary_types=case.boundary_types,
                            rainfall_times_s=case.rain_times_s,
                            rainfall_rates_m_s=case.rain_rates_m_s, dx=case.dx, cfl=case.cfl,
                            dry_tolerance=case.dry_tolerance, max_dt=case.max_dt, device=dev)
    cells = case.bed.size
    if args.window_probe:
        print(json.dumps({"window_probe": window_probe(params, case.h0, args.variant,
                                                       args.rows or 21)}), flush=True)
        return 0
    _progress(f"case {args.size}^2 built; tuning and timing the loop")
    r = time_loop(params, case.h0, args.steps, args.warmup, args.settle, args.segments,
                  args.variant, args.rows)
    _progress(f"timed loop: {r['ms'] / args.steps * 1e3:.1f} us per step, tile {r['tile']}")
    ms_per_step = r["ms"] / args.steps
    value = cells * args.steps / (r["ms"] * 1e-3)
    achieved = BYTES_PER_CELL * cells / (ms_per_step * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"dam_break_urban_{args.size}: {args.size}x{args.size} FP32 cells, dx=1 m, "
                        "wet start 0.05 m + 0.5 m central square, no rain, cfl 0.45, "
                        "adaptive dt, fused step + CFL/mass epilogue",
            "cells": cells,
            "kernel_variant": args.variant,
            "graph_chunk": r["chunk"],
            "tile": r["tile"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_cell_update": BYTES_PER_CELL,
        },
        "settle_steps": args.settle,
        "segments": r["segments"],
        "wall_s": r["wall_s"],
        "mass_balance_rel": r["info"]["mass_balance_rel"],
        "simulated_s": r["info"]["final_time_s"],
    }
    prof = os.path.join(ROOT, "profiles", "step_fast_4096_traffic.json")
    if args.variant == "fast" and os.path.exists(prof):
        with open(prof) as f:
            t = json.load(f)
        line["roofline"]["traffic"] = t.get("hbm_bytes_per_launch")
        line["roofline"]["traffic_source"] = "profiles/step_fast_4096_traffic.json (rocprofv3 --pmc)"
        if "valu" in t:  # the step is co-limited by vector-instruction issue
            line["roofline"]["valu_busy_frac"] = t["valu"]["valu_busy_frac"]
            line["roofline"]["valu_per_wavefront_row"] = t["valu"]["valu_per_wavefront_row"]
    if args.probes > 0:
        line["bandwidth_probes"] = stream_probes(cells)
        _progress("bandwidth probes done")
    if args.extra:
        other = "reference" if args.variant == "fast" else "fast"
        k = min(args.steps, 400)
        ro = time_loop(params, case.h0, k, args.warmup, args.settle, args.segments, other,
                       args.rows)
        line["other_variant"] = {"variant": other, "steps": k,
                                 "cell_updates_per_s": cells * k / (ro["ms"] * 1e-3),
                                 "ms_per_step": ro["ms"] / k}
    mb = {"config3_roofline_run": line["mass_balance_rel"]}
    cfg1 = None
    if args.configs > 0:
        cfg1 = config1_gpu_run(args.variant)
        mb["config1"] = cfg1["diagnostics"]["mass_balance_rel"]
        c2 = config2_mass_balance(args.variant)
        mb["config2"] = c2["mass_balance_rel"]
        c2s = config2_mass_balance(args.variant, "stochastic")
        mb["config2_stochastic_rain"] = c2s["mass_balance_rel"]
        line["config2_run"] = c2
        line["config2_run_stochastic_rain"] = c2s
        _progress("configurations 1 and 2 done")
    if args.ensemble > 0:
        line["ensemble"] = bench_ensemble(args.variant, args.ensemble, 42, args.ensemble_t, value)
        mb["config4_max_abs_over_members"] = line["ensemble"]["max_abs_mass_balance_rel"]
        mb["config4_max_abs_over_members_stochastic_rain"] = (
            line["ensemble"]["stochastic_rain"]["max_abs_mass_balance_rel"])
        _progress("ensemble done")
    if args.adjoint_n > 0:
        line["adjoint"] = bench_adjoint(args.variant, args.adjoint_n, args.adjoint_steps, value)
        mb["config5_forward"] = line["adjoint"]["forward_mass_balance_rel"]
        _progress("adjoint done")
    line["mass_balance_rel_by_config"] = mb
    if args.callers > 0:
        line["callers"] = bench_callers(params, case)
        _progress("callers done")
    if args.cpu_steps > 0:
        if cfg1 is None:
            cfg1 = config1_gpu_run(args.variant)
        _progress("cpu baseline: timing the oracle on the host")
        line["cpu_baseline"], line["accuracy"] = cpu_baseline(case, args.cpu_steps, cfg1,
                                                              args.cpu_reps, args.cpu_omp_steps)
        if args.cpu_config2:
            line["cpu_baseline"]["config2_full"] = cpu_config2_full()
        other = "reference" if args.variant == "fast" else "fast"
        line[f"accuracy_{other}_variant"] = accuracy_record(config1_gpu_run(other))
        _progress(f"accuracy record of the {other} variant done")
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
