"""The benchmark workloads as GPU parity cases: the two smaller ones whole,
and a 512^2 crop of the 4096^2 roofline workload against the FP64 oracle.

* Sloped plane with bumps and a channel, 256^2, dx 2 m, dry start, trapezoid
  rain peaking at 50 mm/h, 600 s, cfl 0.9, dry 1e-3, output every 60 s: both
  kernel variants against the FP32 oracle over the whole run.  The step
  counts must match (the dt sequence follows the same CFL maxima); depth
  within 1e-5 m and momentum within 5e-5 m^2/s of the oracle (measured on
  MI355X: 1.3e-7 / 4e-8 for the reference variant, 3.2e-7 / 3.5e-6 for the
  fast one); FP64 mass balance below 1e-6.
* Urban DEM, 1024^2, dx 2 m, one open (east) edge, 1-min rain series over
  10800 s, 181 outputs: its first 3000 s (1339 steps) against the C/OpenMP
  transcription of the FP32 oracle (same step count and final time, no
  wet-mask mismatch, depth within 2e-6 m, momentum within 5e-7 m^2/s); the
  whole run, too long for a test's oracle, through size-independent
  properties -- every output written, final time, no non-finite value,
  depth >= 0, outflow only through the open edge, the rain volume within 1e-3
  of the hyetograph's integral (the loop applies the rate at each step's
  start), and the FP64 mass balance within twice the per-step drift of the
  float32 oracle on the same urban workload times the step count
  (tests/test_oracle_mass_balance.py; measured 2.5e-6 against ~1.3e-5).
* Dam-break urban crop, 512^2 of the 4096^2 roofline workload, 7 s: both
  variants against the FP64 oracle stepped with the GPU's dt schedule.
* The 4096^2 roofline workload itself: one loop step at the benchmark's tile
  against the FP32 and FP64 oracles, and its first 100 adaptive steps (each
  side taking its own CFL dt) against the C/OpenMP transcription of the FP32
  oracle.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _params(c, dev):
    from hybrid_flood_amd import numerics as nm

    return nm.make_params(c.bed, c.manning_n, c.domain_mask, boundary_types=c.boundary_types,
                          rainfall_times_s=c.rain_times_s, rainfall_rates_m_s=c.rain_rates_m_s,
                          dx=c.dx, cfl=c.cfl, dry_tolerance=c.dry_tolerance, max_dt=c.max_dt,
                          device=dev)


def _zero_state(c, dev):
    import torch

    from hybrid_flood_amd import numerics as nm

    h = torch.as_tensor(c.h0, device=dev)
    return nm.SWEState(h, torch.zeros_like(h), torch.zeros_like(h))


@pytest.fixture(scope="module")
def sloped_bumps_oracle():
    from hybrid_flood_amd import synthetic as sy
    from oracle import swe_oracle as o

    c = sy.sloped_bumps_case(256)
    op = o.OracleParams(bed=c.bed, manning_n=c.manning_n, domain_mask=c.domain_mask,
                        boundary_types=c.boundary_types, rainfall_times_s=c.rain_times_s,
                        rainfall_rates_m_s=c.rain_rates_m_s, dx=c.dx, cfl=c.cfl,
                        dry_tolerance=c.dry_tolerance, max_dt=c.max_dt)
    times = np.arange(0.0, 601.0, 60.0).astype(np.float32)
    z = np.zeros_like(c.h0)
    return c, times, o.integrate_to_outputs(o.OracleState(c.h0, z, z), op, times, max_steps=20000)


@pytest.mark.parametrize("variant", ["reference", "fast"])
def test_sloped_bumps_256_against_oracle(cuda, sloped_bumps_oracle, variant):
    from hybrid_flood_amd import numerics as nm

    c, times, ref = sloped_bumps_oracle
    r = nm.integrate_to_outputs(_zero_state(c, cuda), _params(c, cuda), times, max_steps=20000,
                                variant=variant)
    d = r.diagnostics
    assert int(r.outputs_written) == 11 == ref.outputs_written
    assert int(r.steps_used) == len(ref.dts)
    assert abs(d["mass_balance_rel"]) < 1e-6
    assert d["rain_volume"] == pytest.approx(ref.rain_volume, rel=1e-9)
    tol = {"h": 1e-5, "hu": 5e-5, "hv": 5e-5}
    for f, atol in tol.items():
        got = getattr(r.states, f).cpu().numpy()
        want = getattr(ref.states, f)
        assert np.abs(got - want).max() <= atol, f


def test_urban_1024_181_outputs_properties(cuda):
    import torch

    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd import synthetic as sy

    u = sy.urban_case(1024)
    times = np.arange(0.0, 10801.0, 60.0).astype(np.float32)
    r = nm.integrate_to_outputs(_zero_state(u, cuda), _params(u, cuda), times, max_steps=200000,
                                variant="fast")
    d = r.diagnostics
    assert int(r.outputs_written) == 181 and float(r.final_time_s) == 10800.0
    assert d["nonfinite"] == 0 and bool(torch.isfinite(r.states.h).all())
    assert float(r.states.h.min()) >= 0.0
    assert d["boundary_volume"] < 0.0  # water leaves through the open east edge only
    t = u.rain_times_s.astype(np.float64)
    q = u.rain_rates_m_s.astype(np.float64)
    closed = float(np.sum(0.5 * (q[1:] + q[:-1]) * np.diff(t))) * u.bed.size * u.dx * u.dx
    assert d["rain_volume"] == pytest.approx(closed, rel=1e-3)
    # float32 state over ~27k steps: at most twice the per-step drift of the
    # float32 oracle on the same urban workload (the FP64 oracle closes to
    # 1e-13 there); measured here: 2.5e-6 against a bar of ~1.3e-5
    from test_oracle_mass_balance import URBAN_FP32_DRIFT_PER_STEP

    assert abs(d["mass_balance_rel"]) <= 2.0 * URBAN_FP32_DRIFT_PER_STEP * int(r.steps_used)


@pytest.fixture(scope="module")
def dam_break_crop():
    """A 512^2 window of the 4096^2 dam-break urban workload (rows and columns
    1536..2047: a quarter of it inside the 0.5 m square, so the break runs
    through the window), run as a domain of its own with the workload's
    boundary types.  The crop's edges are walls or open edges where the full
    workload has water continuing: the comparison is of the kernels on the
    workload's terrain, bed and depths, not a cut-out of the 4096^2 run."""
    from hybrid_flood_amd import synthetic as sy

    full = sy.dam_break_urban_case(4096)
    s = slice(1536, 2048)
    return dict(bed=full.bed[s, s].copy(), manning_n=full.manning_n[s, s].copy(),
                domain_mask=full.domain_mask[s, s].copy(), h0=full.h0[s, s].copy(),
                boundary_types=full.boundary_types, rain_t=full.rain_times_s,
                rain_r=full.rain_rates_m_s, dx=full.dx, cfl=full.cfl, dry=full.dry_tolerance,
                max_dt=full.max_dt)


# Max |GPU - FP64 oracle| after the 95 steps of 7 s, measured on MI355X:
# reference variant h 8.1e-7 m, hu 1.1e-6, hv 8.9e-7 m^2/s; fast variant
# 7.0e-7, 1.0e-6, 8.7e-7.  The bars leave a factor of ~6-10.
CROP_TOL = {"reference": {"h": 5e-6, "hu": 1e-5, "hv": 1e-5},
            "fast": {"h": 5e-6, "hu": 1e-5, "hv": 1e-5}}


@pytest.mark.parametrize("variant", ["reference", "fast"])
def test_dam_break_crop_against_fp64_oracle(cuda, dam_break_crop, variant):
    """The roofline workload's terrain and dam break on a 512^2 crop: the GPU
    loop (its own CFL dt schedule, FP64 clock) against the FP64 oracle stepped
    with the same dt schedule (a frozen schedule: the oracle's own CFL dt
    would differ in the last bits and shift every later step)."""
    import torch

    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd.adjoint import rollout_forward
    from oracle import swe_oracle as o

    c = dam_break_crop
    params = nm.make_params(c["bed"], c["manning_n"], c["domain_mask"],
                            boundary_types=c["boundary_types"], rainfall_times_s=c["rain_t"],
                            rainfall_rates_m_s=c["rain_r"], dx=c["dx"], cfl=c["cfl"],
                            dry_tolerance=c["dry"], max_dt=c["max_dt"], device=cuda)
    h = torch.as_tensor(c["h0"], device=cuda)
    state = nm.SWEState(h, torch.zeros_like(h), torch.zeros_like(h))
    times = np.array([0.0, 7.0], np.float32)
    _, res, tape = rollout_forward(params, state, times, max_steps=2000, variant=variant,
                                   record=False)
    assert res.outputs_written == 2 and 40 < tape.steps < 400  # 95 measured
    op = o.OracleParams(bed=c["bed"], manning_n=c["manning_n"], domain_mask=c["domain_mask"],
                        boundary_types=c["boundary_types"], rainfall_times_s=c["rain_t"],
                        rainfall_rates_m_s=c["rain_r"], dx=c["dx"], cfl=c["cfl"],
                        dry_tolerance=c["dry"], max_dt=c["max_dt"], dtype=np.float64)
    z = np.zeros_like(c["h0"])
    st = o.OracleState(c["h0"], z, z)
    # the first dt: FP32 CFL maximum on the GPU, FP64 on the oracle
    assert float(tape.dts[0]) == pytest.approx(o.compute_cfl_dt(st, op), rel=1e-6)
    for k in range(tape.steps):
        st = o.finite_volume_step(st, op, float(tape.dts[k]), float(tape.ts[k]))
    errs = {}
    for f in ("h", "hu", "hv"):
        got = getattr(res.states, f)[-1].cpu().numpy().astype(np.float64)
        errs[f] = float(np.abs(got - getattr(st, f)).max())
    print(variant, tape.steps, errs)
    for f, atol in CROP_TOL[variant].items():
        assert errs[f] <= atol, (f, errs)
    assert float(o.volume(st, op)) == pytest.approx(
        float(res.states.h[-1].double().sum()) * c["dx"] ** 2, rel=1e-6)


@pytest.fixture(scope="module")
def dam_break_4096_after_30(cuda):
    """The 4096^2 roofline workload advanced 30 steps by the packed device loop
    at the benchmark's tile (21-row strips, 20 wavefronts per CU): the state,
    its clock, and the workload."""
    import torch

    from hybrid_flood_amd import _lib
    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd import synthetic as sy

    c = sy.dam_break_urban_case(4096)
    params = _params(c, cuda)
    st = _zero_state(c, cuda)
    st = nm.SWEState(*(x.clone() for x in st))
    print("4096^2: 30 steps of the packed loop", flush=True)
    info = nm.run_steps(st, params, 30, variant="fast", diagnostics=True)
    torch.cuda.synchronize()
    return c, params, st, float(info["final_time_s"])


@pytest.mark.parametrize("variant,rows,waves", [("fast", 21, 20), ("reference", 0, 0)])
def test_one_loop_step_at_4096(cuda, dam_break_4096_after_30, variant, rows, waves):
    """One step of the device loop on the full 4096^2 grid (the product path:
    the packed fast loop at the benchmark's tile; and the reference-order
    loop) from a developed dam-break state, against one oracle step with the
    same dt.  Wet mask bit-identical with the FP32 oracle; depth within
    2e-7 + 2e-5 |x| of the FP64 oracle; the reference-order depth is
    bit-identical with the FP32 oracle.  The momenta exceed that bar by ~1e-7
    at the dam break's pressure fluxes (g/2 h^2 ~ 1.2 for h = 0.5 m), and so
    does the FP32 oracle itself by the same amount: the bar there is the FP32
    oracle's own error against FP64 (max and RMSE)."""
    import ctypes

    import torch

    from hybrid_flood_amd import _lib
    from hybrid_flood_amd import numerics as nm
    from oracle import swe_oracle as o

    c, params, s30, t30 = dam_break_4096_after_30
    ctx = nm.context_for(params, variant)
    ctx.set_tile_rows(rows)
    ctx.set_waves_per_cu(waves)
    lib = _lib.load()
    log = torch.zeros(4, dtype=torch.float64, device=cuda)
    work = nm.SWEState(*(x.clone() for x in s30))
    s = _lib.stream_handle(cuda)
    try:
        _lib.check(lib.swe_set_dt_log(ctx.handle, log.data_ptr(), 4), "swe_set_dt_log")
        st = _lib.state_struct(*work)
        _lib.check(lib.swe_loop_begin(ctx.handle, st, t30, 1, s), "swe_loop_begin")
        _lib.check(lib.swe_loop_advance(ctx.handle, 1, s), "swe_loop_advance")
        info = _lib.SweRunInfo()
        _lib.check(lib.swe_loop_end(ctx.handle, st, ctypes.byref(info), s), "swe_loop_end")
    finally:
        lib.swe_set_dt_log(ctx.handle, None, 0)
        ctx.set_tile_rows(0)
        ctx.set_waves_per_cu(0)
    assert info.steps_used == 1
    dt = float(log[0])
    got = [x.cpu().numpy() for x in work]
    start = o.OracleState(*(x.cpu().numpy() for x in s30))
    kw = dict(bed=c.bed, manning_n=c.manning_n, domain_mask=c.domain_mask,
              boundary_types=c.boundary_types, rainfall_times_s=c.rain_times_s,
              rainfall_rates_m_s=c.rain_rates_m_s, dx=c.dx, cfl=c.cfl,
              dry_tolerance=c.dry_tolerance, max_dt=c.max_dt)
    print(f"oracle steps at 4096^2, dt {dt}", flush=True)
    r32 = o.finite_volume_step(start, o.OracleParams(**kw), dt, t30)
    wet_got = c.domain_mask & (got[0] > np.float32(c.dry_tolerance))
    wet_ref = c.domain_mask & (r32.h > np.float32(c.dry_tolerance))
    assert int(np.count_nonzero(wet_got != wet_ref)) == 0
    if variant == "reference":
        assert np.array_equal(got[0], r32.h)
    r64 = o.finite_volume_step(start, o.OracleParams(**kw, dtype=np.float64), dt, t30)
    for k, f in enumerate(("h", "hu", "hv")):
        want = getattr(r64, f)
        err = np.abs(got[k].astype(np.float64) - want)
        err32 = np.abs(getattr(r32, f).astype(np.float64) - want)
        over = float((err - (2e-7 + 2e-5 * np.abs(want))).max())
        over32 = float((err32 - (2e-7 + 2e-5 * np.abs(want))).max())
        rmse, rmse32 = float(np.sqrt(np.mean(err ** 2))), float(np.sqrt(np.mean(err32 ** 2)))
        print(f"{variant} {f}: device max {err.max():.3e} rmse {rmse:.3e} over-bar {over:.3e}; "
              f"FP32 oracle max {err32.max():.3e} rmse {rmse32:.3e} over-bar {over32:.3e}",
              flush=True)
        if f == "h":
            assert over <= 0.0, f
        # the device's error is the FP32 oracle's own (measured: max 7.19e-7 and
        # RMSE 1.78e-8 on hu for both)
        assert float(err.max()) <= 1.05 * float(err32.max()) + 1e-9, f
        assert rmse <= 1.1 * rmse32 + 1e-12, f


# Max |GPU - C oracle| after the 1339 steps of the first 3000 s, measured on
# MI355X: reference variant h 1.6e-7 m, hu 2.3e-8, hv 1.9e-8 m^2/s; fast
# variant 2.6e-7, 6.4e-8, 3.4e-8; no wet-mask mismatch.  The bars leave ~8x.
URBAN_WINDOW_TOL = {"h": 2e-6, "hu": 5e-7, "hv": 5e-7}


@pytest.fixture(scope="module")
def urban_first_window_oracle():
    """The 1024^2 urban configuration's first 3000 s (50 outputs, ~1300 steps:
    the rain wets two thirds of the cells and the water starts to run), by the
    C/OpenMP transcription of the FP32 oracle on the host's threads (seconds;
    the NumPy oracle would take minutes)."""
    import subprocess
    from pathlib import Path

    from hybrid_flood_amd import synthetic as sy
    from oracle import swe_oracle as o

    subprocess.run(["make", "-s", "-C", str(Path(__file__).resolve().parents[1] / "oracle")],
                   check=True)
    from oracle import swe_step_omp as comp

    u = sy.urban_case(1024)
    op = o.OracleParams(bed=u.bed, manning_n=u.manning_n, domain_mask=u.domain_mask,
                        boundary_types=u.boundary_types, rainfall_times_s=u.rain_times_s,
                        rainfall_rates_m_s=u.rain_rates_m_s, dx=u.dx, cfl=u.cfl,
                        dry_tolerance=u.dry_tolerance, max_dt=u.max_dt)
    times = np.arange(0.0, 3001.0, 60.0).astype(np.float32)
    z = np.zeros_like(u.h0)
    print(f"urban 1024 first 3000 s: C oracle on {comp.threads()} threads", flush=True)
    last, steps, t, written = comp.integrate_to_outputs(o.OracleState(u.h0, z, z), op, times,
                                                        max_steps=200000)
    return u, times, last, steps, t, written


@pytest.mark.parametrize("variant", ["reference", "fast"])
def test_urban_1024_first_window_against_c_oracle(cuda, urban_first_window_oracle, variant):
    from hybrid_flood_amd import numerics as nm

    u, times, last, steps, t, written = urban_first_window_oracle
    r = nm.integrate_to_outputs(_zero_state(u, cuda), _params(u, cuda), times, max_steps=200000,
                                variant=variant)
    got = {f: getattr(r.states, f)[-1].cpu().numpy() for f in ("h", "hu", "hv")}
    want = {"h": last.h, "hu": last.hu, "hv": last.hv}
    diffs = {f: float(np.abs(got[f] - want[f]).max()) for f in got}
    wet = (got["h"] > u.dry_tolerance) != (want["h"] > u.dry_tolerance)
    print(f"{variant}: steps {int(r.steps_used)} vs {steps}, max diffs {diffs}, "
          f"wet-mask mismatches {int(wet.sum())}", flush=True)
    assert int(r.outputs_written) == written == 51
    assert int(r.steps_used) == steps
    assert float(r.final_time_s) == t
    assert int(wet.sum()) == 0
    for f, bar in URBAN_WINDOW_TOL.items():
        assert diffs[f] <= bar, f


# After 100 steps (7.1 s), measured on MI355X: simulated time within 6.2e-9
# (fast) and 4.8e-10 (reference) relative of the oracle's; max |GPU - oracle|
# fast h 1.06e-6 m, hu 1.64e-6, hv 1.19e-6 m^2/s; reference 7.7e-7, 1.25e-6,
# 1.31e-6; no wet-mask mismatch.  The bars leave ~10x.
DAM_BREAK_100_TOL = {"h": 1e-5, "hu": 1.5e-5, "hv": 1.5e-5}


@pytest.fixture(scope="module")
def dam_break_4096_c_oracle_100():
    """The 4096^2 roofline workload's first 100 adaptive steps (CFL dt, then
    the step) by the C/OpenMP transcription of the FP32 oracle on the host's
    threads."""
    import subprocess
    from pathlib import Path

    from hybrid_flood_amd import synthetic as sy
    from oracle import swe_oracle as o

    subprocess.run(["make", "-s", "-C", str(Path(__file__).resolve().parents[1] / "oracle")],
                   check=True)
    from oracle import swe_step_omp as comp

    c = sy.dam_break_urban_case(4096)
    op = o.OracleParams(bed=c.bed, manning_n=c.manning_n, domain_mask=c.domain_mask,
                        boundary_types=c.boundary_types, rainfall_times_s=c.rain_times_s,
                        rainfall_rates_m_s=c.rain_rates_m_s, dx=c.dx, cfl=c.cfl,
                        dry_tolerance=c.dry_tolerance, max_dt=c.max_dt)
    z = np.zeros_like(c.h0)
    print(f"4096^2: 100 C oracle steps on {comp.threads()} threads", flush=True)
    last, t = comp.run_steps(o.OracleState(c.h0, z, z), op, 100)
    return c, last, t


@pytest.mark.parametrize("variant,rows", [("fast", 15), ("reference", 0)])
def test_loop_100_steps_at_4096_against_c_oracle(cuda, dam_break_4096_c_oracle_100, variant,
                                                 rows):
    """100 adaptive steps of the device loop on the full 4096^2 grid (the
    packed fast loop at the benchmark's 15-row tile, and the reference-order
    loop) against the oracle's 100 steps, each side taking its own CFL dt."""
    from hybrid_flood_amd import numerics as nm

    c, last, t = dam_break_4096_c_oracle_100
    params = _params(c, cuda)
    ctx = nm.context_for(params, variant)
    ctx.set_tile_rows(rows)
    try:
        st = nm.SWEState(*(x.clone() for x in _zero_state(c, cuda)))
        info = nm.run_steps(st, params, 100, variant=variant, diagnostics=True)
    finally:
        ctx.set_tile_rows(0)
    got = {f: getattr(st, f).cpu().numpy() for f in ("h", "hu", "hv")}
    diffs = {f: float(np.abs(got[f] - getattr(last, f)).max()) for f in got}
    wet = (got["h"] > np.float32(c.dry_tolerance)) != (last.h > np.float32(c.dry_tolerance))
    dt_rel = abs(float(info["final_time_s"]) - t) / t
    print(f"{variant}: time {info['final_time_s']!r} vs {t!r} (rel {dt_rel:.2e}), max diffs "
          f"{diffs}, wet-mask mismatches {int(wet.sum())}", flush=True)
    assert info["steps_used"] == 100 and info["nonfinite"] == 0
    assert dt_rel <= 1e-7
    assert int(wet.sum()) == 0
    for f, bar in DAM_BREAK_100_TOL.items():
        assert diffs[f] <= bar, f
