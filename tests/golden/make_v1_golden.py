"""Write V1 (NumPy twin) golden vectors for the FP32 step, CFL rule and loop.

Run in the build container only, where the reference checkout exists:

    python tests/golden/make_v1_golden.py /root/reference

It imports ``hybrid_flood.numpy_solver.shallow_water_2d`` (numpy only) from the
reference checkout and records inputs and outputs of ``finite_volume_step``
(shallow_water_2d.py:277-327), ``compute_cfl_dt`` (:267-274) and
``NumpyShallowWaterSolver.run`` (:382-424) on small seeded cases.  Those pin
the FP32 evaluation order of the oracle, which no closed form fixes.  The
fixtures are data only (``v1_*.npz``); nothing from the reference is kept.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent


def _import_v1(reference_root: Path):
    sys.dont_write_bytecode = True
    sys.path.insert(0, str(reference_root / "src"))
    from hybrid_flood.numpy_solver import shallow_water_2d as v1  # noqa: PLC0415

    return v1


def _bumpy_bed(rng, shape, amp):
    ny, nx = shape
    y, x = np.mgrid[0:ny, 0:nx].astype(np.float64)
    bed = np.zeros(shape)
    for _ in range(4):
        cy, cx = rng.uniform(0, ny), rng.uniform(0, nx)
        s = rng.uniform(1.5, 4.0)
        bed += rng.uniform(0.2, 1.0) * amp * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
    return bed.astype(np.float32)


def _case_step(v1, name, *, bed, manning, mask, h, hu, hv, dx, rain_t, rain_r, dt, t,
               cfl=0.45, dry=1.0e-4, max_dt=2.0):
    solver = v1.NumpyShallowWaterSolver(
        bed, manning, mask, resolution_m=dx, rainfall_times_s=rain_t,
        rainfall_rates_m_s=rain_r, cfl=cfl, dry_tolerance_m=dry, max_dt_s=max_dt,
    )
    state = v1.NumpyState(h.astype(np.float32), hu.astype(np.float32), hv.astype(np.float32))
    if dt is None:
        dt = v1.compute_cfl_dt(state, solver.params)
    cfl_dt = v1.compute_cfl_dt(state, solver.params)
    out = v1.finite_volume_step(state, solver.params, dt, t)
    np.savez_compressed(
        OUT / f"v1_step_{name}.npz",
        bed=bed, manning_n=manning, domain_mask=mask, h=state.h, hu=state.hu, hv=state.hv,
        dx=np.float64(dx), rain_t=rain_t, rain_r=rain_r, dt=np.float64(dt), t=np.float64(t),
        cfl=np.float64(cfl), dry=np.float64(dry), max_dt=np.float64(max_dt),
        cfl_dt=np.float64(cfl_dt), out_h=out.h, out_hu=out.hu, out_hv=out.hv,
    )


def main(reference_root: str) -> None:
    v1 = _import_v1(Path(reference_root))

    # 1. The reference's own V1/V2 parity case (tests/test_v1_v2_architecture.py:85-104).
    rng = np.random.default_rng(7)
    bed = rng.uniform(0.0, 0.2, size=(7, 9)).astype(np.float32)
    h = np.full(bed.shape, 0.1, np.float32)
    hu = rng.normal(0.0, 1.0e-3, bed.shape).astype(np.float32)
    hv = rng.normal(0.0, 1.0e-3, bed.shape).astype(np.float32)
    _case_step(
        v1, "parity_7x9", bed=bed, manning=np.full(bed.shape, 0.03, np.float32),
        mask=np.ones(bed.shape, bool), h=h, hu=hu, hv=hv, dx=20.0,
        rain_t=np.array([0.0, 1000.0], np.float32),
        rain_r=np.array([1.2e-6, 1.2e-6], np.float32), dt=0.75, t=20.0, max_dt=2.0,
    )

    # 2. Wet/dry mix with a mask hole, rough bed, rain, friction.
    rng = np.random.default_rng(11)
    shape = (16, 12)
    bed = _bumpy_bed(rng, shape, 0.3)
    mask = np.ones(shape, bool)
    mask[5:9, 4:6] = False
    mask[0, 7:] = False
    eta = 0.18 + rng.normal(0, 0.02, shape)
    h = np.maximum(eta - bed, 0.0).astype(np.float32)
    h[rng.uniform(size=shape) < 0.15] = 0.0
    h[rng.uniform(size=shape) < 0.1] = 5.0e-5  # below the dry tolerance
    h = np.where(mask, h, 0).astype(np.float32)
    hu = (h * rng.normal(0, 0.4, shape)).astype(np.float32)
    hv = (h * rng.normal(0, 0.4, shape)).astype(np.float32)
    _case_step(
        v1, "wetdry_16x12", bed=bed, manning=rng.uniform(0.01, 0.05, shape).astype(np.float32),
        mask=mask, h=h, hu=hu, hv=hv, dx=2.0,
        rain_t=np.array([0.0, 30.0, 90.0, 120.0], np.float32),
        rain_r=np.array([0.0, 2.0e-5, 1.0e-5, 0.0], np.float32), dt=None, t=47.5,
        cfl=0.9, dry=1.0e-4, max_dt=5.0,
    )

    # 3. Fast flows (super- and subcritical) at irregular walls: every HLL branch.
    rng = np.random.default_rng(23)
    shape = (20, 24)
    bed = _bumpy_bed(rng, shape, 0.1)
    mask = rng.uniform(size=shape) > 0.12
    mask[:, 0] = True
    h = np.where(mask, rng.uniform(0.01, 0.4, shape), 0).astype(np.float32)
    hu = (h * rng.normal(0, 3.0, shape)).astype(np.float32)
    hv = (h * rng.normal(0, 3.0, shape)).astype(np.float32)
    _case_step(
        v1, "walls_20x24", bed=bed, manning=np.full(shape, 0.035, np.float32), mask=mask,
        h=h, hu=hu, hv=hv, dx=1.0, rain_t=np.array([0.0, 60.0], np.float32),
        rain_r=np.array([5.0e-6, 5.0e-6], np.float32), dt=None, t=3.0,
        cfl=0.45, dry=1.0e-3, max_dt=10.0,
    )

    # 4. A short adaptive run: dam break next to a mask hole, with rain.
    rng = np.random.default_rng(5)
    shape = (12, 14)
    bed = (0.02 * np.arange(shape[1])[None, :] * np.ones(shape)).astype(np.float32)
    mask = np.ones(shape, bool)
    mask[3:6, 8:10] = False
    h = np.full(shape, 0.05, np.float32)
    h[4:8, 2:6] = 0.4
    h = np.where(mask, h, 0).astype(np.float32)
    manning = np.full(shape, 0.03, np.float32)
    rain_t = np.array([0.0, 1.0, 2.0], np.float32)
    rain_r = np.array([0.0, 1.0e-4, 0.0], np.float32)
    solver = v1.NumpyShallowWaterSolver(
        bed, manning, mask, resolution_m=2.0, rainfall_times_s=rain_t,
        rainfall_rates_m_s=rain_r, cfl=0.45, dry_tolerance_m=1.0e-4, max_dt_s=10.0,
    )
    times = np.array([0.0, 0.5, 1.0, 2.5], np.float32)
    init = solver.initial_state(depth_m=h)
    res = solver.run(init, times, max_steps=1000)
    np.savez_compressed(
        OUT / "v1_run_dambreak_12x14.npz",
        bed=bed, manning_n=manning, domain_mask=mask, h=init.h, hu=init.hu, hv=init.hv,
        dx=np.float64(2.0), rain_t=rain_t, rain_r=rain_r, cfl=np.float64(0.45),
        dry=np.float64(1.0e-4), max_dt=np.float64(10.0), times=times,
        out_h=res.states.h, out_hu=res.states.hu, out_hv=res.states.hv,
        final_time_s=np.float64(res.final_time_s), outputs_written=np.int64(res.outputs_written),
        steps_used=np.int64(res.steps_used),
    )
    print("wrote", sorted(p.name for p in OUT.glob("v1_*.npz")))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
