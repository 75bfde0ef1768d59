"""The C/OpenMP restatement of the step (``oracle/swe_step_omp.c``) against
the NumPy oracle, a second, independent transcription of the same
arithmetic, and against the V1 golden vectors.

Depth and the CFL step size agree bit for bit on every case.  Momentum agrees
bit for bit where the friction term's power is the same function: with no
friction, or with the NumPy oracle's ``np.power`` replaced by libm's
``powf`` (the C code's); with NumPy's own float32 power (SIMD code on AVX-512
hosts, not libm) it agrees to a few ulp.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import random_case

ROOT = Path(__file__).resolve().parents[1]
SIDES = [(0, 0, 0, 0), (1, 1, 1, 1), (0, 1, 0, 1), (1, 0, 1, 0)]
SHAPES = [(7, 9), (16, 16), (33, 24), (50, 130)]


@pytest.fixture(scope="module")
def c():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    from oracle import swe_step_omp

    return swe_step_omp


def _setup(seed, shape, sides, **kw):
    from oracle import swe_oracle as o

    case = random_case(seed, shape, open_sides=sides, speed=1.0)
    case.update(kw)
    p = o.OracleParams(bed=case["bed"], manning_n=case["manning_n"],
                       domain_mask=case["domain_mask"],
                       rainfall_multiplier=case["rainfall_multiplier"],
                       boundary_types=case["boundary_types"], rainfall_times_s=case["rain_t"],
                       rainfall_rates_m_s=case["rain_r"], dx=case["dx"],
                       dry_tolerance=case["dry"])
    return p, o.OracleState(case["h"], case["hu"], case["hv"])


def _ulps(a, b):
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return np.abs(ia - ib)


@pytest.mark.parametrize("k", range(8))
def test_depth_and_cfl_bit_exact(c, k):
    from oracle import swe_oracle as o

    p, st = _setup(200 + k, SHAPES[k % 4], SIDES[k % 4])
    dt = o.compute_cfl_dt(st, p)
    assert c.compute_cfl_dt(st, p) == dt
    a = o.finite_volume_step(st, p, dt, 0.3)
    b = c.finite_volume_step(st, p, dt, 0.3)
    assert np.array_equal(a.h, b.h)
    for x, y in ((a.hu, b.hu), (a.hv, b.hv)):
        assert int(_ulps(x, y).max()) <= 4


@pytest.mark.parametrize("k", range(4))
def test_bit_exact_without_friction(c, k):
    from oracle import swe_oracle as o

    p, st = _setup(210 + k, SHAPES[k], SIDES[k], manning_n=np.zeros(SHAPES[k], np.float32))
    a, ta, _ = o.run_steps(st, p, 12)
    b, tb = c.run_steps(st, p, 12)
    assert tb == ta
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_bit_exact_with_libm_power(c, monkeypatch):
    """Swap the NumPy oracle's float32 power for libm's powf: every field then
    agrees bit for bit, so the two transcriptions differ in nothing else."""
    from oracle import swe_oracle as o

    libm = ctypes.CDLL("libm.so.6")
    libm.powf.restype = ctypes.c_float
    libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]
    real_power = np.power

    def powf(x, y):
        x = np.asarray(x)
        if x.dtype != np.float32:
            return real_power(x, y)
        flat = [libm.powf(float(v), float(y)) for v in x.ravel()]
        return np.array(flat, np.float32).reshape(x.shape)

    monkeypatch.setattr(np, "power", powf)
    for k in range(4):
        p, st = _setup(220 + k, SHAPES[k], SIDES[k])
        dt = o.compute_cfl_dt(st, p)
        a = o.finite_volume_step(st, p, dt, 0.3)
        b = c.finite_volume_step(st, p, dt, 0.3)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_thread_count_does_not_change_results(c):
    p, st = _setup(230, (97, 61), (0, 1, 0, 1))
    n0 = c.threads()
    outs = []
    try:
        for n in (1, 3, max(n0, 2)):
            c.set_threads(n)
            outs.append(c.run_steps(st, p, 6))
    finally:
        c.set_threads(n0)
    for s, t in outs[1:]:
        assert t == outs[0][1]
        for x, y in zip(s, outs[0][0]):
            assert np.array_equal(x, y)


def test_rejects_bad_arguments(c):
    lib = c.load()
    z = ctypes.POINTER(ctypes.c_float)()
    assert lib.swe_omp_step(0, 4, z, z, z, z, z, None, z, None, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                            z, z, z) == -1


def test_output_loop_matches_oracle_without_friction(c):
    from oracle import swe_oracle as o

    p, st = _setup(240, (40, 56), (0, 1, 0, 0), manning_n=np.zeros((40, 56), np.float32))
    times = np.array([0.0, 0.5, 1.25, 2.0])
    ref = o.integrate_to_outputs(st, p, times, max_steps=2000)
    last, steps, t, written = c.integrate_to_outputs(st, p, times, max_steps=2000)
    assert (steps, t, written) == (ref.steps_used, ref.final_time_s, ref.outputs_written)
    for x, y in zip(last, ref.states):
        assert np.array_equal(x, y[-1])


GOLDEN_STEPS = ["v1_step_parity_7x9.npz", "v1_step_wetdry_16x12.npz", "v1_step_walls_20x24.npz"]


@pytest.mark.parametrize("name", GOLDEN_STEPS)
def test_c_step_against_v1_golden_vectors(c, name):
    """Straight against the V1 golden vectors (written from the reference's
    NumPy twin): the CFL dt and the depth bit for bit; momentum, where V1 takes
    np.hypot and V2's (and this) sqrt(hu^2 + hv^2), within the reference's own
    V1/V2 bar (tests/test_v1_v2_architecture.py:104)."""
    from conftest import load_golden, oracle_params_from_golden
    from oracle import swe_oracle as o

    z = load_golden(name)
    p = oracle_params_from_golden(z)
    st = o.OracleState(z["h"], z["hu"], z["hv"])
    assert c.compute_cfl_dt(st, p) == float(z["cfl_dt"])
    out = c.finite_volume_step(st, p, float(z["dt"]), float(z["t"]))
    assert np.array_equal(out.h, z["out_h"])
    np.testing.assert_allclose(out.hu, z["out_hu"], rtol=2e-5, atol=2e-7)
    np.testing.assert_allclose(out.hv, z["out_hv"], rtol=2e-5, atol=2e-7)
