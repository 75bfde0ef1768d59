"""ctypes binding of ``oracle/swe_step_omp.c``: the C/OpenMP restatement of
the float32 step, on all host cores.

TEST INFRASTRUCTURE ONLY.  Only ``tests/`` and the ``cpu_baseline`` leg of
``bench.py`` use it: as a second restatement checked against
``oracle/swe_oracle.py`` (depth bit for bit; momentum to an ulp, where
NumPy's float32 ``power`` and glibc's ``powf`` may round the friction term
differently) and as the multi-core CPU baseline.

``finite_volume_step`` and ``compute_cfl_dt`` take the NumPy oracle's
``OracleState`` / ``OracleParams`` (float32) and round the float64 scalars
(``dt``, ``dt * rate``, ``dt * g``) to float32 exactly as that oracle does
(``swe_oracle.py``: ``finite_volume_step``, ``compute_cfl_dt``).
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from .swe_oracle import OracleParams, OracleState, rainfall_rate

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libswe_step_omp.so")

_f = ctypes.c_float
_fp = ctypes.POINTER(ctypes.c_float)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_lib = None


def build() -> None:
    """Compile the library with ``oracle/Makefile`` (gcc, OpenMP)."""
    import subprocess

    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `make -C oracle`")
        lib = ctypes.CDLL(LIB_PATH)
        lib.swe_omp_step.restype = ctypes.c_int
        lib.swe_omp_step.argtypes = [ctypes.c_int, ctypes.c_int, _fp, _fp, _fp, _fp, _fp, _u8p, _fp,
                                     _i32p, _f, _f, _f, _f, _f, _f, _f, _f, _f, _fp, _fp, _fp]
        lib.swe_omp_cfl_max.restype = ctypes.c_float
        lib.swe_omp_cfl_max.argtypes = [ctypes.c_int, ctypes.c_int, _fp, _fp, _fp, _u8p, _f, _f,
                                        _f, _f]
        lib.swe_omp_threads.restype = ctypes.c_int
        lib.swe_omp_threads.argtypes = []
        lib.swe_omp_set_threads.restype = None
        lib.swe_omp_set_threads.argtypes = [ctypes.c_int]
        _lib = lib
    return _lib


def threads() -> int:
    """OpenMP threads a call uses (OMP_NUM_THREADS, else the host's cores)."""
    return int(load().swe_omp_threads())


def set_threads(n: int) -> None:
    """Use ``n`` OpenMP threads from now on in this process."""
    load().swe_omp_set_threads(int(n))


class Fields:
    """The static fields of ``params`` as contiguous C arrays, made once."""

    def __init__(self, params: OracleParams):
        if params.dtype is not np.float32:
            raise ValueError("the C restatement is float32 only")
        self.params = params
        self.bed = np.ascontiguousarray(params.bed, np.float32)
        self.n = np.ascontiguousarray(params.manning_n, np.float32)
        self.mult = np.ascontiguousarray(params.rainfall_multiplier, np.float32)
        self.mask = np.ascontiguousarray(params.domain_mask, np.uint8)
        self.bt = np.ascontiguousarray(params.boundary_types, np.int32)


def _p(a, t=_fp):
    return a.ctypes.data_as(t)


def finite_volume_step(state: OracleState, params: OracleParams, dt: float, time_s: float,
                       fields: Fields | None = None, out: OracleState | None = None) -> OracleState:
    """One step, as ``swe_oracle.finite_volume_step`` (float32, ``norm="sqrt"``)."""
    T = np.float32
    f = fields or Fields(params)
    h, hu, hv = (np.ascontiguousarray(x, T) for x in state)
    ny, nx = h.shape
    if out is None:
        out = OracleState(np.empty_like(h), np.empty_like(h), np.empty_like(h))
    rate = rainfall_rate(params, time_s)
    rc = load().swe_omp_step(
        ny, nx, _p(h), _p(hu), _p(hv), _p(f.bed), _p(f.n), _p(f.mask, _u8p), _p(f.mult),
        _p(f.bt, _i32p), T(params.gravity), T(0.5 * params.gravity), T(params.dry_tolerance),
        T(params.dx), T(params.dy), T(dt), T(dt * rate), T(dt * params.gravity), T(7.0 / 3.0),
        _p(out.h), _p(out.hu), _p(out.hv))
    if rc != 0:
        raise RuntimeError(f"swe_omp_step failed ({rc})")
    return out


def compute_cfl_dt(state: OracleState, params: OracleParams, fields: Fields | None = None) -> float:
    """As ``swe_oracle.compute_cfl_dt``: the float32 maximum, the quotient in float64."""
    T = np.float32
    f = fields or Fields(params)
    h, hu, hv = (np.ascontiguousarray(x, T) for x in state)
    ny, nx = h.shape
    m = float(load().swe_omp_cfl_max(ny, nx, _p(h), _p(hu), _p(hv), _p(f.mask, _u8p),
                                      T(params.gravity), T(params.dry_tolerance), T(params.dx),
                                      T(params.dy)))
    return min(params.cfl / m, params.max_dt) if m > 0.0 else float(params.max_dt)


def run_steps(state: OracleState, params: OracleParams, n_steps: int, t0: float = 0.0):
    """``n_steps`` adaptive steps (CFL dt, then the step); returns (state, t)."""
    f = Fields(params)
    a = OracleState(*(np.array(x, np.float32) for x in state))
    b = OracleState(*(np.empty_like(x) for x in a))
    t = float(t0)
    for _ in range(n_steps):
        dt = compute_cfl_dt(a, params, f)
        finite_volume_step(a, params, dt, t, f, b)
        a, b = b, a
        t += dt
    return a, t


def integrate_to_outputs(initial: OracleState, params: OracleParams, output_times_s, *,
                         max_steps: int):
    """The output loop of ``swe_oracle.integrate_to_outputs`` (float64 clock,
    dt = min(CFL dt, time to the next output)) over the C step; returns
    (final state, steps taken, final time, outputs written)."""
    from .swe_oracle import OUTPUT_TIME_SLACK_S

    times = np.asarray(output_times_s, np.float64)
    f = Fields(params)
    a = OracleState(*(np.array(x, np.float32) for x in initial))
    b = OracleState(*(np.empty_like(x) for x in a))
    t = float(times[0])
    steps = 0
    written = 1
    for target in times[1:]:
        while t < target - OUTPUT_TIME_SLACK_S and steps < max_steps:
            dt = min(compute_cfl_dt(a, params, f), float(target) - t)
            finite_volume_step(a, params, dt, t, f, b)
            a, b = b, a
            t = t + dt
            steps += 1
        written += int(t >= target - OUTPUT_TIME_SLACK_S)
    return a, steps, t, written
