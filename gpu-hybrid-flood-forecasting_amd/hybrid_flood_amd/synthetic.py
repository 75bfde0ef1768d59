"""Seeded synthetic inputs of the benchmark and parity cases.

The reference builds its DEM from shapefiles and Perlin noise
(dem/synthetic_terrain.py:114-217); the cases here use their own seeded
generators, one ``numpy.random.default_rng(seed)`` per call:

  sloped_bumps_case   256 x 256, dx 2 m: 1 % plane + 8 Gaussian bumps + a
                      4 m x 0.5 m channel; n = 0.03; trapezoidal rain
                      peaking at 50 mm/h; reflective; dry start.
  urban_case          n x n urban terrain: 0.5 % slope + fractal noise
                      (std 0.3 m), ~1500 buildings (per 1024^2 at 2 m) raised
                      10 m with U[8, 40] m footprints between 12 m streets;
                      n = 0.015 on streets, 0.035 elsewhere; one open edge.
  rain_series_1min    seeded piecewise-linear one-minute series whose
                      trapezoid total is 8.35497 mm over 10800 s (the
                      reference's development window, README.md:77-79).
  dam_break_state     wet start h = 0.05 m + a 0.5 m square in the centre.
  ensemble_members    uneven member sizes from {256, 512, 1024, 2048}.
  smooth_field        smooth seeded field mapped into [lo, hi].
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MM_HR_TO_M_S = 1.0e-3 / 3600.0  # jax_solver/shallow_water_2d.py:32
REFERENCE_WINDOW_MM = 8.35497  # README.md:77-79


@dataclass
class Case:
    name: str
    bed: np.ndarray
    manning_n: np.ndarray
    domain_mask: np.ndarray
    dx: float
    boundary_types: tuple
    rain_times_s: np.ndarray
    rain_rates_m_s: np.ndarray
    h0: np.ndarray
    cfl: float = 0.45
    dry_tolerance: float = 1.0e-4
    max_dt: float = 10.0
    duration_s: float = 0.0
    output_interval_s: float = 0.0
    n_steps: int = 0
    seed: int = 0

    @property
    def shape(self):
        return self.bed.shape

    def output_times(self) -> np.ndarray:
        n = int(round(self.duration_s / self.output_interval_s))
        return (np.arange(n + 1) * self.output_interval_s).astype(np.float32)


def _cell_centres(n: int, dx: float):
    c = (np.arange(n, dtype=np.float64) + 0.5) * dx
    return np.meshgrid(c, c)  # X varies along columns (x), Y along rows (y)


def fractal_noise(n: int, rng: np.random.Generator, *, beta: float = 3.0, std: float = 0.3) -> np.ndarray:
    """Spectral-synthesis fractal surface with power spectrum ~ k^-beta, scaled to ``std``."""
    white = rng.standard_normal((n, n))
    spec = np.fft.rfft2(white)
    ky = np.fft.fftfreq(n)[:, None]
    kx = np.fft.rfftfreq(n)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    spec *= k ** (-beta / 2.0)
    spec[0, 0] = 0.0
    surf = np.fft.irfft2(spec, s=(n, n))
    surf -= surf.mean()
    return surf * (std / surf.std())


def smooth_field(n: int, rng: np.random.Generator, lo: float, hi: float, *, beta: float = 4.0) -> np.ndarray:
    f = fractal_noise(n, rng, beta=beta, std=1.0)
    f = (f - f.min()) / max(f.max() - f.min(), 1e-12)
    return (lo + (hi - lo) * f).astype(np.float32)


def trapezoid_total_mm(times_s: np.ndarray, rates_mm_h: np.ndarray) -> float:
    t = np.asarray(times_s, np.float64)
    r = np.asarray(rates_mm_h, np.float64)
    return float(np.sum(0.5 * (r[1:] + r[:-1]) * np.diff(t)) / 3600.0)


def rain_series_1min(rng: np.random.Generator, *, duration_s: float = 10800.0,
                     total_mm: float = REFERENCE_WINDOW_MM):
    """Piecewise-linear one-minute rain (m/s) whose trapezoid total is ``total_mm``."""
    times = np.arange(0.0, duration_s + 30.0, 60.0)
    n = times.size
    # A seeded storm: gamma-shaped envelope times lognormal minute noise.
    peak = rng.uniform(0.3, 0.6) * duration_s
    shape_k = rng.uniform(2.0, 4.0)
    x = np.maximum(times / peak, 1e-6)
    env = x ** shape_k * np.exp(shape_k * (1.0 - x))
    noise = rng.lognormal(0.0, 0.5, n)
    rates_mm_h = env * noise
    rates_mm_h *= total_mm / trapezoid_total_mm(times, rates_mm_h)
    return times.astype(np.float32), (rates_mm_h * MM_HR_TO_M_S).astype(np.float32)


def trapezoid_hyetograph(duration_s: float, peak_mm_h: float):
    t = np.array([0.0, 0.2, 0.8, 1.0]) * duration_s
    r = np.array([0.0, peak_mm_h, peak_mm_h, 0.0])
    return t.astype(np.float32), (r * MM_HR_TO_M_S).astype(np.float32)


def sloped_bumps_case(n: int = 256, dx: float = 2.0, seed: int = 42) -> Case:
    rng = np.random.default_rng(seed)
    X, Y = _cell_centres(n, dx)
    L = n * dx
    bed = 0.01 * (L - X)
    for _ in range(8):
        amp = rng.uniform(0.2, 1.5)
        sig = rng.uniform(5.0, 20.0)
        cx, cy = rng.uniform(0.0, L, 2)
        bed += amp * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2.0 * sig * sig))
    bed[np.abs(Y - 0.5 * L) < 2.0] -= 0.5  # 4 m wide, 0.5 m deep channel along x
    rt, rr = trapezoid_hyetograph(600.0, 50.0)
    return Case(
        name=f"sloped_bumps_{n}", bed=bed.astype(np.float32),
        manning_n=np.full((n, n), 0.03, np.float32), domain_mask=np.ones((n, n), bool), dx=dx,
        boundary_types=(0, 0, 0, 0), rain_times_s=rt, rain_rates_m_s=rr,
        h0=np.zeros((n, n), np.float32), cfl=0.9, dry_tolerance=1.0e-3, max_dt=10.0,
        duration_s=600.0, output_interval_s=60.0, seed=seed,
    )


def urban_terrain(n: int, dx: float, rng: np.random.Generator):
    """Urban DEM, Manning field and street mask on an n x n grid."""
    X, Y = _cell_centres(n, dx)
    L = n * dx
    bed = 0.005 * (L - X) + fractal_noise(n, rng, std=0.3)
    pitch, street = 52.0, 12.0
    streets = ((X % pitch) < street) | ((Y % pitch) < street)
    land = pitch - street
    nb = int(L // pitch)
    buildings = np.zeros((n, n), bool)
    for by in range(nb):
        for bx in range(nb):
            w, h = rng.uniform(8.0, 40.0, 2)
            ox, oy = rng.uniform(0.0, land - w), rng.uniform(0.0, land - h)
            x0 = bx * pitch + street + ox
            y0 = by * pitch + street + oy
            i0, i1 = int(np.floor(x0 / dx)), int(np.ceil((x0 + w) / dx))
            j0, j1 = int(np.floor(y0 / dx)), int(np.ceil((y0 + h) / dx))
            buildings[max(j0, 0):min(j1, n), max(i0, 0):min(i1, n)] = True
    buildings &= ~streets
    bed = bed + 10.0 * buildings
    manning = np.where(streets, 0.015, 0.035).astype(np.float32)
    return bed.astype(np.float32), manning, streets, int(nb * nb)


def urban_case(n: int = 1024, dx: float = 2.0, seed: int = 42, *, rain_scale: float = 1.0,
               n_scale: float = 1.0, duration_s: float = 10800.0) -> Case:
    rng = np.random.default_rng(seed)
    bed, manning, _, _ = urban_terrain(n, dx, rng)
    rt, rr = rain_series_1min(rng, duration_s=10800.0)
    return Case(
        name=f"urban_{n}", bed=bed, manning_n=(manning * np.float32(n_scale)).astype(np.float32),
        domain_mask=np.ones((n, n), bool), dx=dx, boundary_types=(0, 1, 0, 0),
        rain_times_s=rt, rain_rates_m_s=(rr * np.float32(rain_scale)).astype(np.float32),
        h0=np.zeros((n, n), np.float32), cfl=0.45, dry_tolerance=1.0e-4, max_dt=10.0,
        duration_s=duration_s, output_interval_s=60.0, seed=seed,
    )


def dam_break_state(n: int, *, base: float = 0.05, top: float = 0.5, frac: float = 0.125) -> np.ndarray:
    h = np.full((n, n), base, np.float32)
    half = max(1, int(round(n * frac / 2)))
    c = n // 2
    h[c - half:c + half, c - half:c + half] = top
    return h


def dam_break_urban_case(n: int = 4096, dx: float = 1.0, seed: int = 42, n_steps: int = 2000) -> Case:
    """Wet-start urban domain with a central 0.5 m dam-break square; no rain, cfl 0.45."""
    rng = np.random.default_rng(seed)
    bed, manning, _, _ = urban_terrain(n, dx, rng)
    return Case(
        name=f"dam_break_urban_{n}", bed=bed, manning_n=manning, domain_mask=np.ones((n, n), bool),
        dx=dx, boundary_types=(0, 1, 0, 0), rain_times_s=np.zeros(1, np.float32),
        rain_rates_m_s=np.zeros(1, np.float32), h0=dam_break_state(n), cfl=0.45,
        dry_tolerance=1.0e-4, max_dt=10.0, n_steps=n_steps, seed=seed,
    )


def ensemble_members(count: int = 64, seed: int = 42, sides=(256, 512, 1024, 2048)):
    """(side, seed, n_scale, rain_scale) for every member."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        side = int(rng.choice(sides))
        out.append((side, int(rng.integers(0, 2**31 - 1)), float(rng.uniform(0.7, 1.3)),
                    float(rng.uniform(0.5, 2.0))))
    return out
