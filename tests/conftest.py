"""Shared test setup: import paths, the ``gpu`` marker and small case builders."""

from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "gpu-hybrid-flood-forecasting_amd")
for p in (ROOT, PKG_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:  # pragma: no cover
        return False


@pytest.fixture(scope="session")
def cuda():
    if not gpu_available():
        pytest.skip("no GPU visible")
    import torch

    return torch.device("cuda", 0)


def load_golden(name):
    return np.load(os.path.join(GOLDEN_DIR, name))


def oracle_params_from_golden(z, dtype=np.float32):
    from oracle import swe_oracle as o

    return o.OracleParams(
        bed=z["bed"], manning_n=z["manning_n"], domain_mask=z["domain_mask"],
        rainfall_times_s=z["rain_t"], rainfall_rates_m_s=z["rain_r"], dx=float(z["dx"]),
        cfl=float(z["cfl"]), dry_tolerance=float(z["dry"]), max_dt=float(z["max_dt"]),
        dtype=dtype,
    )


def random_case(seed, shape, *, open_sides=(0, 0, 0, 0), mask_frac=0.1, dry_frac=0.15,
                speed=1.0, manning=(0.01, 0.05), rain=2.0e-5, dx=2.0, dry=1.0e-4):
    """A seeded wet/dry case with mask holes, rough bed, friction and rain."""
    rng = np.random.default_rng(seed)
    ny, nx = shape
    y, x = np.mgrid[0:ny, 0:nx].astype(np.float64)
    bed = 0.002 * x * dx
    for _ in range(max(2, (ny * nx) // 400)):
        cy, cx = rng.uniform(0, ny), rng.uniform(0, nx)
        s = rng.uniform(1.5, 6.0)
        bed = bed + rng.uniform(-0.2, 0.4) * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
    bed = bed.astype(np.float32)
    mask = rng.uniform(size=shape) > mask_frac
    eta = 0.25 + rng.normal(0, 0.03, shape)
    h = np.maximum(eta - bed, 0.0)
    h[rng.uniform(size=shape) < dry_frac] = 0.0
    h[rng.uniform(size=shape) < 0.05] = 0.5 * dry
    h = np.where(mask, h, 0.0).astype(np.float32)
    hu = (h * rng.normal(0, speed, shape)).astype(np.float32)
    hv = (h * rng.normal(0, speed, shape)).astype(np.float32)
    hu = np.where(mask, hu, 0).astype(np.float32)
    hv = np.where(mask, hv, 0).astype(np.float32)
    n = rng.uniform(manning[0], manning[1], shape).astype(np.float32)
    mult = np.where(mask, rng.uniform(0.5, 1.5, shape), 0).astype(np.float32)
    return dict(
        bed=bed, manning_n=n, domain_mask=mask, rainfall_multiplier=mult,
        boundary_types=tuple(open_sides), rain_t=np.array([0.0, 50.0, 200.0], np.float32),
        rain_r=np.array([0.0, rain, rain * 0.25], np.float32), dx=dx, dry=dry,
        h=h, hu=hu, hv=hv,
    )
