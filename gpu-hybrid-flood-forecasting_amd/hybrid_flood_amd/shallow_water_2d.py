"""Drop-in ``ShallowWaterSolver`` (jax_solver/shallow_water_2d.py:210-306).

Builds device-resident ``SWEParams`` once, masks the initial state the same
way, validates output times with the reference's messages and raises
``RuntimeError`` when ``max_steps`` was insufficient.  Geospatial loading
(rasterio/geopandas) is not part of this package: ``StructuredGrid`` takes the
rasterised arrays directly.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np
import torch

from . import numerics
from .boundary_conditions import BoundaryConfig, all_reflective
from .numerics import ScanResult, SWEState


@dataclass(frozen=True)
class StructuredGrid:
    """Rasterised inputs (jax_solver/shallow_water_2d.py:35-46)."""

    bed: np.ndarray
    manning_n: np.ndarray
    domain_mask: np.ndarray
    x: np.ndarray | None = None
    y: np.ndarray | None = None
    resolution_m: float = 1.0
    crs: Any = None
    north_up_transform: Any = None


class ShallowWaterSolver:
    """Convenience wrapper around the fused HIP step (shallow_water_2d.py:210-306)."""

    def __init__(self, grid: StructuredGrid, *, rainfall_times_s, rainfall_rates_m_s,
                 default_rainfall_m_s: float = 0.0, rainfall_multiplier=None,
                 boundaries: BoundaryConfig | None = None, gravity_m_s2: float = 9.81,
                 cfl: float = 0.45, dry_tolerance_m: float = 1.0e-4, max_dt_s: float = 10.0,
                 dtype: Any = torch.float32, variant: str | None = None, device=None) -> None:
        if dtype not in (torch.float32, np.float32, "float32"):
            raise ValueError("The HIP step computes in float32; dtype must be float32.")
        self.grid = grid
        mask = np.asarray(grid.domain_mask, dtype=bool)
        multiplier = (
            mask.astype(np.float32)
            if rainfall_multiplier is None
            else np.asarray(rainfall_multiplier, dtype=np.float32)
        )
        if multiplier.shape != np.asarray(grid.bed).shape:
            raise ValueError("Rainfall multiplier must match the structured-grid shape.")
        boundary_config = boundaries or all_reflective()
        self.params = numerics.make_params(
            np.asarray(grid.bed, np.float32), np.asarray(grid.manning_n, np.float32), mask,
            multiplier, boundary_types=np.asarray(torch.as_tensor(boundary_config.types).cpu()),
            rainfall_times_s=rainfall_times_s, rainfall_rates_m_s=rainfall_rates_m_s,
            default_rainfall_m_s=default_rainfall_m_s, dx=grid.resolution_m,
            dy=grid.resolution_m, gravity=gravity_m_s2, cfl=cfl, dry_tolerance=dry_tolerance_m,
            max_dt=max_dt_s, device=device,
        )
        self.dtype = torch.float32
        self.variant = variant

    @property
    def device(self) -> torch.device:
        return self.params.bed.device

    def initial_state(self, *, depth_m=0.0, x_velocity_m_s=0.0, y_velocity_m_s=0.0) -> SWEState:
        """Masked conserved fields from depth and velocities (shallow_water_2d.py:255-276)."""
        shape = tuple(self.params.bed.shape)

        def field(x):
            return torch.as_tensor(np.asarray(x, np.float32) if not isinstance(x, torch.Tensor) else x,
                                   dtype=torch.float32, device=self.device).broadcast_to(shape)

        depth, u, v = field(depth_m), field(x_velocity_m_s), field(y_velocity_m_s)
        mask = self.params.domain_mask.bool()
        depth = torch.where(mask, torch.clamp_min(depth, 0.0), torch.zeros((), device=self.device))
        return SWEState(depth.contiguous(), (depth * u).contiguous(), (depth * v).contiguous())

    def run(self, initial_state: SWEState, output_times_s, *, max_steps: int = 20_000) -> ScanResult:
        """Run the device-resident adaptive loop and verify all requested outputs."""
        host_times = np.asarray(
            output_times_s.detach().cpu().numpy() if isinstance(output_times_s, torch.Tensor)
            else output_times_s
        )
        if host_times.ndim != 1 or len(host_times) < 1:
            raise ValueError("Output times must be a non-empty one-dimensional array.")
        if host_times[0] != 0 or np.any(np.diff(host_times) <= 0):
            raise ValueError("Output times must start at zero and increase strictly.")
        result = numerics.integrate_to_outputs(
            initial_state, self.params, host_times, max_steps=max_steps, variant=self.variant
        )
        if result.diagnostics and result.diagnostics.get("nonfinite"):
            raise RuntimeError(
                f"A non-finite value appeared in the state by t={result.final_time_s:.3f}s."
            )
        if result.outputs_written != len(host_times):
            raise RuntimeError(
                f"max_steps={max_steps} was insufficient: wrote "
                f"{result.outputs_written}/{len(host_times)} outputs and reached "
                f"t={float(result.final_time_s):.3f}s."
            )
        return result


def load_structured_grid(dem_path, roughness_path, domain_path, *, resolution_m: float = 50.0):
    """GeoTIFF DEM + GeoPackage roughness and domain onto one grid
    (shallow_water_2d.py:72-181); implemented in ``ingestion``."""
    from .ingestion import load_structured_grid as _load

    return _load(dem_path, roughness_path, domain_path, resolution_m=resolution_m)


def load_rainfall_series(rainfall_path, *, scenario: str, default_rate_mm_hr: float = 0.0):
    """Parquet mm/h rain series in m/s (shallow_water_2d.py:184-207); see ``ingestion``."""
    from .ingestion import load_rainfall_series as _load

    return _load(rainfall_path, scenario=scenario, default_rate_mm_hr=default_rate_mm_hr)
