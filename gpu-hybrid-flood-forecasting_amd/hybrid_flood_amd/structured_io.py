"""The common forecast schema of the reference's outputs, without xarray.

``result_to_dataset`` / ``conserved_to_dataset`` build the variables of the
reference's ANUGA-compatible NetCDF schema (jax_solver/shallow_water_2d.py:
309-350, structured_io.py:160-211) on the GPU in one pass
(``swe_conserved_to_primitive``): ``elevation`` (y, x), ``depth``,
``x_velocity``, ``y_velocity``, ``velocity`` (time, y, x), float32, NaN
outside the domain mask, velocities ``hu/h`` and ``hv/h`` where ``h > dry``,
speed ``hypot(u, v)``; coordinates ``time`` (float64 s), ``x``, ``y``; the
same attributes.  xarray and netCDF4 are not available here, so the result
is a small ``ForecastDataset``.  ``save_netcdf`` / ``ForecastDataset.to_netcdf``
write it as a NetCDF file (structured_io.py:214-230) in the classic
64-bit-offset format through SciPy's writer, which any NetCDF reader
(xarray, netCDF4, ncdump) opens with the same variables, dimensions,
coordinates and attributes; the reference's netCDF4/HDF5 encoding (zlib level
4, 256 x 256 chunks) needs libraries absent here, so the data are stored
uncompressed.  ``to_npz`` writes the same content to one ``.npz`` file
(attributes as JSON).
"""

from __future__ import annotations

import json
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import numpy as np
import torch

from . import _lib
from .numerics import ScanResult, SWEParams, context_for
from .shallow_water_2d import load_rainfall_series, load_structured_grid  # noqa: F401  (structured_io.py:44-157)

VARIABLE_UNITS = {"elevation": "m", "depth": "m", "x_velocity": "m s-1", "y_velocity": "m s-1",
                  "velocity": "m s-1"}


@dataclass
class ForecastDataset:
    """Variables as (dims, device tensor), coordinates and attributes."""

    data_vars: dict[str, tuple[tuple[str, ...], torch.Tensor]]
    coords: dict[str, np.ndarray]
    attrs: dict[str, Any] = field(default_factory=dict)
    var_attrs: dict[str, dict[str, str]] = field(default_factory=dict)

    def __getitem__(self, name: str) -> torch.Tensor:
        return self.data_vars[name][1]

    def to_npz(self, path: str | Path) -> Path:
        """Write variables (float32), coordinates and attributes to ``path``."""
        p = Path(path)
        p.parent.mkdir(parents=True, exist_ok=True)
        arrays = {name: t.detach().cpu().numpy().astype(np.float32)
                  for name, (_, t) in self.data_vars.items()}
        arrays.update({f"coord_{k}": np.asarray(v) for k, v in self.coords.items()})
        meta = {"attrs": self.attrs, "var_attrs": self.var_attrs,
                "dims": {name: list(d) for name, (d, _) in self.data_vars.items()}}
        np.savez_compressed(p, __meta__=np.array(json.dumps(meta, default=str)), **arrays)
        return p


    def to_netcdf(self, path: str | Path) -> Path:
        """Write the dataset as a classic (64-bit offset) NetCDF file: coordinate
        variables ``time`` (float64), ``x``, ``y``; data variables float32 with
        their ``units``; the global attributes (those that are None are left out)."""
        from scipy.io import netcdf_file

        p = Path(path)
        p.parent.mkdir(parents=True, exist_ok=True)
        f = netcdf_file(str(p), "w", version=2)
        try:
            for name in ("time", "y", "x"):
                v = np.asarray(self.coords[name])
                f.createDimension(name, int(v.shape[0]))
                var = f.createVariable(name, "f8", (name,))
                var[:] = v.astype(np.float64)
                for k, a in self.var_attrs.get(name, {}).items():
                    setattr(var, k, a)
            for name, (dims, t) in self.data_vars.items():
                var = f.createVariable(name, "f4", tuple(dims))
                var[:] = t.detach().cpu().numpy().astype(np.float32)
                for k, a in self.var_attrs.get(name, {}).items():
                    setattr(var, k, a)
            for k, a in self.attrs.items():
                if a is not None:
                    setattr(f, k, a)
        finally:
            f.close()
        return p


def save_netcdf(dataset: ForecastDataset, output_path: str | Path) -> Path:
    """structured_io.py:214-230: write ``dataset`` to ``output_path`` (parents created)."""
    return dataset.to_netcdf(output_path)


def load_netcdf(path: str | Path) -> dict:
    """Variables (arrays), ``__attrs__`` and ``__var_attrs__`` of a file written by ``save_netcdf``."""
    from scipy.io import netcdf_file

    f = netcdf_file(str(path), "r", mmap=False)
    try:
        # NetCDF stores big-endian values; return native-order arrays
        out = {name: np.asarray(v[:], dtype=v[:].dtype.newbyteorder("="))
               for name, v in f.variables.items()}
        out["__attrs__"] = {k: (v.decode() if isinstance(v, bytes) else v)
                            for k, v in f._attributes.items()}
        out["__var_attrs__"] = {name: {k: (a.decode() if isinstance(a, bytes) else a)
                                       for k, a in v._attributes.items()}
                                for name, v in f.variables.items()}
        out["__dims__"] = {name: tuple(v.dimensions) for name, v in f.variables.items()}
    finally:
        f.close()
    return out


def load_npz(path: str | Path) -> dict:
    """Read a file written by ``ForecastDataset.to_npz`` (no pickles)."""
    with np.load(path, allow_pickle=False) as z:
        out = {k: z[k] for k in z.files if k != "__meta__"}
        out["__meta__"] = json.loads(str(z["__meta__"]))
    return out


def _primitive(params: SWEParams, h: torch.Tensor, hu: torch.Tensor, hv: torch.Tensor):
    ctx = context_for(params)
    ny, nx = ctx.shape
    if h.dim() == 2:
        h, hu, hv = h[None], hu[None], hv[None]
    for t in (h, hu, hv):
        if t.shape[1:] != (ny, nx) or t.dtype != torch.float32 or not t.is_cuda:
            raise ValueError(f"states must be float32 (T, {ny}, {nx}) device tensors")
    h, hu, hv = (t.contiguous() for t in (h, hu, hv))
    T = h.shape[0]
    out = [torch.empty_like(h) for _ in range(4)]
    elev = torch.empty((ny, nx), dtype=torch.float32, device=h.device)
    _lib.check(
        _lib.load().swe_conserved_to_primitive(
            ctx.handle, h.data_ptr(), hu.data_ptr(), hv.data_ptr(), T, *(o.data_ptr() for o in out),
            elev.data_ptr(), _lib.stream_handle(h.device)),
        "swe_conserved_to_primitive")
    return elev, out


def _coords(params: SWEParams, times, x=None, y=None) -> dict[str, np.ndarray]:
    ny, nx = params.bed.shape
    dx = float(params.dx)
    dy = float(params.dy)
    return {
        "time": np.asarray(times.detach().cpu().numpy() if isinstance(times, torch.Tensor) else times,
                           dtype=np.float64),
        "x": np.asarray(x if x is not None else (np.arange(nx) + 0.5) * dx),
        "y": np.asarray(y if y is not None else (np.arange(ny) + 0.5) * dy),
    }


def conserved_to_dataset(h, hu, hv, output_times_s, params: SWEParams, *, dry_tolerance_m: float,
                         title: str, backend: str, steps_used: int, x=None, y=None,
                         crs: str | None = None) -> ForecastDataset:
    """structured_io.py:160-211 (``conserved_to_xarray``) on the GPU."""
    elev, (depth, vx, vy, speed) = _primitive(params, h, hu, hv)
    attrs = {"title": title, "crs": crs, "grid_resolution_m": float(params.dx),
             "dry_tolerance_m": dry_tolerance_m,
             "numerical_flux": "HLL with hydrostatic reconstruction",
             "execution_backend": backend, "steps_used": int(steps_used)}
    return ForecastDataset(
        {"elevation": (("y", "x"), elev), "depth": (("time", "y", "x"), depth),
         "x_velocity": (("time", "y", "x"), vx), "y_velocity": (("time", "y", "x"), vy),
         "velocity": (("time", "y", "x"), speed)},
        _coords(params, output_times_s, x, y), attrs,
        {name: {"units": u} for name, u in VARIABLE_UNITS.items()})


def result_to_dataset(result: ScanResult, params: SWEParams, *, dry_tolerance_m: float | None = None,
                      x=None, y=None, crs: str | None = None) -> ForecastDataset:
    """jax_solver/shallow_water_2d.py:309-350 (``result_to_xarray``) on the GPU.

    The dry tolerance of ``params`` is the one the conversion uses; passing a
    different ``dry_tolerance_m`` is rejected rather than silently ignored.
    """
    dry = float(params.dry_tolerance)
    if dry_tolerance_m is not None and float(np.float32(dry_tolerance_m)) != float(np.float32(dry)):
        raise ValueError("dry_tolerance_m must equal the solver's dry tolerance")
    ds = conserved_to_dataset(result.states.h, result.states.hu, result.states.hv,
                              result.output_times_s, params, dry_tolerance_m=dry,
                              title="JAX finite-volume shallow-water raw forecast",
                              backend="hip-gfx950", steps_used=result.steps_used, x=x, y=y, crs=crs)
    ds.attrs.pop("execution_backend")
    ds.var_attrs["time"] = {"units_description": "seconds since simulation start"}
    return ds
