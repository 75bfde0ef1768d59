"""The forecast schema as a NetCDF file (structured_io.save_netcdf), on the CPU.

The reference writes its ANUGA-compatible schema with xarray's netCDF4 engine
(src/hybrid_flood/structured_io.py:214-230).  Here the same variables,
dimensions, coordinates, units and global attributes go into a classic
(64-bit offset) NetCDF file through SciPy's writer; the test reads the file
back with SciPy's reader and with a hand parse of the classic header's magic
and version byte, and checks every value bit for bit (NaN outside the mask).
"""

from __future__ import annotations

import numpy as np
import torch

from hybrid_flood_amd.structured_io import VARIABLE_UNITS, ForecastDataset, load_netcdf, save_netcdf


def _dataset():
    rng = np.random.default_rng(4)
    T, ny, nx = 3, 5, 7
    mask = rng.uniform(size=(ny, nx)) > 0.2
    elev = np.where(mask, rng.normal(size=(ny, nx)), np.nan).astype(np.float32)
    fields = {k: np.where(mask, rng.normal(size=(T, ny, nx)), np.nan).astype(np.float32)
              for k in ("depth", "x_velocity", "y_velocity", "velocity")}
    data = {"elevation": (("y", "x"), torch.from_numpy(elev))}
    data.update({k: (("time", "y", "x"), torch.from_numpy(v)) for k, v in fields.items()})
    coords = {"time": np.array([0.0, 60.0, 120.0]), "x": 1000.0 + 2.0 * np.arange(nx),
              "y": 5000.0 + 2.0 * np.arange(ny)}
    attrs = {"title": "test forecast", "crs": "EPSG:32643", "grid_resolution_m": 2.0,
             "dry_tolerance_m": 1e-3, "numerical_flux": "HLL with hydrostatic reconstruction",
             "execution_backend": "hip-gfx950", "steps_used": 17, "unset": None}
    var_attrs = {name: {"units": u} for name, u in VARIABLE_UNITS.items()}
    var_attrs["time"] = {"units_description": "seconds since simulation start"}
    return ForecastDataset(data, coords, attrs, var_attrs), elev, fields


def test_netcdf_round_trip(tmp_path):
    ds, elev, fields = _dataset()
    path = save_netcdf(ds, tmp_path / "nested" / "forecast.nc")
    head = path.read_bytes()[:4]
    assert head == b"CDF\x02"  # classic format, 64-bit offsets
    back = load_netcdf(path)
    assert np.array_equal(back["elevation"], elev, equal_nan=True)
    for k, v in fields.items():
        assert back[k].dtype == np.float32
        assert np.array_equal(back[k], v, equal_nan=True), k
        assert back["__dims__"][k] == ("time", "y", "x")
    assert back["__dims__"]["elevation"] == ("y", "x")
    np.testing.assert_array_equal(back["time"], [0.0, 60.0, 120.0])
    np.testing.assert_array_equal(back["x"], ds.coords["x"])
    np.testing.assert_array_equal(back["y"], ds.coords["y"])
    a = back["__attrs__"]
    assert a["title"] == "test forecast" and a["crs"] == "EPSG:32643"
    assert a["steps_used"] == 17 and a["grid_resolution_m"] == 2.0
    assert a["numerical_flux"] == "HLL with hydrostatic reconstruction"
    assert "unset" not in a
    for name, u in VARIABLE_UNITS.items():
        assert back["__var_attrs__"][name]["units"] == u
    assert back["__var_attrs__"]["time"]["units_description"] == "seconds since simulation start"
