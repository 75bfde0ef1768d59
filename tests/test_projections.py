"""Reprojection of roughness-layer coordinates (hybrid_flood_amd.projections).

PROJ is absent, so the transverse Mercator and the Mercator are pinned by the
mathematics (each check would fail on a dropped term, a wrong sign or a
transposed axis):
* on the central meridian the easting is the false easting and the northing
  is 0.9996 times the meridian arc, the arc computed here by Gauss-Legendre
  quadrature of the meridional radius of curvature (independent of the
  series);
* the map is conformal: by finite differences the scale east and north is the
  same and the two image directions are at right angles, and the scale grows
  off the central meridian;
* forward and inverse compose to the identity over the zone and between
  adjacent zones, in both hemispheres;
* the Pseudo-Mercator is x = a lon, y = a asinh(tan lat): its x at 180
  degrees is a pi.
"""

from __future__ import annotations

import numpy as np
import pytest
from numpy.polynomial.legendre import leggauss

from hybrid_flood_amd import projections as pj


def _meridian_arc(lat_deg: float) -> float:
    a, e2 = pj.WGS84_A, pj.WGS84_F * (2.0 - pj.WGS84_F)
    phi = np.radians(lat_deg)
    x, w = leggauss(64)
    t = 0.5 * phi * (x + 1.0)
    return float(0.5 * phi * np.sum(w * a * (1.0 - e2) / (1.0 - e2 * np.sin(t) ** 2) ** 1.5))


@pytest.mark.parametrize("lat", [0.0, 5.0, 28.4, 45.0, 66.0, 84.0])
def test_central_meridian_is_the_scaled_meridian_arc(lat):
    for code, fn in ((32643, 0.0), (32743, 10_000_000.0)):
        for sign in (1.0, -1.0):
            x, y = pj.transform(75.0, sign * lat, "EPSG:4326", f"EPSG:{code}")
            assert float(x) == pytest.approx(500_000.0, abs=1e-6)
            assert float(y) - fn == pytest.approx(sign * 0.9996 * _meridian_arc(lat), abs=1e-5)


@pytest.mark.parametrize("lon,lat", [(77.5, 28.4), (72.1, 10.0), (75.4, 60.0), (78.2, -45.0),
                                     (72.0, 0.5), (76.0, 80.0)])
def test_transverse_mercator_is_conformal(lon, lat):
    a, e2 = pj.WGS84_A, pj.WGS84_F * (2.0 - pj.WGS84_F)
    h = 1e-5
    x0, y0 = pj.tm_forward(lon, lat, 75.0)
    xe, ye = pj.tm_forward(lon + h, lat, 75.0)
    xn, yn = pj.tm_forward(lon, lat + h, 75.0)
    phi = np.radians(lat)
    w = np.sqrt(1.0 - e2 * np.sin(phi) ** 2)
    east = np.radians(h) * a / w * np.cos(phi)           # ground distance of the lon step
    north = np.radians(h) * a * (1.0 - e2) / w ** 3      # ground distance of the lat step
    ve = np.array([xe - x0, ye - y0]) / east
    vn = np.array([xn - x0, yn - y0]) / north
    ke, kn = np.linalg.norm(ve), np.linalg.norm(vn)
    assert ke == pytest.approx(kn, rel=2e-6)
    assert abs(float(ve @ vn)) / (ke * kn) < 2e-6
    # the scale is k0 on the central meridian and grows with the distance from it
    assert ke >= 0.9996 * (1.0 - 1e-6)
    assert ke == pytest.approx(0.9996 * (1.0 + 0.5 * ((x0 / (0.9996 * a)) ** 2)), rel=2e-4)


def test_round_trips_in_and_between_zones():
    rng = np.random.default_rng(5)
    lon = 75.0 + rng.uniform(-3.0, 3.0, 5000)
    lat = rng.uniform(-80.0, 84.0, 5000)
    north = lat >= 0
    for code in (32643, 32743):
        sel = north if code == 32643 else ~north
        x, y = pj.transform(lon[sel], lat[sel], "EPSG:4326", f"EPSG:{code}")
        lo, la = pj.transform(x, y, f"EPSG:{code}", "EPSG:4326")
        assert np.abs(lo - lon[sel]).max() < 1e-10 and np.abs(la - lat[sel]).max() < 1e-10
    # through geographic coordinates between adjacent zones
    x, y = pj.transform(lon[north], lat[north], "EPSG:4326", "EPSG:32643")
    x44, y44 = pj.transform(x, y, "EPSG:32643", "EPSG:32644")
    xb, yb = pj.transform(x44, y44, "EPSG:32644", "EPSG:32643")
    assert np.abs(xb - x).max() < 1e-6 and np.abs(yb - y).max() < 1e-6
    lo44, la44 = pj.transform(x44, y44, "EPSG:32644", "EPSG:4326")
    assert np.abs(lo44 - lon[north]).max() < 1e-10 and np.abs(la44 - lat[north]).max() < 1e-10


def test_pseudo_mercator_closed_form():
    x, y = pj.transform(180.0, 0.0, "EPSG:4326", "EPSG:3857")
    assert float(x) == pytest.approx(np.pi * pj.WGS84_A, abs=1e-6) and float(y) == 0.0
    lat = np.array([-70.0, -10.0, 0.0, 33.3, 85.0])
    x, y = pj.transform(np.full(5, -12.5), lat, "EPSG:4326", "EPSG:3857")
    np.testing.assert_allclose(y, pj.WGS84_A * np.log(np.tan(np.pi / 4 + np.radians(lat) / 2)),
                               rtol=1e-12, atol=1e-6)
    np.testing.assert_allclose(x, pj.WGS84_A * np.radians(-12.5), rtol=1e-14)
    lo, la = pj.transform(x, y, "EPSG:3857", "EPSG:4326")
    np.testing.assert_allclose(la, lat, atol=1e-11)
    np.testing.assert_allclose(lo, -12.5, atol=1e-11)
    # Mercator to UTM goes through geographic coordinates
    u = pj.transform(x, y, "EPSG:3857", "EPSG:32629")
    v = pj.transform(lo, la, "EPSG:4326", "EPSG:32629")
    np.testing.assert_allclose(u, v, atol=1e-6)


def test_supported_crs_and_errors():
    assert pj.is_supported("EPSG:32643") and pj.is_supported("epsg:4326") and pj.is_supported(3857)
    assert not pj.is_supported("EPSG:2193") and not pj.is_supported(None)
    with pytest.raises(ValueError, match="not supported"):
        pj.transform(0.0, 0.0, "EPSG:4326", "EPSG:27700")
    x, y = pj.transform(np.array([1.0, 2.0]), np.array([3.0, 4.0]), "EPSG:32643", "EPSG:32643")
    np.testing.assert_array_equal(x, [1.0, 2.0])
    np.testing.assert_array_equal(y, [3.0, 4.0])
