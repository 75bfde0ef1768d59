"""Reprojection of vector coordinates between the WGS 84 CRSs a roughness layer comes in.

The reference moves the LULC roughness layer into the domain's CRS with
geopandas' ``to_crs`` when the two differ
(src/hybrid_flood/jax_solver/shallow_water_2d.py:86, structured_io.py:58);
that calls PROJ, which is not available here.  This module covers the CRSs of
the WGS 84 datum:

* ``EPSG:4326``  geographic; x = longitude, y = latitude in degrees (the
  traditional GIS axis order that geopandas uses);
* ``EPSG:3857``  WGS 84 / Pseudo-Mercator: spherical Mercator on a = 6378137 m;
* ``EPSG:326zz`` / ``EPSG:327zz``  WGS 84 / UTM zone zz, north / south:
  transverse Mercator, central meridian 6 zz - 183 degrees, scale 0.9996,
  false easting 500 km, false northing 0 / 10,000 km (the reference's roughness
  layer is in EPSG:32643, config/jax_solver/shallow_water.yaml:22).

Every transformation goes through geographic coordinates, and vertices move
one by one, as ``to_crs`` moves them.  The transverse Mercator is Krueger's
series in the third flattening n, to order n^6, in the form published by
Karney ("Transverse Mercator with an accuracy of a few nanometers",
J. Geodesy 85, 2011, eqs. 7-9 and 35-36), which is also the series of PROJ's
UTM; its error inside a UTM zone is far below a millimetre.  The inverse finds
the latitude from the conformal one by Newton's method (Karney eqs. 19-21).

Bit parity with PROJ is unpinned (pyproj is absent).  The tests pin the
mathematics: on the central meridian the northing is 0.9996 times the
meridian arc, computed independently by quadrature; the map is conformal
(equal scale east and north, at right angles, by finite differences);
forward and inverse compose to the identity; the Mercator is its closed form.
"""

from __future__ import annotations

import re
from typing import Sequence

import numpy as np

# WGS 84 ellipsoid
WGS84_A = 6378137.0
WGS84_F = 1.0 / 298.257223563
_E2 = WGS84_F * (2.0 - WGS84_F)
_E = np.sqrt(_E2)
_N = WGS84_F / (2.0 - WGS84_F)  # third flattening

UTM_K0 = 0.9996
UTM_FE = 500_000.0
UTM_FN_SOUTH = 10_000_000.0


def _kruger_series(n: float):
    """A (rectifying radius over a) and the alpha / beta series of Krueger, to n^6."""
    n2, n3, n4, n5, n6 = n * n, n ** 3, n ** 4, n ** 5, n ** 6
    a_ratio = (1.0 + n2 / 4.0 + n4 / 64.0 + n6 / 256.0) / (1.0 + n)
    alpha = (
        n / 2 - 2 * n2 / 3 + 5 * n3 / 16 + 41 * n4 / 180 - 127 * n5 / 288 + 7891 * n6 / 37800,
        13 * n2 / 48 - 3 * n3 / 5 + 557 * n4 / 1440 + 281 * n5 / 630 - 1983433 * n6 / 1935360,
        61 * n3 / 240 - 103 * n4 / 140 + 15061 * n5 / 26880 + 167603 * n6 / 181440,
        49561 * n4 / 161280 - 179 * n5 / 168 + 6601661 * n6 / 7257600,
        34729 * n5 / 80640 - 3418889 * n6 / 1995840,
        212378941 * n6 / 319334400,
    )
    beta = (
        n / 2 - 2 * n2 / 3 + 37 * n3 / 96 - n4 / 360 - 81 * n5 / 512 + 96199 * n6 / 604800,
        n2 / 48 + n3 / 15 - 437 * n4 / 1440 + 46 * n5 / 105 - 1118711 * n6 / 3870720,
        17 * n3 / 480 - 37 * n4 / 840 - 209 * n5 / 4480 + 5569 * n6 / 90720,
        4397 * n4 / 161280 - 11 * n5 / 504 - 830251 * n6 / 7257600,
        4583 * n5 / 161280 - 108847 * n6 / 3991680,
        20648693 * n6 / 638668800,
    )
    return a_ratio, alpha, beta


_A_RATIO, _ALPHA, _BETA = _kruger_series(_N)


def _conformal_tan(tau: np.ndarray) -> np.ndarray:
    """tan of the conformal latitude from tau = tan(latitude) (Karney eq. 7-9)."""
    sig = np.sinh(_E * np.arctanh(_E * tau / np.hypot(1.0, tau)))
    return tau * np.hypot(1.0, sig) - sig * np.hypot(1.0, tau)


def _tau_from_conformal(taup: np.ndarray) -> np.ndarray:
    """tan(latitude) from tan(conformal latitude), Newton's method (Karney eqs. 19-21)."""
    tau = taup / (1.0 - _E2)
    for _ in range(20):
        tp = _conformal_tan(tau)
        dtp = ((1.0 - _E2) * np.hypot(1.0, tp) * np.hypot(1.0, tau)
               / (1.0 + (1.0 - _E2) * tau * tau))
        step = (taup - tp) / dtp
        tau = tau + step
        if np.all(np.abs(step) <= 1e-15 * np.maximum(1.0, np.abs(tau))):
            break
    return tau


def tm_forward(lon_deg, lat_deg, lon0_deg: float, k0: float = UTM_K0):
    """Transverse Mercator easting / northing (metres, no false origin) of geographic points."""
    lon = np.radians(np.asarray(lon_deg, np.float64))
    lat = np.radians(np.asarray(lat_deg, np.float64))
    dl = np.remainder(lon - np.radians(lon0_deg) + np.pi, 2.0 * np.pi) - np.pi
    with np.errstate(divide="ignore", invalid="ignore"):
        taup = _conformal_tan(np.tan(lat))
        xip = np.arctan2(taup, np.cos(dl))
        etap = np.arcsinh(np.sin(dl) / np.hypot(taup, np.cos(dl)))
    xi, eta = xip.copy(), etap.copy()
    for j, a in enumerate(_ALPHA, start=1):
        xi += a * np.sin(2 * j * xip) * np.cosh(2 * j * etap)
        eta += a * np.cos(2 * j * xip) * np.sinh(2 * j * etap)
    scale = k0 * WGS84_A * _A_RATIO
    return scale * eta, scale * xi


def tm_inverse(x, y, lon0_deg: float, k0: float = UTM_K0):
    """Geographic longitude / latitude (degrees) of transverse Mercator easting / northing."""
    scale = k0 * WGS84_A * _A_RATIO
    xi = np.asarray(y, np.float64) / scale
    eta = np.asarray(x, np.float64) / scale
    xip, etap = xi.copy(), eta.copy()
    for j, b in enumerate(_BETA, start=1):
        xip -= b * np.sin(2 * j * xi) * np.cosh(2 * j * eta)
        etap -= b * np.cos(2 * j * xi) * np.sinh(2 * j * eta)
    s, c = np.sinh(etap), np.cos(xip)
    taup = np.sin(xip) / np.hypot(s, c)
    lat = np.degrees(np.arctan(_tau_from_conformal(taup)))
    lon = lon0_deg + np.degrees(np.arctan2(s, c))
    return lon, lat


def mercator_forward(lon_deg, lat_deg):
    """EPSG:3857 x / y of geographic points (spherical formulas on a = 6378137 m)."""
    lon = np.radians(np.asarray(lon_deg, np.float64))
    lat = np.radians(np.asarray(lat_deg, np.float64))
    return WGS84_A * lon, WGS84_A * np.arcsinh(np.tan(lat))


def mercator_inverse(x, y):
    return (np.degrees(np.asarray(x, np.float64) / WGS84_A),
            np.degrees(np.arctan(np.sinh(np.asarray(y, np.float64) / WGS84_A))))


def _parse(crs) -> tuple[str, int | None]:
    """("geographic" | "mercator" | "utm", signed zone) of a supported CRS, else ValueError."""
    code = None
    if isinstance(crs, (int, np.integer)):
        code = int(crs)
    elif isinstance(crs, str):
        m = re.fullmatch(r"\s*EPSG\s*:\s*(\d+)\s*", crs, flags=re.IGNORECASE)
        if m:
            code = int(m.group(1))
    if code == 4326:
        return "geographic", None
    if code == 3857:
        return "mercator", None
    if code is not None and (32601 <= code <= 32660 or 32701 <= code <= 32760):
        zone = code % 100
        return "utm", zone if code < 32700 else -zone
    raise ValueError(f"CRS {crs!r} is not supported here (EPSG:4326, EPSG:3857, "
                     "EPSG:32601-32660 and EPSG:32701-32760 are)")


def is_supported(crs) -> bool:
    try:
        _parse(crs)
        return True
    except ValueError:
        return False


def _to_geographic(x, y, kind, zone):
    if kind == "geographic":
        return np.asarray(x, np.float64), np.asarray(y, np.float64)
    if kind == "mercator":
        return mercator_inverse(x, y)
    lon0 = 6.0 * abs(zone) - 183.0
    fn = 0.0 if zone > 0 else UTM_FN_SOUTH
    return tm_inverse(np.asarray(x, np.float64) - UTM_FE, np.asarray(y, np.float64) - fn, lon0)


def _from_geographic(lon, lat, kind, zone):
    if kind == "geographic":
        return lon, lat
    if kind == "mercator":
        return mercator_forward(lon, lat)
    lon0 = 6.0 * abs(zone) - 183.0
    fn = 0.0 if zone > 0 else UTM_FN_SOUTH
    e, n = tm_forward(lon, lat, lon0)
    return e + UTM_FE, n + fn


def transform(x, y, src_crs, dst_crs):
    """Coordinates of points (x, y) of ``src_crs`` in ``dst_crs``."""
    sk, sz = _parse(src_crs)
    dk, dz = _parse(dst_crs)
    if (sk, sz) == (dk, dz):
        return np.asarray(x, np.float64).copy(), np.asarray(y, np.float64).copy()
    lon, lat = _to_geographic(x, y, sk, sz)
    return _from_geographic(lon, lat, dk, dz)


def transform_polygons(polygons: Sequence[Sequence[np.ndarray]], src_crs, dst_crs):
    """Every vertex of every ring (``(n, 2)`` arrays) moved from ``src_crs`` to ``dst_crs``."""
    out = []
    for rings in polygons:
        moved = []
        for r in rings:
            r = np.asarray(r, np.float64)
            if r.size == 0:
                moved.append(r)
                continue
            xs, ys = transform(r[:, 0], r[:, 1], src_crs, dst_crs)
            moved.append(np.stack([xs, ys], axis=1))
        out.append(moved)
    return out
