"""Structured-grid ingestion without GDAL, GEOS or PROJ.

The reference rasterises its inputs once, on the host, before the step ever
runs (jax_solver/shallow_water_2d.py:49-181, duplicated in
structured_io.py:44-157):

* the domain (GeoPackage polygons, or closed perimeter lines) fixes a grid
  aligned to ``resolution_m`` (``_aligned_grid_geometry``) and its cell mask;
* the DEM (a deflate / floating-point-predictor, tiled float32 GeoTIFF,
  dem/export_geotiff.py:31-70) is resampled bilinearly onto the grid, holes
  take the nearest-neighbour resample, remaining holes the value of the
  nearest finite cell (a Euclidean distance transform);
* Manning's n is burnt from the LULC polygons' ``roughness`` attribute;
* rows are flipped so that the row index increases northwards, and cells
  outside the domain get bed 0 and n 0;
* the rain series is read from a parquet table in mm/h and converted to m/s.

rasterio, geopandas, shapely, pyproj and affine are not available here, so
this module reads the two file formats itself (TIFF 6.0 with the GeoTIFF
tags; GeoPackage, which is SQLite with standard WKB geometry) and
implements the raster operations in NumPy:

* polygon burning follows the pixel-centre scanline rule (a cell is inside
  when its centre is, edges half-open), and later features overwrite
  earlier ones, as rasterio's ``rasterize``/``geometry_mask`` do with
  ``all_touched=False``;
* a roughness layer in another CRS than the domain's is moved into it vertex
  by vertex (``projections``: WGS 84 geographic, Pseudo-Mercator and UTM, the
  reference's ``to_crs``); other CRS pairs are refused;
* resampling handles one CRS only (the reference also refuses a DEM whose CRS
  differs from the domain's); "bilinear" is a separable tent whose support
  widens with the downsampling factor, renormalised over valid source cells,
  and "nearest" takes the source cell that contains the destination centre.

These follow the published rules of those libraries but cannot be checked
against them here (they are not installed): bit parity with GDAL's warper
and rasteriser is unpinned; tests pin the closed forms (exact planes, exact
polygon masks, nearest fills by brute force, format round trips).
"""

from __future__ import annotations

import sqlite3
import struct
import zlib
from dataclasses import dataclass, field
from math import ceil, floor
from pathlib import Path
from typing import Any, Iterable, Sequence

import numpy as np

from . import projections
from .shallow_water_2d import StructuredGrid

MM_HR_TO_M_S = 1.0e-3 / 3600.0


# ---------------------------------------------------------------------------
# Geometry containers
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class NorthUpTransform:
    """Affine map from (col, row) to (x, y): x = a*col + b*row + c, y = d*col + e*row + f."""

    a: float
    b: float
    c: float
    d: float
    e: float
    f: float

    @classmethod
    def from_origin(cls, west: float, north: float, xsize: float, ysize: float) -> "NorthUpTransform":
        return cls(float(xsize), 0.0, float(west), 0.0, -float(ysize), float(north))

    def __iter__(self):
        return iter((self.a, self.b, self.c, self.d, self.e, self.f))

    @property
    def axis_aligned(self) -> bool:
        return self.b == 0.0 and self.d == 0.0


Ring = np.ndarray  # (n, 2) float64 vertices; the closing vertex may repeat the first


@dataclass
class VectorLayer:
    """Features of one GeoPackage table: polygons (list of rings each) and attributes."""

    polygons: list[list[Ring]]
    attributes: dict[str, list]
    crs: str | None = None
    projected: bool | None = None
    kinds: list[str] = field(default_factory=list)

    @property
    def columns(self) -> list[str]:
        return list(self.attributes)


@dataclass
class Raster:
    """One band of a GeoTIFF, rows north to south as stored."""

    data: np.ndarray
    transform: NorthUpTransform
    nodata: float | None
    crs: str | None
    projected: bool | None


# ---------------------------------------------------------------------------
# TIFF / GeoTIFF reader
# ---------------------------------------------------------------------------
_TIFF_TYPES = {1: ("B", 1), 2: ("s", 1), 3: ("H", 2), 4: ("I", 4), 5: ("II", 8), 6: ("b", 1),
               7: ("B", 1), 8: ("h", 2), 9: ("i", 4), 10: ("ii", 8), 11: ("f", 4), 12: ("d", 8),
               16: ("Q", 8), 17: ("q", 8)}


def _lzw_decode(data: bytes) -> bytes:
    """TIFF 6.0 LZW (MSB-first codes, early change, 256 = clear, 257 = end)."""
    out = bytearray()
    table: list[bytes] = [bytes([i]) for i in range(256)] + [b"", b""]
    width, buf, nbits, pos, prev = 9, 0, 0, 0, None
    while True:
        while nbits < width:
            if pos >= len(data):
                return bytes(out)
            buf = (buf << 8) | data[pos]
            pos += 1
            nbits += 8
        code = (buf >> (nbits - width)) & ((1 << width) - 1)
        nbits -= width
        if code == 257:
            return bytes(out)
        if code == 256:
            table = table[:258]
            width, prev = 9, None
            continue
        if prev is None:
            entry = table[code]
        elif code < len(table):
            entry = table[code]
            table.append(prev + entry[:1])
        else:
            entry = prev + prev[:1]
            table.append(entry)
        out += entry
        prev = entry
        if len(table) + 1 >= (1 << width) and width < 12:
            width += 1


def _undo_predictor(block: np.ndarray, predictor: int, dtype: np.dtype, spp: int) -> np.ndarray:
    """Undo TIFF predictors on a (rows, bytes_per_row) uint8 block; returns (rows, cols*spp) values."""
    rows = block.shape[0]
    item = dtype.itemsize
    if predictor == 1:
        return block.reshape(rows, -1).view(dtype)
    if predictor == 2:
        vals = block.reshape(rows, -1).view(dtype).reshape(rows, -1, spp)
        return np.cumsum(vals, axis=1, dtype=dtype).reshape(rows, -1)
    if predictor == 3:
        # Floating-point predictor (Adobe TIFF Technote 3): bytes are
        # differenced along the row (stride = samples per pixel), and each
        # row stores the most significant byte of every sample first.
        acc = block.reshape(rows, -1, spp).astype(np.uint64).cumsum(axis=1) & 0xFF
        acc = acc.astype(np.uint8).reshape(rows, -1)
        n = acc.shape[1] // item
        planes = acc.reshape(rows, item, n)  # plane 0 = most significant byte
        big = np.ascontiguousarray(planes.transpose(0, 2, 1)).reshape(rows, n * item)
        return big.view(dtype.newbyteorder(">")).astype(dtype)
    raise ValueError(f"Unsupported TIFF predictor {predictor}.")


def _read_ifd(buf: bytes, off: int, bo: str, big: bool) -> dict[int, Any]:
    tags: dict[int, Any] = {}
    if big:
        (n,) = struct.unpack_from(bo + "Q", buf, off)
        off += 8
        esize, inline = 20, 8
    else:
        (n,) = struct.unpack_from(bo + "H", buf, off)
        off += 2
        esize, inline = 12, 4
    for k in range(n):
        e = off + k * esize
        if big:
            tag, typ, count = struct.unpack_from(bo + "HHQ", buf, e)
            vpos = e + 12
        else:
            tag, typ, count = struct.unpack_from(bo + "HHI", buf, e)
            vpos = e + 8
        if typ not in _TIFF_TYPES:
            continue
        fmt, size = _TIFF_TYPES[typ]
        total = size * count
        if total > inline:
            (vpos,) = struct.unpack_from(bo + ("Q" if big else "I"), buf, vpos)
        if typ == 2:
            tags[tag] = buf[vpos:vpos + count].split(b"\0", 1)[0].decode("ascii", "replace")
        elif typ in (5, 10):
            raw = struct.unpack_from(bo + fmt[0] * (2 * count), buf, vpos)
            tags[tag] = [raw[2 * i] / raw[2 * i + 1] if raw[2 * i + 1] else float("nan")
                         for i in range(count)]
        else:
            tags[tag] = list(struct.unpack_from(bo + fmt * count, buf, vpos))
    return tags


def _geokeys(tags: dict[int, Any]) -> tuple[str | None, bool | None]:
    """(``EPSG:code`` or None, projected?) from the GeoKeyDirectoryTag (34735)."""
    keys = tags.get(34735)
    if not keys or len(keys) < 4:
        return None, None
    model, proj, geog = None, None, None
    for i in range(keys[3]):
        kid, loc, _cnt, val = keys[4 + 4 * i: 8 + 4 * i]
        if loc != 0:
            continue
        if kid == 1024:
            model = val
        elif kid == 3072:
            proj = val
        elif kid == 2048:
            geog = val
    if model == 1 and proj not in (None, 0, 32767):
        return f"EPSG:{proj}", True
    if model == 2 and geog not in (None, 0, 32767):
        return f"EPSG:{geog}", False
    return None, (model == 1) if model in (1, 2) else None


def read_geotiff(path: str | Path, band: int = 1) -> Raster:
    """Read one band of a (Big)TIFF with GeoTIFF georeferencing.

    Strips or tiles; no compression, LZW or deflate; predictors 1, 2, 3;
    8-64 bit integer or IEEE float samples, chunky or planar.  The transform
    comes from ModelPixelScale + ModelTiepoint (or ModelTransformation), the
    no-data value from the GDAL_NODATA tag (42113), the CRS from the
    GeoKeyDirectory.
    """
    buf = Path(path).read_bytes()
    if buf[:2] == b"II":
        bo = "<"
    elif buf[:2] == b"MM":
        bo = ">"
    else:
        raise ValueError(f"{path}: not a TIFF file.")
    (magic,) = struct.unpack_from(bo + "H", buf, 2)
    if magic == 42:
        big = False
        (ifd,) = struct.unpack_from(bo + "I", buf, 4)
    elif magic == 43:
        big = True
        (ifd,) = struct.unpack_from(bo + "Q", buf, 8)
    else:
        raise ValueError(f"{path}: not a TIFF file.")
    t = _read_ifd(buf, ifd, bo, big)
    width, height = t[256][0], t[257][0]
    spp = t.get(277, [1])[0]
    bps = t.get(258, [1])[0]
    fmt = t.get(339, [1])[0]
    comp = t.get(259, [1])[0]
    pred = t.get(317, [1])[0]
    planar = t.get(284, [1])[0]
    if not 1 <= band <= spp:
        raise ValueError(f"{path}: band {band} out of range (1..{spp}).")
    kind = {1: "u", 2: "i", 3: "f"}.get(fmt)
    if kind is None or bps not in (8, 16, 32, 64):
        raise ValueError(f"{path}: unsupported sample format {fmt}/{bps}.")
    dtype = np.dtype(f"{kind}{bps // 8}").newbyteorder(bo)
    if comp == 1:
        decode = bytes
    elif comp in (8, 32946):
        decode = zlib.decompress
    elif comp == 5:
        decode = _lzw_decode
    else:
        raise ValueError(f"{path}: unsupported TIFF compression {comp}.")
    tiled = 322 in t
    if tiled:
        bw, bh = t[322][0], t[323][0]
        offs, counts = t[324], t[325]
    else:
        bw, bh = width, t.get(278, [height])[0]
        offs, counts = t[273], t[279]
    across = (width + bw - 1) // bw
    down = (height + bh - 1) // bh
    per_plane = across * down
    cs = 1 if planar == 2 else spp  # samples interleaved in a block
    out = np.empty((height, width), dtype=dtype.newbyteorder("="))
    plane = band - 1 if planar == 2 else 0
    for by in range(down):
        for bx in range(across):
            k = plane * per_plane + by * across + bx
            raw = decode(buf[offs[k]:offs[k] + counts[k]])
            rows_here = bh if tiled else min(bh, height - by * bh)
            row_bytes = bw * cs * dtype.itemsize
            blk = np.frombuffer(raw, np.uint8, count=rows_here * row_bytes).reshape(rows_here, row_bytes)
            vals = _undo_predictor(blk, pred, dtype, cs).reshape(rows_here, bw, cs)
            vals = vals[:, :, band - 1 if planar != 2 else 0]
            y0, x0 = by * bh, bx * bw
            h = min(rows_here, height - y0)
            w = min(bw, width - x0)
            out[y0:y0 + h, x0:x0 + w] = vals[:h, :w]
    if 34264 in t:
        m = t[34264]
        tr = NorthUpTransform(m[0], m[1], m[3], m[4], m[5], m[7])
    elif 33550 in t and 33922 in t:
        sx, sy = t[33550][0], t[33550][1]
        i, j, _k, x, y, _z = t[33922][:6]
        tr = NorthUpTransform(sx, 0.0, x - i * sx, 0.0, -sy, y + j * sy)
    else:
        tr = NorthUpTransform(1.0, 0.0, 0.0, 0.0, 1.0, 0.0)
    nodata = None
    if 42113 in t and str(t[42113]).strip():
        nodata = float(str(t[42113]).strip())
    crs, projected = _geokeys(t)
    return Raster(out, tr, nodata, crs, projected)


# ---------------------------------------------------------------------------
# GeoPackage reader
# ---------------------------------------------------------------------------
def _wkb_geometry(blob: bytes, pos: int = 0) -> tuple[str, list, int]:
    """Decode one WKB geometry -> (kind, rings or list of polygons, next position)."""
    bo = "<" if blob[pos] == 1 else ">"
    (gtype,) = struct.unpack_from(bo + "I", blob, pos + 1)
    pos += 5
    has_z = has_m = False
    if gtype & 0x80000000 or gtype & 0x40000000:  # EWKB flags
        has_z, has_m = bool(gtype & 0x80000000), bool(gtype & 0x40000000)
        gtype &= 0x0FFFFFFF
    if gtype >= 3000:
        has_z = has_m = True
        gtype -= 3000
    elif gtype >= 2000:
        has_m = True
        gtype -= 2000
    elif gtype >= 1000:
        has_z = True
        gtype -= 1000
    dim = 2 + has_z + has_m

    def points(p):
        (n,) = struct.unpack_from(bo + "I", blob, p)
        p += 4
        arr = np.frombuffer(blob, dtype=np.dtype("f8").newbyteorder(bo), count=n * dim, offset=p)
        return arr.reshape(n, dim)[:, :2].astype(np.float64), p + 8 * n * dim

    if gtype == 2:  # LineString
        pts, pos = points(pos)
        return "LineString", [pts], pos
    if gtype == 3:  # Polygon
        (nr,) = struct.unpack_from(bo + "I", blob, pos)
        pos += 4
        rings = []
        for _ in range(nr):
            r, pos = points(pos)
            rings.append(r)
        return "Polygon", rings, pos
    if gtype in (5, 6, 7):  # MultiLineString, MultiPolygon, GeometryCollection
        (n,) = struct.unpack_from(bo + "I", blob, pos)
        pos += 4
        parts = []
        for _ in range(n):
            kind, geom, pos = _wkb_geometry(blob, pos)
            parts.append((kind, geom))
        return {5: "MultiLineString", 6: "MultiPolygon", 7: "GeometryCollection"}[gtype], parts, pos
    raise ValueError(f"Unsupported WKB geometry type {gtype}.")


def _gpkg_blob_to_wkb(blob: bytes) -> bytes | None:
    """Strip the GeoPackage binary header (``GP``, version, flags, srs id, envelope)."""
    if blob is None:
        return None
    if blob[:2] != b"GP":
        return bytes(blob)
    flags = blob[3]
    if flags & 0x10:  # empty geometry
        return None
    env = (flags >> 1) & 0x07
    env_len = {0: 0, 1: 32, 2: 48, 3: 48, 4: 64}.get(env)
    if env_len is None:
        raise ValueError("Invalid GeoPackage envelope indicator.")
    return bytes(blob[8 + env_len:])


def _polygons_of(kind: str, geom) -> list[list[Ring]]:
    """Polygons (as ring lists) of a decoded geometry; closed lines become polygons."""
    if kind == "Polygon":
        return [geom]
    if kind == "LineString":
        pts = geom[0]
        if len(pts) >= 4 and np.array_equal(pts[0], pts[-1]):
            return [[pts]]
        return []
    if kind == "MultiPolygon":
        return [g for _k, g in geom]
    if kind in ("MultiLineString", "GeometryCollection"):
        out: list[list[Ring]] = []
        for k, g in geom:
            out += _polygons_of(k, g)
        return out
    return []


def read_geopackage(path: str | Path, layer: str | None = None) -> VectorLayer:
    """Features of one feature table (the first one, by default) of a GeoPackage."""
    con = sqlite3.connect(f"file:{Path(path)}?mode=ro", uri=True)
    try:
        rows = con.execute(
            "SELECT table_name, column_name, srs_id FROM gpkg_geometry_columns").fetchall()
        if not rows:
            raise ValueError(f"{path}: no feature table.")
        if layer is None:
            table, gcol, srs_id = rows[0]
        else:
            match = [r for r in rows if r[0] == layer]
            if not match:
                raise ValueError(f"{path}: no layer {layer!r}.")
            table, gcol, srs_id = match[0]
        crs, projected = None, None
        srs = con.execute("SELECT organization, organization_coordsys_id, definition FROM "
                          "gpkg_spatial_ref_sys WHERE srs_id = ?", (srs_id,)).fetchone()
        if srs is not None:
            org, code, definition = srs
            if org and str(org).upper() == "EPSG" and code and int(code) > 0:
                crs = f"EPSG:{int(code)}"
            d = (definition or "").lstrip().upper()
            if d.startswith(("PROJCS", "PROJCRS", "PROJECTEDCRS")):
                projected = True
            elif d.startswith(("GEOGCS", "GEOGCRS", "GEODCRS", "GEOGRAPHICCRS")):
                projected = False
        cols = [r[1] for r in con.execute(f'PRAGMA table_info("{table}")').fetchall()]
        attrs = [c for c in cols if c != gcol and c.lower() != "fid"]
        sel = ", ".join(f'"{c}"' for c in [gcol] + attrs)
        polygons: list[list[Ring]] = []
        kinds: list[str] = []
        values: dict[str, list] = {c: [] for c in attrs}
        for rec in con.execute(f'SELECT {sel} FROM "{table}"'):
            wkb = _gpkg_blob_to_wkb(rec[0])
            if wkb is None:
                kind, geom = "Empty", []
            else:
                kind, geom, _ = _wkb_geometry(wkb)
            kinds.append(kind)
            polygons.append([r for poly in _polygons_of(kind, geom) for r in poly]
                            if kind != "Polygon" else geom)
            for c, v in zip(attrs, rec[1:]):
                values[c].append(v)
        return VectorLayer(polygons, values, crs, projected, kinds)
    finally:
        con.close()


# ---------------------------------------------------------------------------
# Grid geometry, burning, resampling, hole filling
# ---------------------------------------------------------------------------
def _bounds(polygons: Iterable[Sequence[Ring]]) -> tuple[float, float, float, float]:
    pts = [r for poly in polygons for r in poly if len(r)]
    if not pts:
        raise ValueError("Domain must contain polygon geometry or a closed line perimeter.")
    allp = np.concatenate(pts)
    return (float(allp[:, 0].min()), float(allp[:, 1].min()),
            float(allp[:, 0].max()), float(allp[:, 1].max()))


def aligned_grid_geometry(bounds: tuple[float, float, float, float], resolution_m: float):
    """(x, y, north-up transform, (height, width)) of the grid that covers ``bounds``
    with cell edges on multiples of ``resolution_m`` (shallow_water_2d.py:49-70)."""
    xmin, ymin, xmax, ymax = bounds
    ax0 = floor(xmin / resolution_m) * resolution_m
    ay0 = floor(ymin / resolution_m) * resolution_m
    ax1 = ceil(xmax / resolution_m) * resolution_m
    ay1 = ceil(ymax / resolution_m) * resolution_m
    width = int(round((ax1 - ax0) / resolution_m))
    height = int(round((ay1 - ay0) / resolution_m))
    transform = NorthUpTransform.from_origin(ax0, ay1, resolution_m, resolution_m)
    x = ax0 + (np.arange(width) + 0.5) * resolution_m
    y = ay0 + (np.arange(height) + 0.5) * resolution_m
    return x, y, transform, (height, width)


def _edges(rings: Sequence[Ring], transform: NorthUpTransform) -> np.ndarray:
    """Polygon edges in pixel coordinates (col, row), shape (E, 4): c0, r0, c1, r1."""
    if not transform.axis_aligned:
        raise ValueError("Rotated grids are not supported.")
    out = []
    for r in rings:
        r = np.asarray(r, np.float64)
        if len(r) < 3:
            continue
        if not np.array_equal(r[0], r[-1]):
            r = np.vstack([r, r[:1]])
        col = (r[:, 0] - transform.c) / transform.a
        row = (r[:, 1] - transform.f) / transform.e
        out.append(np.stack([col[:-1], row[:-1], col[1:], row[1:]], axis=1))
    return np.concatenate(out) if out else np.zeros((0, 4))


def polygon_cells(rings: Sequence[Ring], shape: tuple[int, int],
                  transform: NorthUpTransform) -> np.ndarray:
    """Cells whose centre lies inside the polygon (even-odd over all rings).

    Scanline rule: for the centre line of row r (row + 0.5), an edge crosses
    it when its end points straddle the line half-open (y0 <= yc < y1 or
    y1 <= yc < y0); crossings are paired after sorting and cell c is inside
    a span [xa, xb) when xa <= c + 0.5 < xb.
    """
    height, width = shape
    mask = np.zeros(shape, dtype=bool)
    e = _edges(rings, transform)
    if not len(e):
        return mask
    c0, r0, c1, r1 = e.T
    lo = np.floor(np.minimum(r0, r1) - 0.5).astype(np.int64)
    hi = np.ceil(np.maximum(r0, r1) - 0.5).astype(np.int64)
    row_lo = max(int(lo.min()), 0)
    row_hi = min(int(hi.max()), height - 1)
    cols = np.arange(width) + 0.5
    for row in range(row_lo, row_hi + 1):
        yc = row + 0.5
        hit = ((r0 <= yc) & (yc < r1)) | ((r1 <= yc) & (yc < r0))
        if not hit.any():
            continue
        t = (yc - r0[hit]) / (r1[hit] - r0[hit])
        xs = np.sort(c0[hit] + t * (c1[hit] - c0[hit]))
        a, b = xs[0::2], xs[1::2]
        inside = ((cols[None, :] >= a[:, None]) & (cols[None, :] < b[:, None])).any(axis=0)
        mask[row] |= inside
    return mask


def geometry_mask(polygons: Iterable[Sequence[Ring]], shape, transform) -> np.ndarray:
    """Union of the polygons' cell masks (rasterio ``geometry_mask(..., invert=True)``)."""
    mask = np.zeros(shape, dtype=bool)
    for poly in polygons:
        mask |= polygon_cells(poly, shape, transform)
    return mask


def rasterize(shapes: Iterable[tuple[Sequence[Ring], float]], shape, transform, *,
              fill: float = np.nan, dtype=np.float32) -> np.ndarray:
    """Burn each polygon's value into its cells, later shapes on top (rasterio ``rasterize``)."""
    out = np.full(shape, fill, dtype=dtype)
    for rings, value in shapes:
        out[polygon_cells(rings, shape, transform)] = value
    return out


def _tent_weights(src_n: int, src_origin: float, src_res: float, dst_centres: np.ndarray,
                  dst_res: float) -> tuple[np.ndarray, np.ndarray]:
    """Sparse (dst, taps) source indices and tent weights along one axis."""
    scale = min(1.0, abs(src_res) / abs(dst_res))  # < 1 when downsampling
    radius = 1.0 / scale
    pos = (dst_centres - src_origin) / src_res - 0.5  # fractional source index of the centre
    taps = int(np.ceil(radius)) * 2 + 1
    base = np.floor(pos).astype(np.int64) - taps // 2 + 1
    idx = base[:, None] + np.arange(taps)[None, :]
    w = np.maximum(0.0, 1.0 - np.abs(idx - pos[:, None]) * scale)
    outside = (idx < 0) | (idx >= src_n)
    w[outside] = 0.0
    return np.clip(idx, 0, src_n - 1), w


def reproject_same_crs(src: np.ndarray, src_transform: NorthUpTransform, src_nodata: float | None,
                       dst_shape: tuple[int, int], dst_transform: NorthUpTransform, *,
                       resampling: str = "bilinear") -> np.ndarray:
    """Resample ``src`` onto the destination grid (same CRS); NaN where no source cell counts."""
    if not (src_transform.axis_aligned and dst_transform.axis_aligned):
        raise ValueError("Rotated grids are not supported.")
    height, width = dst_shape
    data = np.asarray(src, np.float64)
    valid = np.isfinite(data)
    if src_nodata is not None and np.isfinite(src_nodata):
        valid &= data != src_nodata
    xs = dst_transform.c + (np.arange(width) + 0.5) * dst_transform.a
    ys = dst_transform.f + (np.arange(height) + 0.5) * dst_transform.e
    sh, sw = data.shape
    if resampling == "nearest":
        ci = np.floor((xs - src_transform.c) / src_transform.a).astype(np.int64)
        ri = np.floor((ys - src_transform.f) / src_transform.e).astype(np.int64)
        okc = (ci >= 0) & (ci < sw)
        okr = (ri >= 0) & (ri < sh)
        out = np.full(dst_shape, np.nan, np.float64)
        sub = data[np.ix_(ri[okr], ci[okc])]
        vsub = valid[np.ix_(ri[okr], ci[okc])]
        out[np.ix_(okr, okc)] = np.where(vsub, sub, np.nan)
        return out.astype(np.float32)
    if resampling != "bilinear":
        raise ValueError(f"Unsupported resampling {resampling!r}.")
    cx, wx = _tent_weights(sw, src_transform.c, src_transform.a, xs, dst_transform.a)
    ry, wy = _tent_weights(sh, src_transform.f, src_transform.e, ys, dst_transform.e)
    vals = np.where(valid, data, 0.0)
    vmask = valid.astype(np.float64)
    # Separable: along x for every source row, then along y.
    tx_v = np.einsum("rdt,dt->rd", vals[:, cx], wx)
    tx_m = np.einsum("rdt,dt->rd", vmask[:, cx], wx)
    num = np.einsum("yt,ytc->yc", wy, tx_v[ry])
    den = np.einsum("yt,ytc->yc", wy, tx_m[ry])
    with np.errstate(invalid="ignore", divide="ignore"):
        out = np.where(den > 0, num / np.where(den > 0, den, 1.0), np.nan)
    # Destination centres outside the source extent get no value.
    sx = (xs - src_transform.c) / src_transform.a
    sy = (ys - src_transform.f) / src_transform.e
    out[~((sy >= 0) & (sy < sh)), :] = np.nan
    out[:, ~((sx >= 0) & (sx < sw))] = np.nan
    return out.astype(np.float32)


def fill_from_nearest(values: np.ndarray, missing: np.ndarray) -> np.ndarray:
    """Give each ``missing`` cell the value of the nearest finite cell (Euclidean distance
    transform, shallow_water_2d.py:128-143)."""
    from scipy.ndimage import distance_transform_edt

    out = np.array(values, copy=True)
    if not missing.any():
        return out
    finite = np.isfinite(out)
    if not finite.any():
        raise ValueError("DEM resampling produced no finite elevations.")
    idx = distance_transform_edt(~finite, return_distances=False, return_indices=True)
    out[missing] = out[idx[0][missing], idx[1][missing]]
    return out


# ---------------------------------------------------------------------------
# Reference entry points
# ---------------------------------------------------------------------------
def domain_polygons(layer: VectorLayer) -> list[list[Ring]]:
    """Polygons of the domain layer: polygons and closed perimeter lines
    (dem/synthetic_terrain.py:37-54; overlapping parts are merged by the mask union)."""
    polys = [p for p in layer.polygons if p]
    if not polys:
        raise ValueError("Domain must contain polygon geometry or a closed line perimeter.")
    return polys


def grid_from_rasters(north_mask: np.ndarray, north_bed: np.ndarray, north_nearest: np.ndarray,
                      north_roughness: np.ndarray, x: np.ndarray, y: np.ndarray,
                      resolution_m: float, *, crs: Any = None,
                      transform: NorthUpTransform | None = None) -> StructuredGrid:
    """Array half of ``load_structured_grid`` (shallow_water_2d.py:124-181): fill DEM holes,
    check coverage, flip rows south-to-north, zero the fields outside the domain."""
    north_mask = np.asarray(north_mask, bool)
    bed = np.array(north_bed, dtype=np.float32, copy=True)
    missing = north_mask & ~np.isfinite(bed)
    bed[missing] = np.asarray(north_nearest, np.float32)[missing]
    missing = north_mask & ~np.isfinite(bed)
    if missing.any():
        bed = fill_from_nearest(bed, missing)
    if not np.isfinite(bed[north_mask]).all():
        raise ValueError("DEM resampling left non-finite cells inside the domain.")
    rough = np.asarray(north_roughness, np.float32)
    if not np.isfinite(rough[north_mask]).all():
        n_missing = int((north_mask & ~np.isfinite(rough)).sum())
        raise ValueError(f"Roughness rasterization left {n_missing} domain cells unassigned.")
    bed = np.flipud(bed).copy()
    manning = np.flipud(rough).copy()
    mask = np.flipud(north_mask).copy()
    bed[~mask] = 0.0
    manning[~mask] = 0.0
    return StructuredGrid(bed=bed, manning_n=manning, domain_mask=mask,
                          x=np.asarray(x, np.float64), y=np.asarray(y, np.float64),
                          resolution_m=float(resolution_m), crs=crs, north_up_transform=transform)


def load_structured_grid(dem_path: str | Path, roughness_path: str | Path,
                         domain_path: str | Path, *, resolution_m: float = 50.0) -> StructuredGrid:
    """Resample DEM, LULC Manning n and the domain mask onto one aligned grid
    (jax_solver/shallow_water_2d.py:72-181), from a GeoTIFF and two GeoPackages."""
    if resolution_m <= 0:
        raise ValueError("Structured-grid resolution must be positive.")
    domain = read_geopackage(domain_path)
    roughness = read_geopackage(roughness_path)
    if domain.crs is None or domain.projected is False:
        raise ValueError("Domain must use a projected CRS.")
    if roughness.crs is not None and roughness.crs != domain.crs:
        # the reference moves the layer with to_crs (shallow_water_2d.py:86)
        if not (projections.is_supported(roughness.crs) and projections.is_supported(domain.crs)):
            raise ValueError(f"Roughness CRS {roughness.crs} differs from the domain CRS "
                             f"{domain.crs}; reprojection between them is not available here "
                             "(EPSG:4326, EPSG:3857 and the WGS 84 UTM zones are).")
        roughness = VectorLayer(
            projections.transform_polygons(roughness.polygons, roughness.crs, domain.crs),
            roughness.attributes, domain.crs, domain.projected, roughness.kinds)
    if "roughness" not in roughness.attributes:
        raise ValueError("LULC roughness layer must contain the 'roughness' attribute.")
    polys = domain_polygons(domain)
    x, y, transform, shape = aligned_grid_geometry(_bounds(polys), resolution_m)
    north_mask = geometry_mask(polys, shape, transform)
    dem = read_geotiff(dem_path)
    if dem.crs is not None and dem.crs != domain.crs:
        raise ValueError(f"DEM CRS {dem.crs} does not match domain CRS {domain.crs}.")
    north_bed = reproject_same_crs(dem.data, dem.transform, dem.nodata, shape, transform,
                                   resampling="bilinear")
    north_nearest = reproject_same_crs(dem.data, dem.transform, dem.nodata, shape, transform,
                                       resampling="nearest")
    shapes = [(rings, float(v)) for rings, v in zip(roughness.polygons,
                                                   roughness.attributes["roughness"], strict=True)
              if rings and v is not None and np.isfinite(float(v))]
    north_rough = rasterize(shapes, shape, transform, fill=np.nan, dtype=np.float32)
    return grid_from_rasters(north_mask, north_bed, north_nearest, north_rough, x, y,
                             resolution_m, crs=domain.crs, transform=transform)


# Columns of the rainfall table (jax_solver/shallow_water_2d.py:184-207): the
# loader accepts a table only when it carries all of them.
RAINFALL_COLUMNS = ("timestamp", "scenario", "rainfall_mm_hr", "is_missing_timestamp", "units")


def load_rainfall_series(rainfall_path: str | Path, *, scenario: str,
                         default_rate_mm_hr: float = 0.0) -> tuple[np.ndarray, np.ndarray, float]:
    """Explicit mm/h rainfall of one scenario as (elapsed s, rates m/s, default m/s)
    (jax_solver/shallow_water_2d.py:184-207).

    The Parquet table must hold RAINFALL_COLUMNS, every non-null ``units``
    entry must be "mm/hr", the scenario must have rows, and none of them may
    be flagged missing or lack an intensity.  The scenario's rows are put in
    time order; elapsed seconds count from the first one.  Times and rates
    are float32, the default rate a Python float."""
    import pyarrow.parquet as pq

    absent = sorted(set(RAINFALL_COLUMNS) - set(pq.read_schema(rainfall_path).names))
    if absent:
        raise ValueError(f"Rainfall table is missing columns: {absent}")
    table = pq.read_table(rainfall_path, columns=list(RAINFALL_COLUMNS))
    column = {name: table.column(name).to_numpy(zero_copy_only=False) for name in RAINFALL_COLUMNS}
    found = sorted({str(u) for u in column["units"] if u is not None})
    if found != ["mm/hr"]:
        raise ValueError(f"Expected rainfall units 'mm/hr', found {found}")
    rows = np.flatnonzero(column["scenario"] == scenario)
    if rows.size == 0:
        raise ValueError(f"Rainfall scenario {scenario!r} is unavailable.")
    intensity = np.asarray(column["rainfall_mm_hr"][rows], dtype=np.float64)
    flagged = np.asarray(column["is_missing_timestamp"][rows], dtype=bool)
    if flagged.any() or np.isnan(intensity).any():
        raise ValueError("Rainfall contains missing timestamps or intensities.")
    stamps = np.asarray(column["timestamp"][rows], dtype="datetime64[ns]")
    order = np.argsort(stamps, kind="stable")
    stamps, intensity = stamps[order], intensity[order]
    seconds = (stamps - stamps[0]) / np.timedelta64(1, "s")
    rates = intensity.astype(np.float32) * np.float32(MM_HR_TO_M_S)
    return seconds.astype(np.float32), rates, default_rate_mm_hr * MM_HR_TO_M_S
