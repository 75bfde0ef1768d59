"""CPU oracle for the conserved -> primitive output schema.

TEST INFRASTRUCTURE ONLY: the checker for ``hybrid_flood_amd.structured_io``;
only ``tests/`` imports it.

Restates jax_solver/shallow_water_2d.py:309-350 (and the identical
structured_io.py:160-211) in NumPy float32: ``np.divide(q, h, where=h > dry)``
into zeros, ``np.hypot`` for the speed, NaN outside the domain mask for the
four time-varying fields and for the elevation.  Pinned in
``tests/test_oracle_output.py`` by closed forms (3-4-5 speeds, dry cells at
exactly zero velocity, NaN exactly where the mask is false).
"""

from __future__ import annotations

import numpy as np


def conserved_to_primitive(h, hu, hv, bed, domain_mask, dry_tolerance):
    depth = np.array(h, dtype=np.float32, copy=True)
    hu = np.asarray(hu, np.float32)
    hv = np.asarray(hv, np.float32)
    mask = np.asarray(domain_mask, bool)
    wet = depth > np.float32(dry_tolerance)
    vx = np.zeros_like(depth)
    vy = np.zeros_like(depth)
    np.divide(hu, depth, out=vx, where=wet)
    np.divide(hv, depth, out=vy, where=wet)
    speed = np.hypot(vx, vy).astype(np.float32)
    for values in (depth, vx, vy, speed):
        values[..., ~mask] = np.nan
    elevation = np.asarray(bed, np.float32).copy()
    elevation[~mask] = np.nan
    return {"elevation": elevation, "depth": depth, "x_velocity": vx, "y_velocity": vy,
            "velocity": speed}
