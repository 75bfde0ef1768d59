"""Boundary codes and the ghost-cell rule (jax_solver/boundary_conditions.py:9-52).

The fused HIP step applies the ghost layer inside the stencil (it never
materialises padded arrays); ``apply_ghost_cells`` is kept as a tensor utility
with the reference's semantics for callers and tests that use it directly.
"""

from __future__ import annotations

from typing import NamedTuple

import torch

REFLECTIVE = 0
OPEN = 1


class BoundaryConfig(NamedTuple):
    """Boundary codes ordered as left, right, bottom, top."""

    types: torch.Tensor


def all_reflective() -> BoundaryConfig:
    return BoundaryConfig(torch.full((4,), REFLECTIVE, dtype=torch.int32))


def all_open() -> BoundaryConfig:
    return BoundaryConfig(torch.full((4,), OPEN, dtype=torch.int32))


def apply_ghost_cells(h, hu, hv, boundary_types):
    """Pad one edge-copied ghost layer; reflective sides negate the normal momentum."""
    h, hu, hv = (torch.as_tensor(x) for x in (h, hu, hv))
    types = [int(b) for b in torch.as_tensor(boundary_types).reshape(-1).tolist()]
    pad = torch.nn.functional.pad

    def edge(x):
        return pad(x[None, None], (1, 1, 1, 1), mode="replicate")[0, 0]

    hg, hug, hvg = edge(h), edge(hu), edge(hv)
    if types[0] == REFLECTIVE:
        hug[:, 0] = -hug[:, 1]
    if types[1] == REFLECTIVE:
        hug[:, -1] = -hug[:, -2]
    if types[2] == REFLECTIVE:
        hvg[0, :] = -hvg[1, :]
    if types[3] == REFLECTIVE:
        hvg[-1, :] = -hvg[-2, :]
    return hg, hug, hvg
