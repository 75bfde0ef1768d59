"""Feature staging for the CNN on the GPU (comparison/v2_staging.py).

The reference builds the NHWC feature batch in one compiled JAX stage and
hands it to PyTorch through DLPack, then hands the predictions back the same
way for the depth correction (v2_staging.py:17-94).  Here the states already
live in torch tensors on the device, so both exchanges disappear: one kernel
writes the standardised features straight into a channels-last buffer that a
PyTorch convolution reads as NCHW, and one kernel applies
``max(depth + residual, 0)`` reading the NCHW predictions in place.
"""

from __future__ import annotations

from time import perf_counter
from typing import Any

import numpy as np
import torch

from . import _lib
from .coupled_forecast import _f32x3


def _dev_f32(x, dev) -> torch.Tensor:
    return torch.as_tensor(np.asarray(x, np.float32) if not isinstance(x, torch.Tensor) else x,
                           dtype=torch.float32, device=dev).contiguous()


def stage_features(state_t, rainfall_t_mm_hr, terrain_standardized, roughness_standardized,
                   domain_mask, state_mean, state_std, rainfall_scale_mm_hr) -> torch.Tensor:
    """NCHW (count, 6, ny, nx) channels-last view of the standardised features.

    ``state_t`` is the (count, ny, nx, 3) primitive state (depth, u, v);
    channels follow v2_staging.py:17-44: standardised and masked state, the
    sample's rain over ``rainfall_scale_mm_hr``, terrain, roughness.
    """
    st = state_t if isinstance(state_t, torch.Tensor) and state_t.is_cuda else None
    dev = st.device if st is not None else torch.device("cuda", torch.cuda.current_device())
    st = _dev_f32(state_t, dev)
    if st.dim() != 4 or st.shape[-1] != 3:
        raise ValueError("state_t must be (count, ny, nx, 3)")
    count, ny, nx, _ = st.shape
    terrain = _dev_f32(terrain_standardized, dev)
    rough = _dev_f32(roughness_standardized, dev)
    mask = torch.as_tensor(np.asarray(domain_mask) if not isinstance(domain_mask, torch.Tensor)
                           else domain_mask, device=dev).to(torch.uint8).contiguous()
    for f in (terrain, rough, mask):
        if tuple(f.shape) != (ny, nx):
            raise ValueError("static fields must be (ny, nx)")
    rain = np.asarray(rainfall_t_mm_hr.detach().cpu().numpy() if isinstance(rainfall_t_mm_hr, torch.Tensor)
                      else rainfall_t_mm_hr, np.float32).reshape(-1)
    if rain.shape[0] != count:
        raise ValueError("one rainfall value per sample is required")
    scale = np.float32(rainfall_scale_mm_hr.item() if isinstance(rainfall_scale_mm_hr, torch.Tensor)
                       else rainfall_scale_mm_hr)
    rain_ch = torch.as_tensor((rain / scale).astype(np.float32), device=dev)
    out = torch.empty((count, ny, nx, 6), dtype=torch.float32, device=dev)
    N = ny * nx
    _lib.check(
        _lib.load().swe_stage_features(
            count, ny, nx, st.data_ptr(), mask.data_ptr(), terrain.data_ptr(), rough.data_ptr(),
            _f32x3(state_mean), _f32x3(state_std), rain_ch.data_ptr(), out.data_ptr(), 6 * N, 1, 6,
            _lib.stream_handle(dev)),
        "swe_stage_features")
    return out.permute(0, 3, 1, 2)


def correct_depth(raw_depth, predictions_nchw: torch.Tensor) -> torch.Tensor:
    """``max(raw_depth + predictions[:, 0], 0)`` per sample (v2_staging.py:47-52)."""
    pred = predictions_nchw
    if pred.dim() != 4 or pred.shape[1] < 1 or pred.dtype != torch.float32:
        raise ValueError("predictions must be float32 (count, C, ny, nx)")
    count, _, ny, nx = pred.shape
    depth = _dev_f32(raw_depth, pred.device)
    if tuple(depth.shape) != (count, ny, nx):
        raise ValueError("raw_depth must be (count, ny, nx)")
    s0, _, s2, s3 = pred.stride()
    if s2 != nx * s3:
        raise ValueError("prediction rows must be evenly strided")
    out = torch.empty_like(depth)
    _lib.check(
        _lib.load().swe_correct_depth(count, ny, nx, depth.data_ptr(), pred.data_ptr(), s0, s3,
                                      out.data_ptr(), _lib.stream_handle(pred.device)),
        "swe_correct_depth")
    return out


def v2_features_to_torch(dataset: Any) -> tuple[torch.Tensor, float, float]:
    """Reference-shaped entry (v2_staging.py:55-77): features of every sample.

    ``dataset`` needs the ResidualDataset fields ``state_t``,
    ``rainfall_t_mm_hr``, ``terrain_standardized``, ``roughness_standardized``,
    ``domain_mask`` and ``metadata["normalization"]``.  Returns (NCHW view,
    0.0 compile seconds -- nothing is compiled at run time --, execution s).
    """
    norm = dataset.metadata["normalization"]
    t0 = perf_counter()
    feats = stage_features(dataset.state_t, dataset.rainfall_t_mm_hr, dataset.terrain_standardized,
                           dataset.roughness_standardized, dataset.domain_mask, norm["state_mean"],
                           norm["state_std"], norm["rainfall_scale_mm_hr"])
    torch.cuda.synchronize(feats.device)
    return feats, 0.0, perf_counter() - t0


def corrections_to_torch(raw_depth, predictions_nchw: torch.Tensor) -> tuple[torch.Tensor, float, float]:
    """Reference-shaped entry (v2_staging.py:80-94) returning a device tensor."""
    t0 = perf_counter()
    out = correct_depth(raw_depth, predictions_nchw)
    torch.cuda.synchronize(out.device)
    return out, 0.0, perf_counter() - t0
