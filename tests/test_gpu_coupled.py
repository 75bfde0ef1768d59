"""GPU fixed-length advance (hybrid/coupled_forecast.py:123-161) against the
oracle's, and against the device loop of integrate_to_outputs.

Bars: with n = 0 the reference-order kernel reproduces the FP32 oracle bit
for bit (state, step count, clock); either kernel equals the final slot of
integrate_to_outputs over [start, target] bit for bit (same device loop);
the fast kernel vs the FP64 oracle |x - x64| <= 2e-7 + 2e-5 |x64| after a
short advance.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import random_case
from test_gpu_step import _oparams, _params, _state

pytestmark = pytest.mark.gpu

RTOL, ATOL = 2.0e-5, 2.0e-7


@pytest.mark.parametrize("shape", [(12, 14), (40, 130)])
def test_frictionless_advance_bit_exact(cuda, shape):
    from hybrid_flood_amd.coupled_forecast import advance_physics_to_time
    from oracle import swe_oracle as o

    case = random_case(21, shape, open_sides=(1, 0, 0, 1), speed=1.0)
    case["manning_n"] = np.zeros(shape, np.float32)
    params = _params(case, cuda)
    out, steps = advance_physics_to_time(_state(case, cuda), params, 2.0, 3.25, max_substeps=400,
                                         variant="reference")
    ref, rsteps, _ = o.advance_physics_to_time(o.OracleState(case["h"], case["hu"], case["hv"]),
                                               _oparams(case), 2.0, 3.25, max_substeps=400)
    assert steps == rsteps > 0
    for g, r in zip(out, ref):
        assert np.array_equal(g.cpu().numpy(), r)


@pytest.mark.parametrize("variant", ["reference", "fast"])
def test_advance_equals_integrate_final_slot(cuda, variant):
    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd.coupled_forecast import advance_physics_to_time

    case = random_case(5, (33, 70), open_sides=(0, 1, 0, 0), speed=0.8)
    params = _params(case, cuda)
    times = np.array([10.0, 12.5], np.float32)
    res = nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=1000, variant=variant)
    st = _state(case, cuda)
    out, steps = advance_physics_to_time(st, params, 10.0, 12.5, max_substeps=1000, variant=variant,
                                         out=st)  # in place
    assert steps == res.steps_used
    for g, r in zip(out, res.states):
        assert np.array_equal(g.cpu().numpy(), r[-1].cpu().numpy())


def test_fast_advance_against_fp64_oracle(cuda):
    from hybrid_flood_amd.coupled_forecast import advance_physics_to_time
    from oracle import swe_oracle as o

    case = random_case(8, (24, 31), speed=0.5)
    params = _params(case, cuda)
    ref, rsteps, _ = o.advance_physics_to_time(o.OracleState(case["h"], case["hu"], case["hv"]),
                                               _oparams(case, np.float64), 0.0, 0.2, max_substeps=64)
    out, steps = advance_physics_to_time(_state(case, cuda), params, 0.0, 0.2, max_substeps=64)
    assert steps == rsteps
    for g, r in zip(out, ref):
        np.testing.assert_allclose(g.cpu().numpy().astype(np.float64), r, rtol=RTOL, atol=ATOL)


def test_unreached_target_gives_nan(cuda):
    from hybrid_flood_amd.coupled_forecast import advance_physics_to_time
    from oracle import swe_oracle as o

    case = random_case(4, (20, 20), speed=1.0)
    params = _params(case, cuda)
    out, steps = advance_physics_to_time(_state(case, cuda), params, 0.0, 50.0, max_substeps=3,
                                         variant="reference")
    _, rsteps, _ = o.advance_physics_to_time(o.OracleState(case["h"], case["hu"], case["hv"]),
                                             _oparams(case), 0.0, 50.0, max_substeps=3)
    assert steps == rsteps == 3
    assert all(bool(x.isnan().all()) for x in out)


def test_target_before_start_is_identity(cuda):
    from hybrid_flood_amd.coupled_forecast import advance_physics_to_time

    case = random_case(6, (17, 23))
    params = _params(case, cuda)
    st = _state(case, cuda)
    out, steps = advance_physics_to_time(st, params, 5.0, 4.0, max_substeps=10)
    assert steps == 0
    for g, r in zip(out, st):
        assert np.array_equal(g.cpu().numpy(), r.cpu().numpy())


# ---------------------------------------------------------------------------
# Features, residual correction, coupled loop (coupled_forecast.py:53-231)
# ---------------------------------------------------------------------------

def _hybrid_inputs(case, dev, seed=3):
    import torch

    from hybrid_flood_amd.coupled_forecast import HybridInputs

    rng = np.random.default_rng(seed)
    shape = case["bed"].shape
    terrain = rng.normal(0, 1, shape).astype(np.float32)
    rough = rng.normal(0, 1, shape).astype(np.float32)
    return HybridInputs(torch.as_tensor(terrain, device=dev), torch.as_tensor(rough, device=dev),
                        [0.1, 0.02, -0.01], [0.2, 0.3, 0.25], 12.0, [0.01, 0.05, 0.04]), terrain, rough


@pytest.mark.parametrize("shape", [(21, 37), (24, 44)])  # one-cell; four-cell NHWC form
@pytest.mark.parametrize("layout", ["nhwc", "nchw"])
def test_features_bit_exact(cuda, layout, shape):
    from hybrid_flood_amd.coupled_forecast import network_features
    from oracle import coupled_oracle as co
    from oracle import swe_oracle as o

    case = random_case(12, shape)
    params = _params(case, cuda)
    hi, terrain, rough = _hybrid_inputs(case, cuda)
    got = network_features(_state(case, cuda), params, hi, 75.0, layout=layout).cpu().numpy()
    if layout == "nchw":
        got = got.transpose(0, 2, 3, 1)
    want = co.network_features(o.OracleState(case["h"], case["hu"], case["hv"]), _oparams(case),
                               terrain, rough, hi.state_mean, hi.state_std, 12.0, 75.0)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("inplace", [False, True])
def test_residual_correction_matches_oracle(cuda, inplace):
    import torch

    from hybrid_flood_amd.coupled_forecast import apply_residual_correction
    from oracle import coupled_oracle as co
    from oracle import swe_oracle as o

    case = random_case(13, (40, 52))
    params = _params(case, cuda)
    rng = np.random.default_rng(1)
    r_nchw = rng.normal(0, 3.0, (1, 3, 40, 52)).astype(np.float32)
    lim = np.array([1.0, 2.0, 0.5], np.float32)
    pred = _state(case, cuda)
    residual = torch.as_tensor(r_nchw, device=cuda)[0].permute(1, 2, 0)  # strided (H, W, C) view
    out, vols = apply_residual_correction(pred, residual, params, correction_relaxation=0.3,
                                          clip_limits=lim, out=pred if inplace else None,
                                          volumes=True)
    want, wvols = co.apply_residual_correction(o.OracleState(case["h"], case["hu"], case["hv"]),
                                               r_nchw[0].transpose(1, 2, 0), _oparams(case),
                                               correction_relaxation=0.3, clip_limits=lim)
    np.testing.assert_allclose(vols, wvols, rtol=1e-12)
    # the float64 sums differ only in order: at most one float32 rounding step apart
    for g, w in zip(out, want):
        np.testing.assert_allclose(g.cpu().numpy(), w, rtol=2.4e-7, atol=0)
    assert float(out.h.double().sum()) == pytest.approx(wvols[0], rel=1e-6)


@pytest.mark.parametrize("shape,layout", [((40, 52), "nhwc"), ((40, 52), "offset"),
                                          ((41, 53), "nhwc"), ((41, 53), "planar")])
def test_residual_correction_four_cell_and_one_cell_forms(cuda, shape, layout):
    """Interleaved NHWC and planar residuals on a cell count that is a multiple
    of 4 take the four-cell apply kernel; an odd count or a residual one float
    off alignment takes the one-cell kernel; every form matches the oracle."""
    import torch

    from hybrid_flood_amd.coupled_forecast import apply_residual_correction
    from oracle import coupled_oracle as co
    from oracle import swe_oracle as o

    case = random_case(19, shape)
    params = _params(case, cuda)
    r_hwc = np.random.default_rng(2).normal(0, 3.0, shape + (3,)).astype(np.float32)
    lim = np.array([1.5, 0.7, 2.0], np.float32)
    if layout == "planar":
        residual = torch.as_tensor(r_hwc.transpose(2, 0, 1).copy(), device=cuda).permute(1, 2, 0)
    elif layout == "offset":
        buf = torch.empty(r_hwc.size + 1, dtype=torch.float32, device=cuda)
        buf[1:] = torch.as_tensor(r_hwc.ravel(), device=cuda)
        residual = buf[1:].view(r_hwc.shape)
    else:
        residual = torch.as_tensor(r_hwc, device=cuda)
    out, vols = apply_residual_correction(_state(case, cuda), residual, params,
                                          correction_relaxation=0.4, clip_limits=lim, volumes=True)
    want, wvols = co.apply_residual_correction(o.OracleState(case["h"], case["hu"], case["hv"]),
                                               r_hwc, _oparams(case), correction_relaxation=0.4,
                                               clip_limits=lim)
    np.testing.assert_allclose(vols, wvols, rtol=1e-12)
    for g, w in zip(out, want):
        np.testing.assert_allclose(g.cpu().numpy(), w, rtol=2.4e-7, atol=0)


def test_coupled_forecast_matches_oracle(cuda):
    import torch

    from hybrid_flood_amd.coupled_forecast import run_coupled_forecast
    from oracle import coupled_oracle as co
    from oracle import swe_oracle as o

    case = random_case(14, (24, 30), speed=0.5)
    case["manning_n"] = np.zeros((24, 30), np.float32)
    params = _params(case, cuda)
    hi, terrain, rough = _hybrid_inputs(case, cuda)

    def model_t(f):  # (1, 6, ny, nx) -> (1, 3, ny, nx); exact IEEE products and sums only
        return 0.02 * (f[:, 0:3] - 0.5 * f[:, 3:6])

    def model_np(f):  # (1, ny, nx, 6) -> (1, ny, nx, 3)
        return np.float32(0.02) * (f[..., 0:3] - np.float32(0.5) * f[..., 3:6])

    times = np.array([0.0, 0.4, 0.8, 1.2], np.float32)
    res = run_coupled_forecast(_state(case, cuda), params, model_t, hi, times,
                               max_substeps_per_interval=64, correction_relaxation=0.5,
                               correction_clip_sigma=2.0, variant="reference")
    ref, _, steps = co.run_coupled_forecast(
        o.OracleState(case["h"], case["hu"], case["hv"]), _oparams(case), model_np, terrain, rough,
        hi.state_mean, hi.state_std, 12.0, hi.residual_std, times, max_substeps_per_interval=64,
        correction_relaxation=0.5, correction_clip_sigma=2.0)
    assert res.physics_steps_used == steps
    for g, w in zip(res.states, ref):
        np.testing.assert_allclose(g.cpu().numpy(), w, rtol=1e-5, atol=1e-7)
    assert torch.equal(res.output_times_s, torch.as_tensor(times))


# ---------------------------------------------------------------------------
# Output schema (jax_solver/shallow_water_2d.py:309-350)
# ---------------------------------------------------------------------------

def test_result_to_dataset_matches_oracle(cuda, tmp_path):
    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd.structured_io import load_npz, result_to_dataset
    from oracle import output_oracle as oo

    case = random_case(15, (26, 33), speed=1.0)
    params = _params(case, cuda)
    res = nm.integrate_to_outputs(_state(case, cuda), params, np.array([0.0, 0.2, 0.4], np.float32),
                                  max_steps=100)
    ds = result_to_dataset(res, params)
    want = oo.conserved_to_primitive(*(s.cpu().numpy() for s in res.states), case["bed"],
                                     case["domain_mask"], np.float32(case["dry"]))
    for name in ("elevation", "depth", "x_velocity", "y_velocity"):
        assert np.array_equal(ds[name].cpu().numpy(), want[name], equal_nan=True), name
    # hypot: the device and the host library may differ in the last bit
    np.testing.assert_allclose(ds["velocity"].cpu().numpy(), want["velocity"], rtol=1.2e-7,
                               equal_nan=True)
    assert ds.attrs["steps_used"] == res.steps_used
    np.testing.assert_array_equal(ds.coords["time"], [0.0, np.float32(0.2), np.float32(0.4)])
    back = load_npz(ds.to_npz(tmp_path / "out.npz"))
    assert np.array_equal(back["depth"], want["depth"], equal_nan=True)
    assert back["__meta__"]["var_attrs"]["velocity"]["units"] == "m s-1"
    from hybrid_flood_amd.structured_io import load_netcdf, save_netcdf

    nc = load_netcdf(save_netcdf(ds, tmp_path / "sub" / "out.nc"))
    for name in ("elevation", "depth", "x_velocity", "y_velocity"):
        assert np.array_equal(nc[name], want[name], equal_nan=True), name
    assert nc["__dims__"]["depth"] == ("time", "y", "x")
    assert nc["__attrs__"]["steps_used"] == res.steps_used


@pytest.mark.parametrize("shape,offset", [((24, 36), 0), ((24, 36), 1), ((25, 37), 0), ((64, 4), 0)])
def test_conserved_to_dataset_vector_and_scalar_forms(cuda, shape, offset):
    """Four-cell form (cell count a multiple of 4, aligned arrays) and the
    per-cell form (odd counts, or inputs one float off alignment) give the
    oracle's schema exactly, across several output times."""
    import torch

    from hybrid_flood_amd.structured_io import conserved_to_dataset
    from oracle import output_oracle as oo

    case = random_case(16, shape, speed=1.0)
    params = _params(case, cuda)
    rng = np.random.default_rng(17)
    T = 3
    h, hu, hv = (rng.normal(0.0, 1.0, (3, T) + shape).astype(np.float32) *
                 np.array([0.05, 0.2, 0.2], np.float32)[:, None, None, None])
    h = np.abs(h)
    h[:, ::3, ::2] = 0.5 * np.float32(case["dry"])  # dry cells with momentum: zero velocity

    def dev(a):
        buf = torch.empty(a.size + offset, dtype=torch.float32, device=cuda)
        buf[offset:] = torch.as_tensor(a.ravel(), device=cuda)
        return buf[offset:].view(a.shape)

    ds = conserved_to_dataset(dev(h), dev(hu), dev(hv), np.arange(T, dtype=np.float32), params,
                              dry_tolerance_m=float(case["dry"]), title="t", backend="hip",
                              steps_used=0)
    want = oo.conserved_to_primitive(h, hu, hv, case["bed"], case["domain_mask"],
                                     np.float32(case["dry"]))
    for name in ("elevation", "depth", "x_velocity", "y_velocity"):
        assert np.array_equal(ds[name].cpu().numpy(), want[name], equal_nan=True), name
    np.testing.assert_allclose(ds["velocity"].cpu().numpy(), want["velocity"], rtol=1.2e-7,
                               equal_nan=True)
    assert not np.isnan(ds["depth"].cpu().numpy()[:, case["domain_mask"]]).any()


# ---------------------------------------------------------------------------
# Feature staging (comparison/v2_staging.py)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("ny,nx", [(19, 45), (20, 44)])  # one-cell; four-cell NHWC form
def test_stage_features_and_correct_depth_bit_exact(cuda, ny, nx):
    import torch

    from hybrid_flood_amd.staging import correct_depth, stage_features
    from oracle import coupled_oracle as co

    rng = np.random.default_rng(9)
    count = 4
    st = rng.normal(0, 1, (count, ny, nx, 3)).astype(np.float32)
    mask = rng.uniform(size=(ny, nx)) > 0.2
    terrain, rough = rng.normal(size=(2, ny, nx)).astype(np.float32)
    rain = rng.uniform(0, 40, count).astype(np.float32)
    got = stage_features(torch.as_tensor(st, device=cuda), rain, terrain, rough, mask,
                         [0.1, -0.2, 0.05], [0.3, 0.7, 0.9], 25.0)
    assert tuple(got.shape) == (count, 6, ny, nx)
    want = co.stage_features(st, rain, terrain, rough, mask, [0.1, -0.2, 0.05], [0.3, 0.7, 0.9], 25.0)
    assert np.array_equal(got.permute(0, 2, 3, 1).cpu().numpy(), want)
    pred = torch.as_tensor(rng.normal(0, 0.5, (count, 3, ny, nx)).astype(np.float32), device=cuda)
    depth = np.abs(rng.normal(0, 0.4, (count, ny, nx))).astype(np.float32)
    out = correct_depth(depth, pred)
    assert np.array_equal(out.cpu().numpy(),
                          co.correct_depth(depth, pred.permute(0, 2, 3, 1).cpu().numpy()))


def _flat_case(ny, nx, seed):
    """Random state and fields without the terrain generator (fast at 1M cells)."""
    rng = np.random.default_rng(seed)
    h = np.abs(rng.normal(0, 0.3, (ny, nx))).astype(np.float32)
    h[rng.uniform(size=(ny, nx)) < 0.1] = 0.0
    return dict(bed=rng.normal(0, 1, (ny, nx)).astype(np.float32),
                manning_n=np.full((ny, nx), 0.03, np.float32),
                domain_mask=rng.uniform(size=(ny, nx)) > 0.1,
                rainfall_multiplier=None, boundary_types=(0, 0, 0, 0),
                rain_t=np.array([0.0, 100.0], np.float32), rain_r=np.array([1e-5, 2e-5], np.float32),
                dx=2.0, dry=1e-4, h=h, hu=(h * rng.normal(0, 1, (ny, nx))).astype(np.float32),
                hv=(h * rng.normal(0, 1, (ny, nx))).astype(np.float32))


@pytest.mark.parametrize("layout", ["nhwc", "nchw"])
def test_features_bit_exact_past_one_grid_sweep(cuda, layout):
    """An odd-sized grid above 4096 x 256 cells: every workgroup loops over
    several tiles, and the NHWC path stages its tiles through LDS."""
    from hybrid_flood_amd.coupled_forecast import network_features
    from oracle import coupled_oracle as co
    from oracle import swe_oracle as o

    case = _flat_case(1031, 1029, 21)
    params = _params(case, cuda)
    hi, terrain, rough = _hybrid_inputs(case, cuda, seed=5)
    got = network_features(_state(case, cuda), params, hi, 30.0, layout=layout).cpu().numpy()
    if layout == "nchw":
        got = got.transpose(0, 2, 3, 1)
    want = co.network_features(o.OracleState(case["h"], case["hu"], case["hv"]), _oparams(case),
                               terrain, rough, hi.state_mean, hi.state_std, 12.0, 30.0)
    assert np.array_equal(got, want)


def test_stage_features_bit_exact_on_a_large_odd_batch(cuda):
    """Two samples of an odd cell count: the second sample's output starts
    8 bytes off a 16-byte boundary and takes the strided path."""
    import torch

    from hybrid_flood_amd.staging import stage_features
    from oracle import coupled_oracle as co

    rng = np.random.default_rng(17)
    count, ny, nx = 2, 1031, 1029
    st = rng.normal(0, 1, (count, ny, nx, 3)).astype(np.float32)
    mask = rng.uniform(size=(ny, nx)) > 0.2
    terrain, rough = rng.normal(size=(2, ny, nx)).astype(np.float32)
    rain = np.array([3.0, 17.0], np.float32)
    got = stage_features(torch.as_tensor(st, device=cuda), rain, terrain, rough, mask,
                         [0.1, -0.2, 0.05], [0.3, 0.7, 0.9], 25.0)
    want = co.stage_features(st, rain, terrain, rough, mask, [0.1, -0.2, 0.05], [0.3, 0.7, 0.9], 25.0)
    assert np.array_equal(got.permute(0, 2, 3, 1).cpu().numpy(), want)
