"""The oracle's fixed-length advance (hybrid/coupled_forecast.py:123-161),
pinned by closed forms of the loop rather than by its own formula:

  * lake at rest: the state never changes, so every step has the same CFL dt
    (cfl*dx / (2 sqrt(g h)), capped at max_dt) and the step count is the
    ceiling of the interval over that dt, with the last step clamped to the
    target;
  * a step budget one short of that count leaves the clock short of the
    target, which the reference turns into an all-NaN state;
  * a target at or before the start takes no step and keeps the state.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

G = 9.81


def _o():
    from oracle import swe_oracle as o

    return o


def _lake(shape=(9, 11), depth=0.3, dx=2.0, cfl=0.45, max_dt=10.0):
    o = _o()
    bed = np.zeros(shape, np.float32)
    p = o.OracleParams(bed=bed, manning_n=np.full(shape, 0.03, np.float32),
                       domain_mask=np.ones(shape, bool), boundary_types=(0, 0, 0, 0),
                       rainfall_times_s=np.array([0.0, 1e4], np.float32),
                       rainfall_rates_m_s=np.zeros(2, np.float32), dx=dx, cfl=cfl,
                       dry_tolerance=1e-4, max_dt=max_dt, dtype=np.float64)
    h = np.full(shape, depth)
    z = np.zeros(shape)
    return o.OracleState(h, z, z), p


@pytest.mark.parametrize("depth,dx,interval", [(0.3, 2.0, 7.3), (2.0, 1.0, 1.0), (0.05, 5.0, 60.0)])
def test_lake_at_rest_step_count_closed_form(depth, dx, interval):
    o = _o()
    st, p = _lake(depth=depth, dx=dx)
    dt = min(0.45 / (2.0 * math.sqrt(G * depth) / dx), 10.0)
    want = math.ceil(interval / dt - 1e-12)
    out, steps, t = o.advance_physics_to_time(st, p, 100.0, 100.0 + interval, max_substeps=want + 5)
    assert steps == want
    assert t == pytest.approx(100.0 + interval, abs=1e-9)
    np.testing.assert_array_equal(out.h, st.h)
    np.testing.assert_array_equal(out.hu, 0.0)
    # one step short of the count: the target is missed and the state is NaN
    out, steps, t = o.advance_physics_to_time(st, p, 100.0, 100.0 + interval, max_substeps=want - 1)
    assert steps == want - 1
    assert t < 100.0 + interval - 1e-6
    assert all(np.isnan(x).all() for x in out)


def test_target_not_after_start_takes_no_step():
    o = _o()
    st, p = _lake()
    for target in (5.0, 5.0 + 5e-8, 4.0):
        out, steps, t = o.advance_physics_to_time(st, p, 5.0, target, max_substeps=10)
        assert steps == 0 and t == 5.0
        np.testing.assert_array_equal(out.h, st.h)


def test_zero_budget_with_open_interval_is_nan():
    o = _o()
    st, p = _lake()
    out, steps, _ = o.advance_physics_to_time(st, p, 0.0, 1.0, max_substeps=0)
    assert steps == 0 and np.isnan(out.h).all()


# ---------------------------------------------------------------------------
# Residual correction and features (coupled_forecast.py:53-120)
# ---------------------------------------------------------------------------

def _co():
    from oracle import coupled_oracle as co

    return co


def _field_case(seed=0, shape=(13, 17)):
    o = _o()
    rng = np.random.default_rng(seed)
    mask = rng.uniform(size=shape) > 0.15
    h = np.where(mask, rng.uniform(0, 0.5, shape), 0).astype(np.float32)
    h[rng.uniform(size=shape) < 0.1] = 0.0
    hu = (h * rng.normal(0, 0.5, shape)).astype(np.float32)
    hv = (h * rng.normal(0, 0.5, shape)).astype(np.float32)
    p = o.OracleParams(bed=np.zeros(shape, np.float32), manning_n=np.full(shape, 0.03, np.float32),
                       domain_mask=mask, boundary_types=(0, 0, 0, 0),
                       rainfall_times_s=np.array([0.0, 60.0, 120.0], np.float32),
                       rainfall_rates_m_s=np.array([0.0, 1.0e-5, 4.0e-6], np.float32),
                       dx=2.0, dry_tolerance=1e-4, dtype=np.float32)
    return o.OracleState(h, hu, hv), p, rng


def test_zero_residual_keeps_masked_depth():
    co = _co()
    st, p, _ = _field_case()
    out, (tgt, prov, scale, rescaled) = co.apply_residual_correction(
        st, np.zeros(st.h.shape + (3,), np.float32), p, correction_relaxation=0.05)
    assert tgt == prov and scale == 1.0 and rescaled == 1.0
    np.testing.assert_array_equal(out.h, np.where(p.domain_mask, st.h, 0))
    wet = p.domain_mask & (st.h > 1e-4)
    # h * (hu / h) is hu up to the rounding of the quotient and the product
    np.testing.assert_allclose(out.hu[wet], st.hu[wet], rtol=3e-7)
    assert np.all(out.hu[~wet] == 0) and np.all(out.hv[~wet] == 0)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_correction_preserves_physics_volume_and_positivity(seed):
    co = _co()
    st, p, rng = _field_case(seed)
    r = rng.normal(0, 2.0, st.h.shape + (3,)).astype(np.float32)
    out, (tgt, _, _, _) = co.apply_residual_correction(st, r, p, correction_relaxation=0.2)
    assert np.all(out.h >= 0)
    assert np.all(out.h[~p.domain_mask] == 0)
    # each cell's float32 rounding bounds the volume error
    assert abs(float(np.sum(out.h.astype(np.float64))) - tgt) <= 1e-6 * tgt


def test_clip_equals_preclipped_residual():
    co = _co()
    st, p, rng = _field_case(4)
    r = rng.normal(0, 5.0, st.h.shape + (3,)).astype(np.float32)
    lim = np.array([0.5, 1.0, 2.0], np.float32)
    a, _ = co.apply_residual_correction(st, r, p, correction_relaxation=0.1, clip_limits=lim)
    b, _ = co.apply_residual_correction(st, np.clip(r, -lim, lim), p, correction_relaxation=0.1)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_dry_domain_returns_predicted_depth():
    co = _co()
    st, p, rng = _field_case(5)
    z = np.zeros_like(st.h)
    dry = st._replace(h=z, hu=z, hv=z)
    out, (_, _, _, rescaled) = co.apply_residual_correction(
        dry, rng.normal(0, 1, z.shape + (3,)).astype(np.float32), p, correction_relaxation=0.05)
    assert rescaled == 0.0
    np.testing.assert_array_equal(out.h, z)


def test_features_uniform_state_closed_form():
    co = _co()
    st, p, _ = _field_case(6)
    shape = st.h.shape
    h = np.full(shape, 0.5, np.float32)
    s = st._replace(h=h, hu=np.full(shape, 0.2, np.float32), hv=np.full(shape, -0.1, np.float32))
    mean, std = [0.1, 0.0, 0.05], [0.2, 0.5, 0.25]
    terrain = np.arange(np.prod(shape), dtype=np.float32).reshape(shape)
    f = co.network_features(s, p, terrain, -terrain, mean, std, 20.0, 90.0)[0]
    m = p.domain_mask
    np.testing.assert_allclose(f[m, 0], (0.5 - 0.1) / 0.2, rtol=1e-6)
    np.testing.assert_allclose(f[m, 1], (0.4 - 0.0) / 0.5, rtol=1e-6)
    np.testing.assert_allclose(f[m, 2], (-0.2 - 0.05) / 0.25, rtol=1e-6)
    assert np.all(f[~m, :3] == 0)
    # rain at t=90 s: halfway between 1e-5 (60 s) and 4e-6 (120 s) m/s
    np.testing.assert_allclose(f[..., 3], 7.0e-6 * 3.6e6 / 20.0, rtol=1e-6)
    np.testing.assert_array_equal(f[..., 4], terrain)
    np.testing.assert_array_equal(f[..., 5], -terrain)


def test_coupled_loop_zero_model_is_physics_for_depth():
    co = _co()
    o = _o()
    st, p = _lake(shape=(8, 9), depth=0.2)
    st = o.OracleState(st.h.astype(np.float32), st.hu.astype(np.float32), st.hv.astype(np.float32))
    p32 = o.OracleParams(bed=p.bed, manning_n=p.manning_n, domain_mask=p.domain_mask,
                         boundary_types=p.boundary_types, rainfall_times_s=p.rainfall_times_s,
                         rainfall_rates_m_s=p.rainfall_rates_m_s, dx=p.dx, cfl=p.cfl,
                         dry_tolerance=p.dry_tolerance, max_dt=p.max_dt, dtype=np.float32)
    z = np.zeros(st.h.shape, np.float32)
    stack, _, steps = co.run_coupled_forecast(
        st, p32, lambda f: np.zeros(f.shape[:3] + (3,), np.float32), z, z, [0, 0, 0], [1, 1, 1],
        1.0, [1, 1, 1], [0.0, 5.0, 10.0], max_substeps_per_interval=64)
    ref, rsteps, _ = o.advance_physics_to_time(st, p32, 0.0, 5.0, max_substeps=64)
    np.testing.assert_array_equal(stack.h[1], ref.h)
    assert steps >= rsteps > 0


def test_stage_features_closed_form():
    co = _co()
    rng = np.random.default_rng(0)
    st = rng.normal(size=(3, 5, 7, 3)).astype(np.float32)
    mask = rng.uniform(size=(5, 7)) > 0.3
    f = co.stage_features(st, [0.0, 12.0, 30.0], np.ones((5, 7)), -np.ones((5, 7)), mask,
                          [1.0, 0.0, -1.0], [2.0, 1.0, 4.0], 6.0)
    assert f.shape == (3, 5, 7, 6)
    np.testing.assert_allclose(f[:, mask, 0], (st[:, mask, 0] - 1.0) / 2.0, rtol=1e-7)
    np.testing.assert_allclose(f[:, mask, 2], (st[:, mask, 2] + 1.0) / 4.0, rtol=1e-7)
    assert np.all(f[:, ~mask, :3] == 0)
    np.testing.assert_array_equal(f[:, 2, 3, 3], [0.0, 2.0, 5.0])
    assert np.all(f[..., 4] == 1) and np.all(f[..., 5] == -1)
    d = co.correct_depth(np.array([[0.5, 0.1]], np.float32), np.array([[[0.25, 9], [-0.3, 9]]], np.float32))
    np.testing.assert_array_equal(d, [[0.75, 0.0]])
