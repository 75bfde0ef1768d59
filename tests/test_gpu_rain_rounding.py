"""Stochastic rounding of the rain source on the GPU (swe_set_rain_rounding;
an option of this build, off by default).

* The reference-order loop with stochastic rain equals the FP32 oracle with
  the same option bit for bit where the oracle's loop is reproduced exactly
  (no friction: the friction divisor is exactly 1), and differs from its own
  round-to-nearest run: the draw is the oracle's (cell index, step index).
* The fast loop's stochastic draw does not depend on the tile or the sweep
  order, and an ensemble member equals its solo run with the option.
* The 1024^2 urban configuration (10800 s of rain, ~27k steps): the FP64 mass
  balance closes below the 1e-6 target with stochastic rain, where round to
  nearest leaves ~2.5e-6 (the float32 rounding of the rain addition, the same
  on every ponded cell; tests/test_oracle_rain_rounding.py).
* The option excludes adjoint checkpoints, whose replays round to nearest.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import random_case

pytestmark = pytest.mark.gpu


def _params(case, dev):
    from hybrid_flood_amd import numerics as nm

    return nm.make_params(case["bed"], case["manning_n"], case["domain_mask"],
                          case["rainfall_multiplier"], boundary_types=case["boundary_types"],
                          rainfall_times_s=case["rain_t"], rainfall_rates_m_s=case["rain_r"],
                          dx=case["dx"], dry_tolerance=case["dry"], cfl=0.45, max_dt=10.0,
                          device=dev)


def _oparams(case):
    from oracle import swe_oracle as o

    return o.OracleParams(bed=case["bed"], manning_n=case["manning_n"],
                          domain_mask=case["domain_mask"],
                          rainfall_multiplier=case["rainfall_multiplier"],
                          boundary_types=case["boundary_types"], rainfall_times_s=case["rain_t"],
                          rainfall_rates_m_s=case["rain_r"], dx=case["dx"], cfl=0.45,
                          dry_tolerance=case["dry"], max_dt=10.0)


def _state(case, dev):
    import torch

    from hybrid_flood_amd import numerics as nm

    return nm.SWEState(*(torch.as_tensor(case[k], device=dev).contiguous() for k in ("h", "hu", "hv")))


def _rainy(seed, shape, **kw):
    c = random_case(seed, shape, **kw)
    # a steady rate large enough that every step adds rain
    c["rain_t"] = np.array([0.0, 1000.0], np.float32)
    c["rain_r"] = np.array([3e-5, 3e-5], np.float32)
    return c


@pytest.mark.parametrize("shape", [(12, 14), (50, 130)])
def test_reference_loop_with_stochastic_rain_matches_oracle(cuda, shape):
    from hybrid_flood_amd import numerics as nm
    from oracle import swe_oracle as o

    case = _rainy(91, shape, open_sides=(0, 1, 0, 1), speed=1.0)
    case["manning_n"] = np.zeros(shape, np.float32)
    params = _params(case, cuda)
    times = np.array([0.0, 0.5, 1.0, 1.7], np.float32)
    res = nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=500,
                                  variant="reference", rain_rounding="stochastic")
    ref = o.integrate_to_outputs(o.OracleState(case["h"], case["hu"], case["hv"]), _oparams(case),
                                 times.astype(np.float64), max_steps=500,
                                 rain_rounding="stochastic")
    assert res.steps_used == ref.steps_used > 3 and res.outputs_written == 4
    assert res.final_time_s == ref.final_time_s
    for g, r in zip(res.states, ref.states):
        assert np.array_equal(g.cpu().numpy(), r)
    near = nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=500,
                                   variant="reference")
    assert not np.array_equal(near.states.h.cpu().numpy(), res.states.h.cpu().numpy())
    assert abs(res.diagnostics["mass_balance_rel"]) < 1e-6


def test_fast_loop_stochastic_rain_is_tile_and_order_invariant(cuda):
    import torch

    from hybrid_flood_amd import numerics as nm

    case = _rainy(92, (150, 260), open_sides=(0, 1, 0, 0), speed=0.8)
    params = _params(case, cuda)
    ctx = nm.context_for(params, "fast")
    times = np.array([0.0, 2.0], np.float32)
    outs = []
    for rows, alt in ((16, 1), (5, 1), (23, 0)):
        ctx.set_tile_rows(rows)
        ctx.set_sweep_order(alt)
        r = nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=5000,
                                    variant="fast", rain_rounding="stochastic")
        outs.append(r)
    ctx.set_sweep_order(1)
    for r in outs[1:]:
        assert r.steps_used == outs[0].steps_used
        for a, b in zip(r.states, outs[0].states):
            assert torch.equal(a, b)
    near = nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=5000,
                                   variant="fast")
    assert not torch.equal(near.states.h, outs[0].states.h)


def test_fast_loop_stochastic_rain_near_oracle(cuda):
    """The fast variant with stochastic rain against the FP32 oracle with the
    same option: the same step count, depth within 1e-5 m, momentum within
    5e-5 m^2/s (the bars of the workload tests)."""
    from hybrid_flood_amd import numerics as nm
    from oracle import swe_oracle as o

    case = _rainy(93, (64, 96), open_sides=(0, 1, 0, 0), speed=0.6)
    params = _params(case, cuda)
    times = np.array([0.0, 1.0, 3.0], np.float32)
    res = nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=5000,
                                  variant="fast", rain_rounding="stochastic")
    ref = o.integrate_to_outputs(o.OracleState(case["h"], case["hu"], case["hv"]), _oparams(case),
                                 times.astype(np.float64), max_steps=5000,
                                 rain_rounding="stochastic")
    assert res.steps_used == ref.steps_used
    for g, r, tol in zip(res.states, ref.states, (1e-5, 5e-5, 5e-5)):
        assert np.abs(g.cpu().numpy() - r).max() <= tol


def test_urban_1024_mass_balance_below_target_with_stochastic_rain(cuda):
    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd import synthetic as sy

    import torch

    u = sy.urban_case(1024)
    params = nm.make_params(u.bed, u.manning_n, u.domain_mask, boundary_types=u.boundary_types,
                            rainfall_times_s=u.rain_times_s, rainfall_rates_m_s=u.rain_rates_m_s,
                            dx=u.dx, cfl=u.cfl, dry_tolerance=u.dry_tolerance, max_dt=u.max_dt,
                            device=cuda)
    times = np.arange(0.0, 10801.0, 60.0).astype(np.float32)
    out = {}
    for mode in ("nearest", "stochastic"):
        h = torch.as_tensor(u.h0, device=cuda)
        st = nm.SWEState(h, torch.zeros_like(h), torch.zeros_like(h))
        out[mode] = nm.integrate_to_outputs(st, params, times, max_steps=200000, variant="fast",
                                            rain_rounding=mode)
    dn, ds = out["nearest"].diagnostics, out["stochastic"].diagnostics
    print(f"urban 1024: mass balance nearest {dn['mass_balance_rel']:.3e}, "
          f"stochastic {ds['mass_balance_rel']:.3e}", flush=True)
    assert out["stochastic"].outputs_written == 181 and ds["nonfinite"] == 0
    assert abs(dn["mass_balance_rel"]) > 1e-6
    assert abs(ds["mass_balance_rel"]) < 1e-6
    assert abs(ds["steps_used"] - dn["steps_used"]) <= 0.01 * dn["steps_used"]


def test_ensemble_member_with_stochastic_rain_equals_solo(cuda):
    import torch

    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd.ensemble import EnsembleRunner, Member

    cases = [_rainy(94 + k, s, open_sides=(0, 1, 0, 0), speed=0.7)
             for k, s in enumerate([(40, 70), (33, 90)])]
    cases[1]["manning_n"] = np.where(cases[1]["domain_mask"], 0.03, 0.02).astype(np.float32)
    cases[1]["rainfall_multiplier"] = cases[1]["domain_mask"].astype(np.float32)
    members = [Member(_params(c, cuda), _state(c, cuda)) for c in cases]
    solo = [nm.integrate_to_outputs(m.state, m.params, np.array([0.0, 2.0], np.float32),
                                    max_steps=5000, variant="fast", rain_rounding="stochastic")
            for m in members]
    ens = EnsembleRunner(members, variant="fast")
    ens.set_rain_rounding("stochastic")
    infos = ens.run(2.0, max_steps=5000)
    torch.cuda.synchronize()
    assert ens.static_layout(1) == 2
    for m, r, info in zip(members, solo, infos):
        assert info["steps_used"] == r.steps_used
        for got, want in zip(m.state, r.states):
            assert torch.equal(got, want[-1])


def test_stochastic_rain_excludes_checkpoints(cuda):
    import torch

    from hybrid_flood_amd import _lib
    from hybrid_flood_amd import numerics as nm

    lib = _lib.load()
    case = _rainy(95, (20, 30))
    ctx = nm.context_for(_params(case, cuda), "fast")
    buf = [torch.zeros((2, 20, 30), device=cuda) for _ in range(3)]

    def checkpoints(every, capacity):
        ptrs = [b.data_ptr() for b in buf] if every else [None, None, None]
        _lib.check(lib.swe_set_checkpoints(ctx.handle, *ptrs, every, capacity), "swe_set_checkpoints")

    ctx.set_rain_rounding("stochastic")
    with pytest.raises(ValueError):
        checkpoints(3, 2)
    ctx.set_rain_rounding("nearest")
    checkpoints(3, 2)
    with pytest.raises(ValueError):
        ctx.set_rain_rounding("stochastic")
    checkpoints(0, 0)
    ctx.set_rain_rounding("stochastic")
    ctx.set_rain_rounding("nearest")
    with pytest.raises(ValueError):
        ctx.set_rain_rounding("upward")
