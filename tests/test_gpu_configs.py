"""The two smaller benchmark workloads as GPU parity cases.

* Sloped plane with bumps and a channel, 256^2, dx 2 m, dry start, trapezoid
  rain peaking at 50 mm/h, 600 s, cfl 0.9, dry 1e-3, output every 60 s: both
  kernel variants against the FP32 oracle over the whole run.  The step
  counts must match (the dt sequence follows the same CFL maxima); depth
  within 1e-5 m and momentum within 5e-5 m^2/s of the oracle (measured on
  MI355X: 1.3e-7 / 4e-8 for the reference variant, 3.2e-7 / 3.5e-6 for the
  fast one); FP64 mass balance below 1e-6.
* Urban DEM, 1024^2, dx 2 m, one open (east) edge, 1-min rain series over
  10800 s, 181 outputs: too long for the oracle, so size-independent
  properties -- every output written, final time, no non-finite value,
  depth >= 0, outflow only through the open edge, the rain volume within 1e-3
  of the hyetograph's integral (the loop applies the rate at each step's
  start), and the FP64 mass balance below 1e-5 relative (float32 state over
  ~27k steps; measured 2.5e-6).
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _params(c, dev):
    from hybrid_flood_amd import numerics as nm

    return nm.make_params(c.bed, c.manning_n, c.domain_mask, boundary_types=c.boundary_types,
                          rainfall_times_s=c.rain_times_s, rainfall_rates_m_s=c.rain_rates_m_s,
                          dx=c.dx, cfl=c.cfl, dry_tolerance=c.dry_tolerance, max_dt=c.max_dt,
                          device=dev)


def _zero_state(c, dev):
    import torch

    from hybrid_flood_amd import numerics as nm

    h = torch.as_tensor(c.h0, device=dev)
    return nm.SWEState(h, torch.zeros_like(h), torch.zeros_like(h))


@pytest.fixture(scope="module")
def sloped_bumps_oracle():
    from hybrid_flood_amd import synthetic as sy
    from oracle import swe_oracle as o

    c = sy.sloped_bumps_case(256)
    op = o.OracleParams(bed=c.bed, manning_n=c.manning_n, domain_mask=c.domain_mask,
                        boundary_types=c.boundary_types, rainfall_times_s=c.rain_times_s,
                        rainfall_rates_m_s=c.rain_rates_m_s, dx=c.dx, cfl=c.cfl,
                        dry_tolerance=c.dry_tolerance, max_dt=c.max_dt)
    times = np.arange(0.0, 601.0, 60.0).astype(np.float32)
    z = np.zeros_like(c.h0)
    return c, times, o.integrate_to_outputs(o.OracleState(c.h0, z, z), op, times, max_steps=20000)


@pytest.mark.parametrize("variant", ["reference", "fast"])
def test_sloped_bumps_256_against_oracle(cuda, sloped_bumps_oracle, variant):
    from hybrid_flood_amd import numerics as nm

    c, times, ref = sloped_bumps_oracle
    r = nm.integrate_to_outputs(_zero_state(c, cuda), _params(c, cuda), times, max_steps=20000,
                                variant=variant)
    d = r.diagnostics
    assert int(r.outputs_written) == 11 == ref.outputs_written
    assert int(r.steps_used) == len(ref.dts)
    assert abs(d["mass_balance_rel"]) < 1e-6
    assert d["rain_volume"] == pytest.approx(ref.rain_volume, rel=1e-9)
    tol = {"h": 1e-5, "hu": 5e-5, "hv": 5e-5}
    for f, atol in tol.items():
        got = getattr(r.states, f).cpu().numpy()
        want = getattr(ref.states, f)
        assert np.abs(got - want).max() <= atol, f


def test_urban_1024_181_outputs_properties(cuda):
    import torch

    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd import synthetic as sy

    u = sy.urban_case(1024)
    times = np.arange(0.0, 10801.0, 60.0).astype(np.float32)
    r = nm.integrate_to_outputs(_zero_state(u, cuda), _params(u, cuda), times, max_steps=200000,
                                variant="fast")
    d = r.diagnostics
    assert int(r.outputs_written) == 181 and float(r.final_time_s) == 10800.0
    assert d["nonfinite"] == 0 and bool(torch.isfinite(r.states.h).all())
    assert float(r.states.h.min()) >= 0.0
    assert d["boundary_volume"] < 0.0  # water leaves through the open east edge only
    t = u.rain_times_s.astype(np.float64)
    q = u.rain_rates_m_s.astype(np.float64)
    closed = float(np.sum(0.5 * (q[1:] + q[:-1]) * np.diff(t))) * u.bed.size * u.dx * u.dx
    assert d["rain_volume"] == pytest.approx(closed, rel=1e-3)
    assert abs(d["mass_balance_rel"]) < 1e-5
