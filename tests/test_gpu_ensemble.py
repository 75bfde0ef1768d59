"""GPU ensemble: every member of a batched, uneven-size ensemble must equal the
same member run alone through the single-domain device loop, bit for bit
(same kernels, same per-member dt sequence)."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import random_case

pytestmark = pytest.mark.gpu


def _member(case, dev):
    import torch

    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd.ensemble import Member

    p = nm.make_params(case["bed"], case["manning_n"], case["domain_mask"], case["rainfall_multiplier"],
                       boundary_types=case["boundary_types"], rainfall_times_s=case["rain_t"],
                       rainfall_rates_m_s=case["rain_r"], dx=case["dx"], dry_tolerance=case["dry"],
                       cfl=0.45, max_dt=10.0, device=dev)
    st = nm.SWEState(*(torch.as_tensor(case[k], device=dev).contiguous() for k in ("h", "hu", "hv")))
    return Member(p, st)


@pytest.mark.parametrize("variant", ["fast", "reference"])
@pytest.mark.parametrize("rows", [16, 5])
def test_ensemble_members_equal_solo_runs(cuda, variant, rows):
    import torch

    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd.ensemble import EnsembleRunner

    shapes = [(40, 70), (129, 64), (17, 200), (1, 9), (63, 62)]
    sides = [(0, 0, 0, 0), (0, 1, 0, 0), (1, 1, 1, 1), (0, 0, 0, 0), (1, 0, 0, 1)]
    cases = [random_case(300 + k, s, open_sides=b, speed=1.0 + 0.5 * k)
             for k, (s, b) in enumerate(zip(shapes, sides))]
    members = [_member(c, cuda) for c in cases]
    solo = []
    t_end = 2.5
    for m in members:
        r = nm.integrate_to_outputs(m.state, m.params, np.array([0.0, t_end], np.float32),
                                    max_steps=5000, variant=variant)
        solo.append(r)
    ens = EnsembleRunner(members, variant=variant)
    ens.set_tile_rows(rows)
    infos = ens.run(t_end, max_steps=5000)
    torch.cuda.synchronize()
    for k, (m, r, info) in enumerate(zip(members, solo, infos)):
        assert info["steps_used"] == r.steps_used, k
        assert info["final_time_s"] == r.final_time_s, k
        for got, want in zip(m.state, r.states):
            assert torch.equal(got, want[-1]), f"member {k} differs"
        assert abs(info["mass_balance_rel"]) < 1e-6
        assert info["volume_final"] == pytest.approx(r.diagnostics["volume_final"], rel=1e-12)


def test_ensemble_respects_max_steps(cuda):
    from hybrid_flood_amd.ensemble import EnsembleRunner

    members = [_member(random_case(400 + k, s, speed=1.0), cuda) for k, s in enumerate([(30, 40), (50, 20)])]
    ens = EnsembleRunner(members)
    with pytest.raises(RuntimeError):
        ens.run(1000.0, max_steps=7)
