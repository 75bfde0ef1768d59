"""A grid past 2^31 bytes per field: 20000 x 30000 cells (2.4 GB per float32
array, ~32 GB of device memory in all).

The case is invariant along x: a y-direction dam break over a bed with bumps
that vary in y only, a band of inactive rows, reflective sides, one Manning
value.  Every x face then joins two equal states, so its flux leaves the
x momentum exactly zero and every column evolves as a one-column problem
would: after 10 steps of the fast device loop every column of the large grid
must equal column 0 bit for bit, and column 0 must equal the same case on a
64-column grid (the per-cell arithmetic does not depend on the strip
geometry).  An address computed in 32 bits anywhere on the path -- packing,
the class-indexed record, the strip resources, the reductions, unpacking --
breaks one of the two.
"""

from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu

NY, NX = 20000, 30000


def _case(nx, dev):
    import torch

    from hybrid_flood_amd import numerics as nm

    y = torch.arange(NY, device=dev, dtype=torch.float32).view(NY, 1)
    bed_col = 0.2 * torch.sin(y / 37.0) ** 2 + 1.0e-5 * y
    bed = bed_col.expand(NY, nx).contiguous()
    mask = torch.ones((NY, nx), dtype=torch.bool, device=dev)
    mask[5000:5012, :] = False
    eta = torch.where(y < NY / 2, torch.tensor(1.0, device=dev), torch.tensor(0.6, device=dev))
    h = torch.clamp(eta - bed_col, min=0.0).expand(NY, nx).contiguous()
    h = torch.where(mask, h, torch.zeros_like(h))
    params = nm.make_params(bed, torch.full((NY, nx), 0.03, device=dev), mask,
                            boundary_types=(0, 0, 0, 0), dx=1.0, cfl=0.45, max_dt=1.0, device=dev)
    return params, nm.SWEState(h, torch.zeros_like(h), torch.zeros_like(h))


def test_x_invariant_dam_break_past_2gb_per_field(cuda):
    import torch

    from hybrid_flood_amd import numerics as nm

    print(f"large grid {NY} x {NX}", flush=True)
    params, st = _case(NX, cuda)
    assert st.h.numel() * 4 > 2**31
    info = nm.run_steps(st, params, 10, variant="fast", diagnostics=True)
    assert info["nonfinite"] == 0 and info["steps_used"] == 10
    ctx = nm.context_for(params, "fast")
    assert ctx.static_layout() == 2  # bed + Manning class
    for f in st:
        assert bool((f == f[:, :1]).all()), "a column of the large grid differs from column 0"
    assert float(st.hu.abs().max()) == 0.0
    assert float(st.hv.abs().max()) > 0.0  # the break has started
    cols = [f[:, 0].clone() for f in st]
    t_big = info["final_time_s"]
    del st, params, ctx
    nm._CONTEXTS.clear()
    torch.cuda.empty_cache()
    params_s, st_s = _case(64, cuda)
    info_s = nm.run_steps(st_s, params_s, 10, variant="fast", diagnostics=True)
    assert info_s["final_time_s"] == t_big
    for a, f in zip(cols, st_s):
        assert torch.equal(a, f[:, 0]), "the large grid's column differs from the 64-column run"
