"""The fast device loop's packed layout (state and static fields interleaved,
the mask carried by the sign of the packed n^2 field) against the
caller-layout step.

The loop packs the caller's arrays at every start and unpacks at the end; its
arithmetic is the caller-layout step's, so a replay of the loop's recorded dt
schedule with ``finite_volume_step`` must give the same bits -- also with
negative Manning coefficients (the packed copy stores n^2, as the friction law
uses it) -- and fields changed in place between two runs must be used by the
second run.  The loop's alternating strip order (odd steps sweep in reverse)
must not change a bit either.  Nor may the static record's form: a one-byte
Manning class, (bed, n^2), or (bed, n^2, multiplier)."""

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
                          dx=case["dx"], dry_tolerance=case["dry"], device=dev)


def _state(case, dev):
    import torch

    from hybrid_flood_amd import numerics as nm

    return nm.SWEState(*(torch.as_tensor(case[k], device=dev).contiguous() for k in ("h", "hu", "hv")))


def _set_multiplier(case, form, rng):
    """field: per-cell values (the 3-float static record); uniform: 1.3 on
    every active cell and junk on inactive ones (the 2-float record with the
    value as a constant; inactive cells never use it); mask: the reference's
    default, the mask as float32."""
    mask = case["domain_mask"]
    if form == "uniform":
        junk = rng.uniform(-5.0, 5.0, mask.shape)
        case["rainfall_multiplier"] = np.where(mask, 1.3, junk).astype(np.float32)
    elif form == "mask":
        case["rainfall_multiplier"] = None
    return case


@pytest.mark.parametrize("mult", ["field", "uniform", "mask"])
@pytest.mark.parametrize("shape,sides", [((41, 77), (0, 1, 0, 0)), ((64, 130), (1, 1, 1, 1))])
def test_loop_equals_step_replay_with_negative_n(cuda, shape, sides, mult):
    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd.adjoint import rollout_forward

    case = random_case(17, shape, open_sides=sides, speed=0.8, rain=2.0e-4)
    rng = np.random.default_rng(3)
    flip = rng.uniform(size=shape) < 0.3
    case["manning_n"] = np.where(flip, -case["manning_n"], case["manning_n"]).astype(np.float32)
    case = _set_multiplier(case, mult, rng)
    params = _params(case, cuda)
    times = np.array([0.0, 2.0, 5.0], np.float32)
    _, res, tape = rollout_forward(params, _state(case, cuda), times, max_steps=3000, variant="fast",
                                   record=False)
    assert tape.steps > 10
    cur = _state(case, cuda)
    for k in range(tape.steps):
        cur = nm.finite_volume_step(cur, params, tape.dts[k], tape.ts[k], variant="fast")
    for got, want in zip((res.states.h[-1], res.states.hu[-1], res.states.hv[-1]), cur):
        assert np.array_equal(got.cpu().numpy(), want.cpu().numpy())


@pytest.mark.parametrize("rows", [5, 16])
def test_sweep_order_does_not_change_the_state(cuda, rows):
    from hybrid_flood_amd import numerics as nm

    case = random_case(29, (90, 150), open_sides=(1, 0, 1, 0), speed=0.8)
    params = _params(case, cuda)
    times = np.array([0.0, 1.0, 4.0], np.float32)
    ctx = nm.context_for(params, "fast")
    ctx.set_tile_rows(rows)
    try:
        runs = []
        for alternate in (1, 0, 1):
            ctx.set_sweep_order(alternate)
            runs.append(nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=3000,
                                                variant="fast"))
    finally:
        ctx.set_sweep_order(1)
        ctx.set_tile_rows(0)
    for other in runs[1:]:
        assert other.outputs_written == runs[0].outputs_written == len(times)
        for a, b in zip(runs[0].states, other.states):
            assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    with pytest.raises(ValueError):
        ctx.set_sweep_order(2)


def test_fields_changed_in_place_are_used(cuda):
    import torch

    from hybrid_flood_amd import numerics as nm

    case = random_case(23, (48, 70), open_sides=(0, 1, 0, 0), speed=0.8)
    params = _params(case, cuda)
    times = np.array([0.0, 3.0], np.float32)
    first = nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=2000, variant="fast")
    params.manning_n.mul_(1.7)
    params.bed.add_(0.01)
    params.rainfall_multiplier.mul_(0.5)
    second = nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=2000, variant="fast")
    case2 = dict(case, manning_n=params.manning_n.cpu().numpy(), bed=params.bed.cpu().numpy(),
                 rainfall_multiplier=params.rainfall_multiplier.cpu().numpy())
    fresh = nm.integrate_to_outputs(_state(case, cuda), _params(case2, cuda), times, max_steps=2000,
                                    variant="fast")
    for a, b in zip(second.states, fresh.states):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    assert not np.array_equal(first.states.h[-1].cpu().numpy(), second.states.h[-1].cpu().numpy())
    # the FP64 rain volume follows the changed multiplier too (its sum is
    # taken at every run start), so the mass balance still closes
    assert second.diagnostics["rain_volume"] == pytest.approx(fresh.diagnostics["rain_volume"],
                                                              rel=1e-12)
    assert second.diagnostics["rain_volume"] != pytest.approx(first.diagnostics["rain_volume"],
                                                              rel=1e-3)
    assert abs(second.diagnostics["mass_balance_rel"]) < 1e-6
    # a per-cell multiplier made uniform in place switches the packed static
    # record to the 2-float form, and back
    params.rainfall_multiplier.copy_(torch.where(params.domain_mask != 0, 0.8, -3.0))
    third = nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=2000, variant="fast")
    case3 = dict(case2, rainfall_multiplier=params.rainfall_multiplier.cpu().numpy())
    fresh3 = nm.integrate_to_outputs(_state(case, cuda), _params(case3, cuda), times, max_steps=2000,
                                     variant="fast")
    for a, b in zip(third.states, fresh3.states):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    assert third.diagnostics["rain_volume"] == pytest.approx(fresh3.diagnostics["rain_volume"],
                                                             rel=1e-12)
    assert abs(third.diagnostics["mass_balance_rel"]) < 1e-6
    params.rainfall_multiplier.copy_(torch.as_tensor(case2["rainfall_multiplier"], device=cuda))
    fourth = nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=2000, variant="fast")
    for a, b in zip(fourth.states, second.states):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())


def test_variant_switch_on_one_context(cuda):
    """One context runs fast (packed), reference (caller layout) and fast
    loops in turn; each equals the run of a context of that variant alone, and
    the variant cannot change while a loop is active."""
    import torch

    from hybrid_flood_amd import _lib
    from hybrid_flood_amd import numerics as nm

    case = random_case(29, (40, 66), open_sides=(1, 0, 0, 1), speed=0.8)
    params = _params(case, cuda)
    ctx = nm.context_for(params, "fast")
    lib = _lib.load()

    def run(variant):
        ctx.set_variant(variant)
        st = _state(case, cuda)
        _lib.check(lib.swe_run_steps(ctx.handle, _lib.state_struct(*st), 25, 0.0, None,
                                     _lib.stream_handle()), "swe_run_steps")
        torch.cuda.synchronize()
        return [x.cpu().numpy() for x in st]

    fast1, ref, fast2 = run("fast"), run("reference"), run("fast")
    alone = {}
    for variant in ("fast", "reference"):
        st = _state(case, cuda)
        nm.run_steps(st, _params(case, cuda), 25, variant=variant)
        alone[variant] = [x.cpu().numpy() for x in st]
    for a, b, c in zip(fast1, fast2, alone["fast"]):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    for a, b in zip(ref, alone["reference"]):
        assert np.array_equal(a, b)
    st = _state(case, cuda)
    _lib.check(lib.swe_loop_begin(ctx.handle, _lib.state_struct(*st), 0.0, 4, _lib.stream_handle()),
               "swe_loop_begin")
    with pytest.raises(ValueError):
        ctx.set_variant("reference")
    _lib.check(lib.swe_loop_end(ctx.handle, _lib.state_struct(*st), None, _lib.stream_handle()),
               "swe_loop_end")
    ctx.set_variant("fast")


@pytest.mark.parametrize("shape", [(1, 9), (9, 1), (2, 3), (3, 130), (130, 3), (65, 63)])
def test_degenerate_shapes_tile_invariant_and_near_oracle(cuda, shape):
    """Single-row and single-column grids and strips of 1-3 rows (south-marching
    bands of one row, a last band at the north edge, one partial strip
    column): every tile height gives the same bits, and the loop stays
    within the fast variant's bar of the FP64 oracle stepped with its own dt
    schedule."""
    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd.adjoint import rollout_forward
    from oracle import swe_oracle as o

    case = random_case(71 + shape[0], shape, open_sides=(1, 0, 0, 1), speed=0.8,
                       mask_frac=0.05)
    params = _params(case, cuda)
    ctx = nm.context_for(params, "fast")
    times = np.array([0.0, 0.7, 1.5], np.float32)
    results = []
    try:
        for rows in (1, 2, 3, 0):
            ctx.set_tile_rows(rows)
            results.append(rollout_forward(params, _state(case, cuda), times, max_steps=3000,
                                           variant="fast", record=False))
    finally:
        ctx.set_tile_rows(0)
    _, base, tape = results[0]
    assert base.outputs_written == len(times)
    for _, r, t in results[1:]:
        assert t.steps == tape.steps
        for a, b in zip(r.states, base.states):
            assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    op = o.OracleParams(bed=case["bed"], manning_n=case["manning_n"], domain_mask=case["domain_mask"],
                        rainfall_multiplier=case["rainfall_multiplier"],
                        boundary_types=case["boundary_types"], rainfall_times_s=case["rain_t"],
                        rainfall_rates_m_s=case["rain_r"], dx=case["dx"],
                        dry_tolerance=case["dry"], dtype=np.float64)
    st = o.OracleState(case["h"], case["hu"], case["hv"])
    for k in range(tape.steps):
        st = o.finite_volume_step(st, op, float(tape.dts[k]), float(tape.ts[k]))
    for f, atol in (("h", 1e-5), ("hu", 5e-5), ("hv", 5e-5)):
        got = getattr(base.states, f)[-1].cpu().numpy().astype(np.float64)
        assert np.abs(got - getattr(st, f)).max() <= atol, f



def _palette_manning(mask, n_values, rng):
    """Manning values drawn from ``n_values`` distinct float32 bit patterns on
    the active cells (0.0 and -0.0 among them, negative values included, each
    on some cell), and 7.5, a value outside the palette, on the inactive ones
    (never read there: counted, it would make 256 classes of 255)."""
    pal = rng.uniform(0.01, 0.2, max(n_values, 2)).astype(np.float32)
    pal[rng.uniform(size=pal.size) < 0.3] *= -1.0
    pal[0], pal[1] = np.float32(0.0), np.float32(-0.0)
    pal = pal[:n_values]
    assert np.unique(pal.view(np.uint32)).size == n_values
    idx = rng.integers(0, n_values, mask.shape)
    active = np.flatnonzero(mask)
    assert active.size >= n_values
    idx.ravel()[active[:n_values]] = np.arange(n_values)
    return np.where(mask, pal[idx], np.float32(7.5)).astype(np.float32)


@pytest.mark.parametrize("n_values,mult,layout", [(1, "uniform", 2), (37, "mask", 2),
                                                  (255, "uniform", 2), (256, "uniform", 1),
                                                  (37, "field", 0)])
def test_static_record_forms_give_the_same_bits(cuda, n_values, mult, layout):
    """The loop start stores the bed and a one-byte Manning class when the
    multiplier is uniform and the active cells carry at most 255 distinct
    Manning bit patterns, else (bed, n^2) or (bed, n^2, multiplier); every
    form gives the bits of the 8- or 12-byte record and of a replay of the
    loop's dt schedule with the caller-layout step."""
    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd.adjoint import rollout_forward

    rng = np.random.default_rng(100 + n_values)
    case = random_case(41, (70, 133), open_sides=(0, 1, 0, 1), speed=0.8, rain=2.0e-4)
    case["manning_n"] = _palette_manning(case["domain_mask"], n_values, rng)
    case = _set_multiplier(case, mult, rng)
    params = _params(case, cuda)
    ctx = nm.context_for(params, "fast")
    times = np.array([0.0, 1.5, 4.0], np.float32)
    runs, layouts = [], []
    try:
        for enable in (1, 0):
            ctx.set_static_classes(enable)
            runs.append(rollout_forward(params, _state(case, cuda), times, max_steps=3000,
                                        variant="fast", record=False))
            layouts.append(ctx.static_layout())
    finally:
        ctx.set_static_classes(1)
    print(f"classes {n_values}, multiplier {mult}: layouts {layouts}", flush=True)
    assert layouts == [layout, min(layout, 1)]
    (_, with_cls, tape), (_, without, tape2) = runs
    assert tape.steps == tape2.steps > 10
    for a, b in zip(with_cls.states, without.states):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    cur = _state(case, cuda)
    for k in range(tape.steps):
        cur = nm.finite_volume_step(cur, params, tape.dts[k], tape.ts[k], variant="fast")
    for got, want in zip((with_cls.states.h[-1], with_cls.states.hu[-1], with_cls.states.hv[-1]),
                         cur):
        assert np.array_equal(got.cpu().numpy(), want.cpu().numpy())
    with pytest.raises(ValueError):
        ctx.set_static_classes(2)


def test_manning_classes_follow_in_place_changes(cuda):
    """A Manning field changed in place between runs re-chooses the record:
    300 values fall back to (bed, n^2), a new 12-value palette returns to
    classes, and every run equals a fresh context's."""
    import torch

    from hybrid_flood_amd import numerics as nm

    rng = np.random.default_rng(7)
    case = _set_multiplier(random_case(43, (50, 90), open_sides=(1, 0, 0, 0), speed=0.8),
                           "uniform", rng)
    mask = case["domain_mask"]
    case["manning_n"] = _palette_manning(mask, 20, rng)
    params = _params(case, cuda)
    ctx = nm.context_for(params, "fast")
    times = np.array([0.0, 3.0], np.float32)
    for n_values, layout in ((20, 2), (300, 1), (12, 2)):
        if n_values != 20:
            case["manning_n"] = _palette_manning(mask, n_values, rng)
            params.manning_n.copy_(torch.as_tensor(case["manning_n"], device=cuda))
        got = nm.integrate_to_outputs(_state(case, cuda), params, times, max_steps=2000,
                                      variant="fast")
        assert ctx.static_layout() == layout
        fresh = nm.integrate_to_outputs(_state(case, cuda), _params(case, cuda), times,
                                        max_steps=2000, variant="fast")
        for a, b in zip(got.states, fresh.states):
            assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())


@pytest.mark.parametrize("variant", ["fast", "reference"])
def test_loop_finish_modes_agree(cuda, variant):
    """The three ways a device-loop step closes (fused: the next step derives
    dt from CFL slots; split: a finish kernel after each step; last
    workgroup: the step kernel's last workgroup finishes) give the same
    states, dt schedule and FP64 diagnostics; the mode cannot change while a
    loop is active."""
    from hybrid_flood_amd import _lib
    from hybrid_flood_amd import numerics as nm
    from hybrid_flood_amd.adjoint import rollout_forward

    case = random_case(53, (75, 140), open_sides=(1, 1, 0, 0), speed=0.8, rain=2.0e-4)
    params = _params(case, cuda)
    ctx = nm.context_for(params, variant)
    times = np.array([0.0, 1.0, 3.0], np.float32)
    runs = {}
    try:
        for mode in (1, 2, 0):
            ctx.set_loop_finish(mode)
            print(f"finish mode {mode}, variant {variant}", flush=True)
            runs[mode] = rollout_forward(params, _state(case, cuda), times, max_steps=3000,
                                         variant=variant, record=False)
    finally:
        ctx.set_loop_finish(1)
    _, base, tape = runs[1]
    assert tape.steps > 10
    for mode in (2, 0):
        _, r, t = runs[mode]
        assert t.steps == tape.steps
        assert np.array_equal(np.asarray(t.dts), np.asarray(tape.dts))
        for a, b in zip(r.states, base.states):
            assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
        for key in ("mass_balance_rel", "rain_volume", "boundary_volume", "clip_volume",
                    "volume_final"):
            assert r.diagnostics[key] == base.diagnostics[key], key
    with pytest.raises(ValueError):
        ctx.set_loop_finish(3)
    st = _state(case, cuda)
    lib = _lib.load()
    _lib.check(lib.swe_loop_begin(ctx.handle, _lib.state_struct(*st), 0.0, 4, _lib.stream_handle()),
               "swe_loop_begin")
    with pytest.raises(ValueError):
        ctx.set_loop_finish(2)
    _lib.check(lib.swe_loop_end(ctx.handle, _lib.state_struct(*st), None, _lib.stream_handle()),
               "swe_loop_end")
