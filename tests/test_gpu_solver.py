"""The reference's own solver tests, run through the drop-in ShallowWaterSolver.

Mirrors tests/test_jax_solver_conservation.py:73-122 (flat-basin mass, lake at
rest over a bump, lake at rest at irregular mask walls, same bars),
tests/test_v1_v2_architecture.py:107-117 (lake at rest over a sine bump) and
tests/test_hybrid_rollout.py:37-64 (the correction keeps the physics volume),
plus the run-time errors of jax_solver/shallow_water_2d.py:278-306 and the
record of jax_solver/benchmarking.py:17-97.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VARIANTS = ["fast", "reference"]


def _solver_for_bed(bed, *, resolution_m=1.0, manning_n=0.0, mask=None, variant="fast", cuda=None):
    from hybrid_flood_amd.shallow_water_2d import ShallowWaterSolver, StructuredGrid

    if mask is None:
        mask = np.ones(bed.shape, dtype=bool)
    grid = StructuredGrid(bed=bed.astype(np.float32),
                          manning_n=np.full(bed.shape, manning_n, dtype=np.float32),
                          domain_mask=mask, resolution_m=resolution_m)
    return ShallowWaterSolver(grid, rainfall_times_s=np.array([0.0, 100.0], np.float32),
                              rainfall_rates_m_s=np.array([0.0, 0.0], np.float32), cfl=0.4,
                              dry_tolerance_m=1.0e-6, max_dt_s=0.1, variant=variant, device=cuda)


@pytest.mark.parametrize("variant", VARIANTS)
def test_reflective_flat_basin_conserves_water_mass(cuda, variant):
    bed = np.zeros((20, 20), np.float32)
    depth = np.full_like(bed, 0.1)
    depth[8:12, 8:12] = 0.2
    solver = _solver_for_bed(bed, variant=variant, cuda=cuda)
    result = solver.run(solver.initial_state(depth_m=depth), np.array([0.0, 1.0, 2.0], np.float32),
                        max_steps=1_000)
    volume = result.states.h.double().sum(dim=(1, 2)).cpu().numpy()
    np.testing.assert_allclose(volume, volume[0], rtol=2.0e-7, atol=1.0e-5)


@pytest.mark.parametrize("variant", VARIANTS)
def test_lake_at_rest_remains_well_balanced(cuda, variant):
    coordinate = np.linspace(-1.0, 1.0, 20)
    xx, yy = np.meshgrid(coordinate, coordinate)
    bed = (0.03 * np.exp(-5.0 * (xx**2 + yy**2))).astype(np.float32)
    solver = _solver_for_bed(bed, variant=variant, cuda=cuda)
    result = solver.run(solver.initial_state(depth_m=0.2 - bed), np.array([0.0, 1.0, 2.0], np.float32),
                        max_steps=1_000)
    final_surface = result.states.h[-1].cpu().numpy() + bed
    np.testing.assert_allclose(final_surface, 0.2, atol=2.0e-7)
    assert float(result.states.hu[-1].abs().max()) < 1.0e-6
    assert float(result.states.hv[-1].abs().max()) < 1.0e-6


@pytest.mark.parametrize("variant", VARIANTS)
def test_lake_at_rest_remains_balanced_at_irregular_mask_walls(cuda, variant):
    bed = np.zeros((12, 12), np.float32)
    mask = np.ones_like(bed, dtype=bool)
    mask[4:8, 5:7] = False
    solver = _solver_for_bed(bed, mask=mask, variant=variant, cuda=cuda)
    result = solver.run(solver.initial_state(depth_m=0.1), np.array([0.0, 0.5], np.float32),
                        max_steps=1_000)
    final = [x[-1].cpu().numpy() for x in result.states]
    np.testing.assert_allclose(final[0][mask], 0.1, atol=2.0e-7)
    assert np.max(np.abs(final[1][mask])) < 1.0e-6
    assert np.max(np.abs(final[2][mask])) < 1.0e-6


@pytest.mark.parametrize("variant", VARIANTS)
def test_lake_at_rest_over_sine_bump(cuda, variant):
    """tests/test_v1_v2_architecture.py:107-117 (the NumPy twin's bar, atol 2e-6)."""
    from hybrid_flood_amd.shallow_water_2d import ShallowWaterSolver, StructuredGrid

    x = np.linspace(0.0, 1.0, 12, dtype=np.float32)
    bed = (0.2 * np.outer(np.sin(np.pi * x) ** 2, np.sin(np.pi * x) ** 2)).astype(np.float32)
    grid = StructuredGrid(bed=bed, manning_n=np.full(bed.shape, 0.03, np.float32),
                          domain_mask=np.ones(bed.shape, bool), resolution_m=2.0)
    solver = ShallowWaterSolver(grid, rainfall_times_s=np.array([0.0, 100.0], np.float32),
                                rainfall_rates_m_s=np.zeros(2, np.float32), variant=variant, device=cuda)
    h0 = (np.float32(0.35) - bed).astype(np.float32)
    result = solver.run(solver.initial_state(depth_m=h0), np.array([0.0, 20.0], np.float32),
                        max_steps=100)
    np.testing.assert_allclose(result.states.h[-1].cpu().numpy(), h0, atol=2.0e-6)
    assert float(result.states.hu[-1].abs().max()) < 2.0e-6
    assert float(result.states.hv[-1].abs().max()) < 2.0e-6


def test_run_rejects_bad_times_and_reports_max_steps(cuda):
    bed = np.zeros((10, 10), np.float32)
    solver = _solver_for_bed(bed, cuda=cuda)
    st = solver.initial_state(depth_m=0.1)
    with pytest.raises(ValueError, match="start at zero and increase strictly"):
        solver.run(st, np.array([1.0, 2.0], np.float32))
    with pytest.raises(ValueError, match="start at zero and increase strictly"):
        solver.run(st, np.array([0.0, 2.0, 2.0], np.float32))
    with pytest.raises(RuntimeError, match="max_steps=3 was insufficient"):
        solver.run(st, np.array([0.0, 5.0], np.float32), max_steps=3)


def test_benchmark_compiled_steps_record(cuda, tmp_path):
    from hybrid_flood_amd.benchmarking import benchmark_compiled_steps

    bed = np.zeros((64, 96), np.float32)
    solver = _solver_for_bed(bed, cuda=cuda)
    report = benchmark_compiled_steps(solver.initial_state(depth_m=0.1), solver.params, iterations=5,
                                      dt_s=0.01, output_directory=tmp_path)
    gpu = [r for r in report["benchmarks"] if r["platform"] == "gpu"][0]
    assert gpu["iterations"] == 5 and gpu["grid_shape"] == [64, 96]
    assert gpu["steps_per_second"] > 0 and gpu["includes_jit_compilation"] is False
    assert report["gpu_benchmark_available"] is True
    assert (tmp_path / report["output_path"].split("/")[-1]).exists()


def test_neural_correction_preserves_physics_water_volume(cuda):
    """tests/test_hybrid_rollout.py:37-64 on a synthetic grid (same bar, rtol 2e-6)."""
    import torch

    from hybrid_flood_amd import synthetic as sy
    from hybrid_flood_amd.coupled_forecast import apply_residual_correction
    from hybrid_flood_amd.numerics import SWEState
    from hybrid_flood_amd.shallow_water_2d import ShallowWaterSolver, StructuredGrid

    case = sy.urban_case(96, dx=50.0, seed=3)
    yy, xx = np.mgrid[0:96, 0:96]
    mask = (xx - 47.5) ** 2 + (yy - 47.5) ** 2 < 44.0 ** 2  # an irregular (round) domain
    grid = StructuredGrid(bed=case.bed, manning_n=case.manning_n, domain_mask=mask, resolution_m=50.0)
    solver = ShallowWaterSolver(grid, rainfall_times_s=np.array([0.0, 1.0], np.float32),
                                rainfall_rates_m_s=np.zeros(2, np.float32), device=cuda)
    m = torch.as_tensor(mask, device=cuda)
    depth = torch.where(m, 0.02, 0.0).float()
    predicted = SWEState(depth, torch.zeros_like(depth), torch.zeros_like(depth))
    residual = torch.zeros(depth.shape + (3,), device=cuda)
    residual[..., 0] = torch.where(m, -0.01, 0.0)
    corrected = apply_residual_correction(predicted, residual, solver.params, correction_relaxation=1.0)
    np.testing.assert_allclose(float(corrected.h.double().sum()), float(predicted.h.double().sum()),
                               rtol=2.0e-6)
