"""MI355X-native explicit 2-D shallow-water step (drop-in for the reference's
``hybrid_flood.jax_solver`` step/run entry points).

Modules mirror the reference layout:
  numerics            SWEState, SWEParams, ScanResult, finite_volume_step,
                      compute_cfl_dt, rainfall_rate, integrate_to_outputs
  boundary_conditions REFLECTIVE, OPEN, BoundaryConfig, all_reflective,
                      all_open, apply_ghost_cells
  shallow_water_2d    StructuredGrid, ShallowWaterSolver
  benchmarking        benchmark_compiled_steps
  synthetic           seeded DEM / rainfall generators of the benchmark cases
"""

__version__ = "0.1.0"
