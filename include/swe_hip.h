/*
 * swe_hip.h -- C ABI of libswe_hip.so, the MI355X (gfx950) explicit 2-D
 * shallow-water finite-volume step.
 *
 * The reference exposes this path as pure Python functions over NamedTuple
 * pytrees (no FFI exists there).  Each entry point below replaces one of
 * them; the Python package hybrid_flood_amd binds them with ctypes and keeps
 * the reference's names, argument meaning and exceptions.
 *
 * Conventions
 *   - Fields are row-major (ny, nx) float32 device arrays, x contiguous, row
 *     index increasing northward (jax_solver/shallow_water_2d.py:165-171).
 *   - domain_mask is a uint8 device array (0 inactive, 1 active); the
 *     reference stores it as bool (jax_solver/numerics.py:34).
 *   - boundary_types are ordered left, right, bottom, top with
 *     SWE_REFLECTIVE = 0 and SWE_OPEN = 1 (jax_solver/boundary_conditions.py:9-16).
 *   - Time, dt, the rain rate and the CFL quotient are float64; the state is
 *     float32.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     Calls are asynchronous on it unless documented otherwise.
 *   - Every entry point returns SWE_OK or an error code; swe_last_error()
 *     returns a thread-local message for the last failure.
 *   - A context is not thread-safe: use one per host thread / stream.
 */
#ifndef SWE_HIP_H
#define SWE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWE_OK 0
#define SWE_EINVAL 1      /* bad argument: shapes, times, pointers */
#define SWE_EMAXSTEPS 2   /* max_steps reached before the last output time */
#define SWE_EHIP 3        /* HIP runtime error */
#define SWE_ENONFINITE 4  /* a non-finite value appeared in the state */
#define SWE_ENODEV 5      /* no usable gfx950 device */

#define SWE_REFLECTIVE 0
#define SWE_OPEN 1

/* Kernel variants of the same step arithmetic.
 * SWE_VARIANT_REFERENCE: reference evaluation order, correctly rounded
 *   division and square root, no FMA contraction; bit-identical depth and
 *   wet mask with the FP32 oracle.
 * SWE_VARIANT_FAST: one reciprocal per quotient, hardware sqrt/log/exp,
 *   HLL velocities reused instead of re-divided; within the stated
 *   tolerance of the FP64 oracle. */
#define SWE_VARIANT_REFERENCE 0
#define SWE_VARIANT_FAST 1

typedef struct SweContext SweContext;

/* Conserved state (jax_solver/numerics.py:21-26): three device arrays. */
typedef struct SweState {
  float* h;
  float* hu;
  float* hv;
} SweState;

/* A cotangent of the state held in float64 (the adjoint's *_f64 entries). */
typedef struct SweStateF64 {
  double* h;
  double* hu;
  double* hv;
} SweStateF64;

/* Static fields and controls (jax_solver/numerics.py:29-45, built by
 * ShallowWaterSolver.__init__ at jax_solver/shallow_water_2d.py:213-253).
 * Device pointers must stay valid for the life of the context; the rainfall
 * series is copied at creation. */
typedef struct SweParamsDesc {
  int32_t nx, ny;
  const float* bed;                 /* device (ny, nx) */
  const float* manning_n;           /* device (ny, nx) */
  const uint8_t* domain_mask;       /* device (ny, nx) */
  const float* rainfall_multiplier; /* device (ny, nx) */
  int32_t boundary_types[4];        /* left, right, bottom, top */
  const float* rainfall_times_s;    /* host, n_rainfall, strictly increasing */
  const float* rainfall_rates_m_s;  /* host, n_rainfall */
  int32_t n_rainfall;               /* >= 1 */
  double default_rainfall_m_s;
  double dx, dy, gravity, cfl, dry_tolerance, max_dt;
} SweParamsDesc;

/* Result record of a run (ScanResult, jax_solver/numerics.py:48-55, plus the
 * FP64 mass balance of anuga_solver/run_simulation.py:100-107). */
typedef struct SweRunInfo {
  int32_t outputs_written;
  int32_t steps_used;
  double final_time_s;
  double volume_initial;   /* m^3 over active cells */
  double volume_final;
  double rain_volume;      /* rain applied to active cells */
  double boundary_volume;  /* signed inflow through open sides */
  double clip_volume;      /* created by the positivity clip max(h, 0) */
  double mass_balance_error;  /* V - V0 - rain - boundary - clip */
  double mass_balance_rel;    /* error / max(|V0|+|rain|+|boundary|+|clip|, 1) */
  int32_t nonfinite;       /* 1 if a NaN/Inf was seen in the state */
  int32_t _pad;
} SweRunInfo;

/* Library identity and the last error of the calling thread. */
const char* swe_version(void);
const char* swe_last_error(void);

/* Create a context for one grid (ShallowWaterSolver.__init__,
 * jax_solver/shallow_water_2d.py:213-253).  Validates shapes and controls,
 * copies the rainfall series to the device, allocates the per-block partials,
 * the device control block and the ping-pong work state. */
int swe_create(const SweParamsDesc* params, int32_t variant, SweContext** out);
int swe_destroy(SweContext* ctx);
int swe_set_variant(SweContext* ctx, int32_t variant);
/* Rows per wavefront strip of the fused step (tuning knob; 0 = default:
 * about sqrt(nx * ny) / 170, between 4 and 16). */
int swe_set_tile_rows(SweContext* ctx, int32_t rows);
/* Cap the fused step at `waves` resident wavefronts per compute unit by
 * reserving LDS per workgroup (tuning knob: fewer concurrent row streams can
 * raise HBM efficiency).  0 = no cap; else a multiple of 4, >= 8.  No
 * reference counterpart (the reference has no tiling). */
int swe_set_waves_per_cu(SweContext* ctx, int32_t waves);
/* Rows per wavefront strip of the adjoint step (tuning knob; 0 = default,
 * the same size rule as swe_set_tile_rows).  No
 * reference counterpart (the reference differentiates with jax.grad). */
int swe_set_adjoint_rows(SweContext* ctx, int32_t rows);
/* How the device loop closes a step (tuning knob).  1 (default) = fused:
 * every workgroup folds its CFL maximum into 64 slots by atomic max and
 * stores its FP64 mass partials; every workgroup of the next step derives dt
 * from the slots itself, and that step's workgroup 0 plans the step after and
 * accounts the one before (one launch per step, no finish kernel).  2 = every
 * workgroup stores its partials and a one-workgroup kernel reduces them and
 * plans the next step.  0 = the last workgroup of the step kernel does it
 * (one launch per step, but a release/acquire across workgroups).  All three
 * give the same states, step sizes and FP64 balance. */
int swe_set_loop_finish(SweContext* ctx, int32_t split);
/* Static record of the fast device loop (tuning knob).  With enable = 1
 * (default) a loop start whose rain multiplier is uniform over the active
 * cells and whose active cells carry at most 255 distinct Manning values
 * stores each cell as its bed elevation and a one-byte class (5 bytes), the
 * step reading n^2 from a 255-entry table; otherwise (bed, n^2) with the
 * multiplier as a constant (8 bytes) when it is uniform, else (bed, n^2,
 * multiplier) (12 bytes).  All give bit-identical states.  Refused while a
 * device loop is active.  No reference counterpart: the reference keeps the
 * Manning raster as a dense float array (src/hybrid_flood/jax_solver/
 * numerics.py:33, applied at :359), rasterized from the roughness of a
 * land-use class layer (src/hybrid_flood/structured_io.py:103-113, lookup
 * src/hybrid_flood/anuga_solver/build_domain.py:187-205), so it holds a few
 * distinct values. */
int swe_set_static_classes(SweContext* ctx, int32_t enable);
/* The static record the last fast device loop start chose: 2 = bed + class,
 * 1 = (bed, n^2), 0 = (bed, n^2, multiplier).  Synchronises the device (every
 * stream) before it reads the record.  No reference counterpart. */
int swe_get_static_layout(SweContext* ctx, int32_t* layout);
/* Strip order of the device loop (tuning knob): 1 (default) = odd steps
 * sweep the strips in reverse, so each step starts where the previous one
 * ended and reads what is still in the memory-side cache; 0 = every step
 * sweeps in the same order.  States are bit-identical either way.  No
 * reference counterpart (XLA schedules the reference's step). */
int swe_set_sweep_order(SweContext* ctx, int32_t alternate);
/* Rounding of the rain source in the device loops (swe_run_to_outputs,
 * swe_advance_to_time, swe_run_steps, swe_loop_*): 0 (default) = round to
 * nearest, as the reference; 1 = stochastic: a step with rain adds it to each
 * cell's depth rounded up or down to a float neighbour with the probabilities
 * that make the expected depth exact, drawn from a hash of (cell index, step
 * index) -- deterministic and independent of the tile.  Round-to-nearest adds
 * the same error on every ponded cell of a binade, which is the float32
 * state's drift of the FP64 mass balance on long rainy runs; stochastic
 * rounding decorrelates those errors.  swe_step and the replays of the
 * adjoint always round to nearest, so it cannot be combined with
 * swe_set_checkpoints (SWE_EINVAL either way round).  No reference
 * counterpart. */
int swe_set_rain_rounding(SweContext* ctx, int32_t mode);

/* One step with caller-given dt and time (finite_volume_step,
 * jax_solver/numerics.py:312-369).  `in` and `out` must not alias. */
int swe_step(SweContext* ctx, SweState in, SweState out, double dt, double time_s,
             void* stream);

/* CFL step of a state (compute_cfl_dt, jax_solver/numerics.py:297-309).
 * Synchronises the stream and writes the float64 dt to *dt_out. */
int swe_cfl_dt(SweContext* ctx, SweState in, double* dt_out, void* stream);

/* Per-cell inverse CFL time (|u|+c)/dx + (|v|+c)/dy, 0 outside the mask
 * (jax_solver/numerics.py:299-303), into a device float array. */
int swe_cfl_field(SweContext* ctx, SweState in, float* inv_dt_out, void* stream);

/* Rainfall rate at time_s, float64 linear interpolation with the default
 * outside the series (rainfall_rate, jax_solver/numerics.py:286-294; the
 * float64 form of numpy_solver/shallow_water_2d.py:254-264).  Host only. */
int swe_rainfall_rate(const SweContext* ctx, double time_s, double* rate_out);
int swe_interp_rate(const float* times_s, const float* rates_m_s, int32_t n, double dflt,
                    double time_s, double* rate_out);

/* Adaptive loop to fixed output times (integrate_to_outputs,
 * jax_solver/numerics.py:372-434; ShallowWaterSolver.run,
 * jax_solver/shallow_water_2d.py:278-306).  The time loop runs on the device:
 * each step kernel reduces the CFL maximum of its new state and its last
 * workgroup sets the next dt, time, output slot and stop flag; the host only
 * launches chunks of steps and polls the stop flag between chunks.
 *   output_times_s: host float64[n_times], n_times >= 1, finite (else
 *                   SWE_EINVAL); times[0] is the initial clock.  A target the
 *                   clock has already passed is reached at once, as in the
 *                   reference's scan; ShallowWaterSolver.run additionally
 *                   requires times to start at 0 and increase strictly.
 *   out_h/out_hu/out_hv: device float32[n_times][ny][nx]; slot 0 receives
 *                   the initial state; slots the loop did not reach receive
 *                   the last state (as the reference's scan stacks it).
 *   final_state:    optional device state receiving the last state.
 * Synchronises the stream.  Returns SWE_EMAXSTEPS (with `info` filled) when
 * max_steps ended the loop early, SWE_ENONFINITE if the state went non-finite. */
int swe_run_to_outputs(SweContext* ctx, SweState initial, const double* output_times_s,
                       int32_t n_times, int32_t max_steps, float* out_h, float* out_hu,
                       float* out_hv, SweState final_state, SweRunInfo* info, void* stream);

/* Advance adaptively from t_start to t_target in at most max_substeps CFL
 * steps (hybrid/coupled_forecast.py:123-161, advance_physics_to_time): while
 * t < t_target - 1e-7, dt = min(CFL dt, t_target - t).  `out` receives the
 * final state, or NaN everywhere when the clock ended more than 1e-6 s short
 * of t_target (coupled_forecast.py:152-156); `in` and `out` may alias.
 * *steps_used / *t_final (either may be NULL) receive the step count and the
 * float64 clock.  Runs the device loop; synchronises the stream.  A target at
 * or before t_start takes no step. */
int swe_advance_to_time(SweContext* ctx, SweState in, double t_start, double t_target,
                        int32_t max_substeps, SweState out, int32_t* steps_used, double* t_final,
                        void* stream);

/* Exactly n_steps adaptive steps (dt = CFL dt capped at max_dt, no output
 * target) from time t0, state updated in place.  The device loop of
 * swe_run_to_outputs without outputs; used by the throughput benchmark
 * (jax_solver/benchmarking.py:17-97 times a step loop the same way).
 * Asynchronous unless `info` is non-NULL. */
int swe_run_steps(SweContext* ctx, SweState state, int32_t n_steps, double t0,
                  SweRunInfo* info, void* stream);

/* The device loop of swe_run_steps in three parts, for callers that drive it
 * themselves (a throughput harness brackets swe_loop_advance with events; a
 * coupled model interleaves its own work, as hybrid/coupled_forecast.py:123-161
 * does between fixed-length substep loops).
 *   swe_loop_begin:   copy `state` into the context's work buffers, reduce its
 *                     CFL maximum and volume, prime the control block at t0
 *                     with a budget of max_steps steps (no output target).
 *   swe_loop_advance: queue exactly n_launches step kernels (whole graph
 *                     chunks where they fit, plain launches for the rest);
 *                     launches after the budget is spent are no-ops.
 *   swe_loop_end:     copy the current state back to `state`; with `info`
 *                     non-NULL, synchronise and fill it.
 * All three are asynchronous on `stream` unless `info` is given. */
int swe_loop_begin(SweContext* ctx, SweState state, double t0, int32_t max_steps, void* stream);
int swe_loop_advance(SweContext* ctx, int32_t n_launches, void* stream);
int swe_loop_end(SweContext* ctx, SweState state, SweRunInfo* info, void* stream);

/* Record the device loop's step sizes: the float64 dt of step k (counted
 * from the loop's start) is written to dt_log[k] for k < capacity.  The
 * schedule lets a caller replay the exact same steps with swe_step (the
 * adjoint recomputes its checkpoints that way).  capacity 0 turns it off. */
int swe_set_dt_log(SweContext* ctx, double* dt_log, int32_t capacity);

/* Record checkpoints during the device loop: while set, the state after
 * s * every steps (s = 1, 2, ...; s < capacity) is written to slot s of the
 * caller's arrays h, hu, hv (each capacity x nx*ny floats), except when that
 * step also writes an output-time slot of swe_run_to_outputs, whose cube then
 * holds the same state.  every = 0 or capacity = 0 turns it off.  No reference
 * counterpart: the reference differentiates with jax.grad; the checkpointed
 * adjoint uses it to skip a replay of the forward. */
int swe_set_checkpoints(SweContext* ctx, float* h, float* hu, float* hv, int32_t every,
                        int32_t capacity);

/* Use HIP graphs of `chunk` captured step launches for the device loop
 * (0 = plain launches).  Default 64. */
int swe_set_graph_chunk(SweContext* ctx, int32_t chunk);

/* Adjoint of one step (no counterpart in the reference: it never
 * differentiates its solver; this is the reverse of finite_volume_step,
 * jax_solver/numerics.py:312-369, with dt treated as data).  Given the state
 * `state` entering a step of size dt at time_s and the cotangent `lam_next`
 * of the state it produces, writes the cotangent `lam_out` of `state` and
 * ADDS this step's sensitivities to the per-cell Manning field and rain
 * multiplier into grad_manning / grad_mult (either may be NULL).  lam_out
 * must not alias lam_next.  Asynchronous on `stream`. */
int swe_adjoint_step(SweContext* ctx, SweState state, SweState lam_next, SweState lam_out,
                     double dt, double time_s, float* grad_manning, float* grad_mult,
                     void* stream);

/* Replay a recorded schedule (dts[k], ts[k]: float64 step size and start
 * time of step k, e.g. from swe_set_dt_log): `out` receives the state after
 * n steps from `in` (they may alias).  Bit-identical to the device loop's
 * steps.  swe_replay_store keeps every intermediate state: slot k of the
 * (n, ny, nx) cubes is the state after step k.  Asynchronous. */
int swe_replay(SweContext* ctx, SweState in, const double* dts, const double* ts, int32_t n,
               SweState out, void* stream);
int swe_replay_store(SweContext* ctx, SweState in, const double* dts, const double* ts, int32_t n,
                     float* cube_h, float* cube_hu, float* cube_hv, void* stream);

/* Reverse n consecutive steps: states_* hold the n states entering steps
 * 0..n-1 of the segment as (n, ny, nx) cubes, dts/ts their schedule.  `lam`
 * holds the cotangent of the state after the segment on entry and of the
 * state entering it on return (lam_tmp is scratch of the same shape);
 * sensitivities are added into grad_manning / grad_mult.  Asynchronous. */
int swe_adjoint_sweep(SweContext* ctx, const float* states_h, const float* states_hu,
                      const float* states_hv, int32_t n, const double* dts, const double* ts,
                      SweState lam, SweState lam_tmp, float* grad_manning, float* grad_mult,
                      void* stream);

/* swe_adjoint_step / swe_adjoint_sweep with the cotangents, the reverse
 * arithmetic and the sensitivities in float64 (the forward recompute stays
 * the float32 kernel's, branch for branch) for callers that want the
 * reverse sweep free of float32 rounding; gfx950 runs float64 add/mul/FMA at
 * the float32 rate.  No reference counterpart (the reference has no
 * adjoint). */
int swe_adjoint_step_f64(SweContext* ctx, SweState state, SweStateF64 lam_next,
                         SweStateF64 lam_out, double dt, double time_s, double* grad_manning,
                         double* grad_mult, void* stream);
int swe_adjoint_sweep_f64(SweContext* ctx, const float* states_h, const float* states_hu,
                          const float* states_hv, int32_t n, const double* dts, const double* ts,
                          SweStateF64 lam, SweStateF64 lam_tmp, double* grad_manning,
                          double* grad_mult, void* stream);

/* ------------------------------------------------------------------------
 * Batched ensemble (no counterpart in the reference, which runs one domain;
 * it runs many independent rainfall/roughness scenarios of the same step).
 * Each member is a full domain of its own size with its own fields, rain
 * series and controls.  One launch (one wavefront per strip of every
 * running member) advances every unfinished member by one of its own CFL
 * steps; chunks of launches are captured in a HIP graph.
 * ------------------------------------------------------------------------ */
typedef struct SweEnsemble SweEnsemble;

typedef struct SweMemberDesc {
  SweParamsDesc params; /* as for swe_create; device fields stay valid for the ensemble's life */
  SweState state;       /* device: initial state in (read by swe_ensemble_begin), final state out */
} SweMemberDesc;

int swe_ensemble_create(const SweMemberDesc* members, int32_t n_members, int32_t variant,
                        SweEnsemble** out);
int swe_ensemble_destroy(SweEnsemble* ens);
/* Rows per wavefront strip for every member (tuning knob). */
int swe_ensemble_set_tile_rows(SweEnsemble* ens, int32_t rows);
/* Launches per captured HIP graph (0 = plain launches). */
int swe_ensemble_set_graph_chunk(SweEnsemble* ens, int32_t chunk);
/* Copy every member's state in and prime its clock at 0 with target t_end. */
int swe_ensemble_begin(SweEnsemble* ens, double t_end, int32_t max_steps, void* stream);
/* Queue n_launches ensemble steps (rounded up to the graph chunk). */
int swe_ensemble_advance(SweEnsemble* ens, int32_t n_launches, void* stream);
/* Synchronise and count members still running. */
int swe_ensemble_poll(SweEnsemble* ens, int32_t* running, void* stream);
/* Drop finished members from the launch (and shorten strips when few cells
 * remain).  Synchronises; call between chunks.  swe_ensemble_run does this
 * whenever a member finishes. */
int swe_ensemble_compact(SweEnsemble* ens, void* stream);
/* Copy final states out; fill infos[n_members]; SWE_EMAXSTEPS if a member
 * did not reach t_end, SWE_ENONFINITE if one went non-finite.  Synchronises. */
int swe_ensemble_end(SweEnsemble* ens, SweRunInfo* infos, void* stream);
/* begin + advance/poll until every member reaches t_end + end. */
int swe_ensemble_run(SweEnsemble* ens, double t_end, int32_t max_steps, SweRunInfo* infos,
                     void* stream);
/* Strips (wavefronts of work) per launch and the launch's workgroup count. */
int swe_ensemble_total_waves(const SweEnsemble* ens, int32_t* waves, int32_t* grid);
/* The static record member `member`'s last start chose (fast variant; the
 * codes of swe_get_static_layout: 2 = bed + Manning class, 1 = (bed, n^2),
 * 0 = (bed, n^2, multiplier)).  Synchronises the device before it reads the
 * record.  No reference counterpart. */
int swe_ensemble_get_static_layout(SweEnsemble* ens, int32_t member, int32_t* layout);
/* Rounding of every member's rain source (swe_set_rain_rounding), drawn for
 * each member's own step index.  Not while a run is active.  No reference
 * counterpart. */
int swe_ensemble_set_rain_rounding(SweEnsemble* ens, int32_t mode);

/* ---- Hybrid coupled forecast (hybrid/coupled_forecast.py) ---------------- */

/* network_features (coupled_forecast.py:53-79): the six normalised CNN input
 * channels of `state` -- (h, u, v) standardised with state_mean/state_std
 * (host float32[3]) and zeroed outside the domain mask, the rain channel
 * float32(rate(time_s) * 3.6e6) / rainfall_scale_mm_hr (rate from the
 * context's series, as swe_rainfall_rate), terrain_std and roughness_std
 * (device float32[ny][nx]).  Channel c of cell k (row-major) is written to
 * features[c * channel_stride + k * cell_stride]: NHWC is channel_stride 1,
 * cell_stride 6; NCHW is channel_stride N, cell_stride 1.  Asynchronous. */
int swe_network_features(SweContext* ctx, SweState state, const float* terrain_std,
                         const float* roughness_std, const float* state_mean,
                         const float* state_std, double rainfall_scale_mm_hr, double time_s,
                         float* features, int64_t channel_stride, int64_t cell_stride,
                         void* stream);

/* apply_residual_correction (coupled_forecast.py:82-120) with the residual
 * clip of coupled_forecast.py:199-201 fused in.  residual channel c of cell
 * k is residual[c * channel_stride + k * cell_stride]; each channel is
 * clipped to +-clip_limits[c] (host float32[3]; NULL = no clip).  The
 * provisional depth max(h + relaxation*r0, 0) (0 outside the mask) is
 * rescaled to the masked physics volume with float64 sums; velocities are
 * corrected by relaxation*r1, relaxation*r2 and re-multiplied where wet.
 * `corrected` may alias `predicted`.  volumes (optional host double[4])
 * receives (physics volume, provisional volume, scale, 1 if rescaled) in
 * cell-depth units and synchronises the stream; otherwise asynchronous. */
int swe_apply_residual_correction(SweContext* ctx, SweState predicted, const float* residual,
                                  int64_t channel_stride, int64_t cell_stride, float relaxation,
                                  const float* clip_limits, SweState corrected, double* volumes,
                                  void* stream);

/* swe_apply_residual_correction without any device allocation, for callers
 * that cannot allocate: code under HIP graph capture, tight serving loops
 * that alternate the CNN and the physics, harnesses that forbid hipMalloc.
 * No counterpart in the reference (jax allocates inside its jitted call).
 * Same arguments, kernels and bits as swe_apply_residual_correction; its
 * float64 volume partials live in a static device array of 4 + 2 x 2048
 * doubles (32.8 KB, held by every device the library runs on until the
 * process ends) instead of the context's lazily allocated scratch.  The
 * array is resolved per device by swe_create.  Returns SWE_EINVAL if the
 * grid would need more than 2048 partials (it never does: the partial count
 * is capped at 2048 for any grid).  Every call on one device shares that
 * array, so calls that overlap on one device (two streams) race; calls in
 * one stream are ordered.  With `volumes` non-NULL the call synchronises the
 * stream (do not capture it then); otherwise it is asynchronous. */
int swe_apply_residual_correction_noalloc(SweContext* ctx, SweState predicted,
                                          const float* residual, int64_t channel_stride,
                                          int64_t cell_stride, float relaxation,
                                          const float* clip_limits, SweState corrected,
                                          double* volumes, void* stream);

/* Feature staging of `count` primitive states (comparison/v2_staging.py:17-44):
 * state_nhwc is device float32 [count][ny][nx][3] (depth, u, v); output
 * channels (standardised state, masked; rain_channel[s] (device float32
 * [count], the sample's rain / scale); terrain_std; roughness_std) of sample
 * s, cell k, channel c go to features[s*sample_stride + c*channel_stride +
 * k*cell_stride].  No context needed.  Asynchronous. */
int swe_stage_features(int32_t count, int32_t ny, int32_t nx, const float* state_nhwc,
                       const uint8_t* domain_mask, const float* terrain_std,
                       const float* roughness_std, const float* state_mean, const float* state_std,
                       const float* rain_channel, float* features, int64_t sample_stride,
                       int64_t channel_stride, int64_t cell_stride, void* stream);

/* out[s][k] = max(depth[s][k] + residual channel 0, 0) for `count` samples
 * (comparison/v2_staging.py:47-52); the residual of (s, k) is read at
 * residual[s*sample_stride + k*cell_stride], so an NCHW network output is
 * used in place.  Asynchronous. */
int swe_correct_depth(int32_t count, int32_t ny, int32_t nx, const float* depth,
                      const float* residual, int64_t sample_stride, int64_t cell_stride,
                      float* out, void* stream);

/* ---- Measurement ---------------------------------------------------------- */

/* Streaming probe of the achievable HBM bandwidth (no counterpart in the
 * reference, which never measured a device; the benchmark runs it beside the
 * step).  Asynchronous on `stream`; time it with events.
 *   mode 0: copy n floats src -> dst (n % 4 == 0, 16-byte aligned);
 *   mode 1: read n floats of src (n % 4 == 0, 16-byte aligned), write one
 *           float per 1024 read to dst (ceil(n / 1024) floats);
 *   mode 2: the packed device loop's streams without its arithmetic: n cells,
 *           src = 3n floats of packed state, src2 = 2n floats of packed static
 *           fields, dst = 3n floats written;
 *   mode 3: as mode 2 with the class-indexed static record: src2 = ceil(n / 64)
 *           blocks of 320 bytes (64 float beds, then 64 one-byte classes);
 *   mode 4: clock sample (src, src2 unused): writes the shader clock counter
 *           and the constant-rate real-time counter (hipDeviceAttributeWallClockRate)
 *           as two uint64 to dst slot n - 1 (dst 8-byte aligned, n >= 1); two
 *           samples around other work on the stream give its average shader
 *           clock. */
int swe_stream_probe(int32_t mode, const void* src, const void* src2, void* dst, int64_t n,
                     void* stream);

/* ---- Output schema (jax_solver/shallow_water_2d.py:309-350,
 *      structured_io.py:160-211) ------------------------------------------ */

/* Primitive fields of n_times stacked states (device float32
 * [n_times][ny][nx] each): depth = h; x/y velocity = hu/h, hv/h where
 * h > dry_tolerance, else 0; velocity = hypot(u, v); all four NaN outside the
 * domain mask.  elevation (optional, device float32[ny][nx]) = bed, NaN
 * outside the mask.  Asynchronous. */
int swe_conserved_to_primitive(SweContext* ctx, const float* h, const float* hu, const float* hv,
                               int32_t n_times, float* depth, float* x_velocity,
                               float* y_velocity, float* velocity, float* elevation,
                               void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SWE_HIP_H */
