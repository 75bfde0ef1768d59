// Device-loop control shared by the single-domain and ensemble kernels: the
// control block each step kernel reads on entry, the FP64 partials of its
// epilogue, and the single-thread finish that plans the next step
// (jax_solver/numerics.py:398-418 with the float64 scalars of
// numpy_solver/shallow_water_2d.py:267-274, 402-412).
#pragma once

#include <cmath>

#include "swe_device.hpp"

namespace swe {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kDefaultRows = 16;

// Strip height when none is chosen: about sqrt(cells) / 170 rows, between 4
// and kDefaultRows.  Shorter strips give small grids enough wavefronts to
// hide memory latency, at the price of more halo rows per cell; measured on
// MI355X, the best heights were 6 rows at 1024^2, 12 at 2048^2 and 16-17 at
// 4096^2, for the forward step and for the adjoint step alike.
inline int default_rows(int nx, int ny) {
  const double r = std::sqrt((double)nx * (double)ny) / 170.0 + 0.5;
  return r < 4.0 ? 4 : (r > (double)kDefaultRows ? kDefaultRows : (int)r);
}

// Plan of one step of the fused device loop (MODE_LOOP_FUSED): what every
// workgroup of the step reads on entry, besides the CFL maximum of the state,
// to derive the step's dt, rain rate and output slots itself.  Written by
// workgroup 0 of the step before (or by the loop start), double-buffered by
// step parity.
struct StepPlan {
  double t;          // clock at the start of the step
  double rate;       // rain rate at t (float64 interpolation)
  double tw[4];      // times[oi], times[oi+1], ... (+inf past the last output time)
  double prev_dt;    // dt of the step before (for its FP64 accounting)
  float prev_dtrate; // float32(dt * rate) of the step before
  int steps;         // steps taken before this one (-1: no such step)
  int oi;            // first output slot not reached at t
  int done;          // every output reached or max_steps taken: the loop stops here
};

// Device-resident state of the time loop.  The first fields are what every
// step kernel reads on entry.
struct LoopControl {
  float dt_f, dtrate_f, dtg_f;  // float32(dt), float32(dt*rate), float32(dt*g)
  float dtx_f;                  // float32(dt/dx), the fast variant's divergence scale
  int done;
  int rec_lo, rec_hi;  // output slots the pending step writes
  int ck_slot;         // checkpoint slot the pending step writes (-1: none)
  int steps, out_idx;
  double t, dt, rate;
  double vol0, vol, rain, bnd, clip;
  double max_inv;
  int nonfinite;
  unsigned counter;  // workgroups finished in the current launch
  StepPlan plan[2];  // fused loop: plans of the steps of parity 0 and 1
  int acc_steps;     // fused loop: steps whose FP64 accounting is in the fields above
  int _pad;
};

struct LoopConst {
  double cfl, max_dt, gravity, cell_area, rain_default, dxd, dyd;
  // FP64 sum of the rain multiplier over active cells (device-resident, so a
  // loop start can refresh it without a host round trip).  The single-domain
  // context also keeps [1] = 1 if the multiplier has one value on every
  // active cell and [2] = that value, read by the packed device loop.
  const double* sum_mult;
  const double* rain_t;
  const double* rain_r;
  const double* times;
  double* dt_log;  // optional: dt of step k is written to dt_log[k] (k < dt_log_cap)
  // Optional checkpoints: the state after s * ck_every steps goes to slot s of
  // [ck_h | ck_hu | ck_hv] (each ck_cap x N), unless that step also writes an
  // output slot (the output cube then holds it).
  float* ck_h;
  float* ck_hu;
  float* ck_hv;
  int n_rain, n_times, max_steps, dt_log_cap;
  int ck_every, ck_cap;
  int record;      // 0: output times only stop/clamp the loop, no state is stored
  int rain_sr;     // 1: a step with rain adds it with stochastic rounding (sr_add)
};

struct Partials {
  float* inv;
  double* vol;
  double* clip;
  double* bnd;
  int* nf;
  // Fused loop: the partials of a step of parity p are at [p * stride, ...),
  // and the CFL maximum of the state a step produces goes to the 64 slots
  // cfl_slots[64 (1 - p) ...] as (steps after it) << 32 | float bits,
  // 0xFFFFFFFF for a non-finite state.
  unsigned long long* cfl_slots;
  int stride;
};

constexpr int kCflSlots = 64;
constexpr unsigned kNonfiniteBits = 0xFFFFFFFFu;

// MODE_STEP: caller-given dt, no epilogue.  MODE_LOOP: device loop, the last
// workgroup finishes the step.  MODE_LOOP_SPLIT: device loop, every
// workgroup stores its partials and a one-workgroup kernel finishes the step.
// MODE_LOOP_FUSED: device loop without a finish kernel: each workgroup
// folds its CFL maximum into the step's slots with an atomic max; every
// workgroup of the next step reads them and derives dt itself, and its
// workgroup 0 plans the step after and does the FP64 accounting of the one
// before (k_close accounts the last step when the loop ends).
enum { MODE_STEP = 0, MODE_LOOP = 1, MODE_LOOP_SPLIT = 2, MODE_LOOP_FUSED = 3 };
enum { RED_CFL = 0, RED_INIT = 1 };

template <class T>
__device__ __forceinline__ T coherent_load(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Prepare the control block for the step that starts at t_new with the CFL
// maximum m of the current state (numerics.py:405 with the float64 quotient
// of numpy_solver/shallow_water_2d.py:274).  Single thread.
__device__ void plan_next_step(LoopControl* ctl, const LoopConst& L, float m, double t_new,
                               int oi, int rain_j) {
  const double md = (double)m;
  const double cfl_dt = md > 0.0 ? fmin(L.cfl / md, L.max_dt) : L.max_dt;
  const double target = L.times[oi];
  const double dt = fmin(cfl_dt, target - t_new);
  const double t_end = t_new + dt;
  int hi = oi;
  while (hi < L.n_times && t_end >= L.times[hi] - kOutputSlack) ++hi;
  const double rate = interp_at(L.rain_t, L.rain_r, L.n_rain, L.rain_default, t_new, rain_j);
  ctl->dt = dt;
  ctl->rate = rate;
  ctl->dt_f = (float)dt;
  ctl->dtrate_f = (float)(dt * rate);
  ctl->dtg_f = (float)(dt * L.gravity);
  ctl->dtx_f = (float)(dt / L.dxd);
  ctl->rec_lo = oi;
  ctl->rec_hi = L.record ? hi : oi;
  const int after = ctl->steps + 1;
  const int slot = L.ck_every > 0 ? after / L.ck_every : 0;
  ctl->ck_slot = (L.ck_every > 0 && after % L.ck_every == 0 && slot < L.ck_cap &&
                  ctl->rec_lo == ctl->rec_hi)
                     ? slot
                     : -1;
  if (L.dt_log && ctl->steps < L.dt_log_cap) L.dt_log[ctl->steps] = dt;
}

// Block-wide reduction of (max, sum, sum, sum, or) into thread 0.
struct Totals {
  float inv;
  double vol, clip, bnd;
  int nf;
};

template <int WAVES = kWavesPerBlock>
__device__ __forceinline__ Totals block_reduce(float inv, double vol, double clip, double bnd, int nf) {
  __shared__ float s_inv[WAVES];
  __shared__ double s_vol[WAVES], s_clip[WAVES], s_bnd[WAVES];
  __shared__ int s_nf[WAVES];
  inv = wave_max(inv);
  // Sums that are zero on every lane (the clip and open-boundary volumes of
  // most wavefronts, the volume where the caller does not sum it) skip the
  // shuffles: a uniform branch on one comparison.
  if (__builtin_amdgcn_ballot_w64(vol != 0.0) != 0) vol = wave_sum(vol);
  if (__builtin_amdgcn_ballot_w64(clip != 0.0) != 0) clip = wave_sum(clip);
  if (__builtin_amdgcn_ballot_w64(bnd != 0.0) != 0) bnd = wave_sum(bnd);
  nf = wave_any(nf);
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    s_inv[w] = inv;
    s_vol[w] = vol;
    s_clip[w] = clip;
    s_bnd[w] = bnd;
    s_nf[w] = nf;
  }
  __syncthreads();
  Totals t{0.0f, 0.0, 0.0, 0.0, 0};
  if (threadIdx.x == 0) {
    for (int k = 0; k < WAVES; ++k) {
      t.inv = fmaxf(t.inv, s_inv[k]);
      t.vol += s_vol[k];
      t.clip += s_clip[k];
      t.bnd += s_bnd[k];
      t.nf |= s_nf[k];
    }
  }
  __syncthreads();
  return t;
}

// Publish this workgroup's partials; returns true (block-uniform) in the last
// workgroup of the launch, after which every workgroup's partials are visible
// to it (release on the counter increment, acquire in the last workgroup).
__device__ bool publish_partials(const Partials& P, LoopControl* ctl, const Totals& t) {
  __shared__ int s_last;
  if (threadIdx.x == 0) {
    const int b = blockIdx.x;
    __hip_atomic_store(P.inv + b, t.inv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(P.vol + b, t.vol, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(P.clip + b, t.clip, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(P.bnd + b, t.bnd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(P.nf + b, t.nf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned prev = __hip_atomic_fetch_add(&ctl->counter, 1u, __ATOMIC_RELEASE,
                                                 __HIP_MEMORY_SCOPE_AGENT);
    s_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (s_last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return s_last != 0;
}

// Sum all partials (last workgroup only) into thread 0.
__device__ Totals gather_partials(const Partials& P, int nparts) {
  float inv = 0.0f;
  double vol = 0.0, clip = 0.0, bnd = 0.0;
  int nf = 0;
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
    inv = fmaxf(inv, coherent_load(P.inv + k));
    vol += coherent_load(P.vol + k);
    clip += coherent_load(P.clip + k);
    bnd += coherent_load(P.bnd + k);
    nf |= coherent_load(P.nf + k);
  }
  return block_reduce(inv, vol, clip, bnd, nf);
}

// Index j with rain_t[j] <= t < rain_t[j+1], searched by the whole block.
__device__ int rain_bracket(const LoopConst& L, double t) {
  __shared__ int s_j;
  if (threadIdx.x == 0) s_j = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < L.n_rain - 1; k += blockDim.x) {
    if (L.rain_t[k] <= t && t < L.rain_t[k + 1]) s_j = k;
  }
  __syncthreads();
  return s_j;
}

// FP64 accounting of one finished step (thread 0): rain, open-boundary and
// clip volumes, the volume after it and the non-finite flag.  dt and dtrate
// are the step's own (finish_step and the fused loop share this arithmetic).
__device__ __forceinline__ void account_step(LoopControl* ctl, const LoopConst& L,
                                             const Totals& tot, double dt, float dtrate) {
  ctl->rain += (double)dtrate * *L.sum_mult * L.cell_area;
  ctl->bnd += dt * tot.bnd;
  ctl->clip += tot.clip * L.cell_area;
  ctl->vol = tot.vol * L.cell_area;
  ctl->nonfinite |= tot.nf;
}

// Close the step just taken (single thread): account its rain, open-boundary
// and clip volumes in FP64, advance the clock and step count, move past the
// output times it reached (numerics.py:398-412), then stop or plan the next
// step.  Resets the workgroup counter for the next launch.
__device__ inline void finish_step(LoopControl* ctl, const LoopConst& L, const Totals& tot,
                                   double t_new, int rain_j) {
  account_step(ctl, L, tot, ctl->dt, ctl->dtrate_f);
  ctl->max_inv = (double)tot.inv;
  const int steps = ctl->steps + 1;
  ctl->steps = steps;
  ctl->t = t_new;
  int oi = ctl->out_idx;
  while (oi < L.n_times && t_new >= L.times[oi] - kOutputSlack) ++oi;
  ctl->out_idx = oi;
  if (oi >= L.n_times || steps >= L.max_steps || ctl->nonfinite) {
    ctl->done = 1;
    ctl->rec_lo = ctl->rec_hi = 0;
    ctl->ck_slot = -1;
  } else {
    plan_next_step(ctl, L, tot.inv, t_new, oi, rain_j);
  }
  __hip_atomic_store(&ctl->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block-wide totals of the partials [first, first + count) into thread 0.
// Thread t adds entries t, t + BLOCK, t + 2 BLOCK, ... in that order whatever
// UNROLL is (UNROLL only sets how many loads are in flight), so every caller
// sums in the same order and gets the same bits.  INV: also the float CFL
// maximum (the fused loop keeps it in cfl_slots instead).
template <int BLOCK, int UNROLL, bool INV, class Parts>
__device__ Totals reduce_partials(const Parts& P, int first, int count) {
  const int tid = threadIdx.x;
  float inv = 0.0f;
  double vol = 0.0, clip = 0.0, bnd = 0.0;
  int nf = 0;
  for (int base = tid; base < count; base += BLOCK * UNROLL) {
    float pi[UNROLL];
    double pv[UNROLL], pc[UNROLL], pb[UNROLL];
    int pn[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      const int k = base + u * BLOCK;
      const bool in = k < count;
      const int q = first + (in ? k : 0);
      pi[u] = (INV && in) ? P.inv[q] : 0.0f;
      pv[u] = in ? P.vol[q] : 0.0;
      pc[u] = in ? P.clip[q] : 0.0;
      pb[u] = in ? P.bnd[q] : 0.0;
      pn[u] = in ? P.nf[q] : 0;
    }
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      inv = fmaxf(inv, pi[u]);
      vol += pv[u];
      clip += pc[u];
      bnd += pb[u];
      nf |= pn[u];
    }
  }
  return block_reduce<BLOCK / kWave>(inv, vol, clip, bnd, nf);
}

// ---- fused device loop -------------------------------------------------------

// The CFL maximum of the state entering step `steps`, as float bits
// (kNonfiniteBits for a non-finite state), from the 64 slots that the step
// before filled.  Slots of an older step (smaller tag) count as 0.  Called by
// every lane of a wavefront; the result is wave-uniform.
__device__ __forceinline__ unsigned cfl_bits(const unsigned long long* slots, int steps) {
  const int lane = threadIdx.x & (kWave - 1);
  const unsigned long long w =
      __hip_atomic_load(slots + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned b = ((unsigned)(w >> 32) == (unsigned)steps) ? (unsigned)w : 0u;
  for (int o = kWave / 2; o > 0; o >>= 1) b = max(b, (unsigned)__shfl_xor((int)b, o, kWave));
  return __builtin_amdgcn_readfirstlane(b);
}

// What a workgroup of the fused loop derives on entry from the step's plan
// and the CFL maximum m of its state: the same float64 arithmetic as
// plan_next_step (numerics.py:405 with the float64 quotient of
// numpy_solver/shallow_water_2d.py:274), so every workgroup gets the same dt.
struct StepGo {
  double dt;
  float dt_f, dtrate_f, dtg_f, dtx_f, m;
  int rec_lo, rec_hi, ck_slot;
  bool stop, nonfinite;
};

__device__ __forceinline__ StepGo step_go(const StepPlan& pl, unsigned bits, const LoopConst& L) {
  StepGo g;
  g.nonfinite = bits == kNonfiniteBits;
  g.stop = pl.done != 0 || g.nonfinite;
  g.m = __uint_as_float(g.nonfinite ? 0u : bits);
  g.dt = 0.0;
  g.dt_f = g.dtrate_f = g.dtg_f = g.dtx_f = 0.0f;
  g.rec_lo = g.rec_hi = 0;
  g.ck_slot = -1;
  if (g.stop) return g;
  const double md = (double)g.m;
  const double cfl_dt = md > 0.0 ? fmin(L.cfl / md, L.max_dt) : L.max_dt;
  const double dt = fmin(cfl_dt, pl.tw[0] - pl.t);
  const double t_end = pl.t + dt;
  // Output times the step reaches: the plan's window of four, then (rare:
  // more than four output times in one step) the series itself.
  int hi = pl.oi;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (hi == pl.oi + k && hi < L.n_times && t_end >= pl.tw[k] - kOutputSlack) ++hi;
  if (hi == pl.oi + 4)
    while (hi < L.n_times && t_end >= L.times[hi] - kOutputSlack) ++hi;
  g.rec_lo = pl.oi;
  g.rec_hi = L.record ? hi : pl.oi;
  const int after = pl.steps + 1;
  const int slot = L.ck_every > 0 ? after / L.ck_every : 0;
  g.ck_slot = (L.ck_every > 0 && after % L.ck_every == 0 && slot < L.ck_cap && g.rec_lo == g.rec_hi)
                  ? slot
                  : -1;
  g.dt = dt;
  g.dt_f = (float)dt;
  g.dtrate_f = (float)(dt * pl.rate);
  g.dtg_f = (float)(dt * L.gravity);
  g.dtx_f = (float)(dt / L.dxd);
  return g;
}

// The plan of the step after `pl` (workgroup 0 of the step, every thread):
// clock, output slot and rain rate after the step, as finish_step derives
// them; also the step's entry of the dt log.
__device__ inline void write_next_plan(StepPlan* __restrict__ nx, const LoopConst& L,
                                       const StepPlan& pl, const StepGo& g) {
  const double t1 = pl.t + g.dt;
  const int rj = rain_bracket(L, t1);
  if (threadIdx.x == 0) {
    int oi = pl.oi;
    while (oi < L.n_times && t1 >= L.times[oi] - kOutputSlack) ++oi;
    nx->t = t1;
    nx->rate = interp_at(L.rain_t, L.rain_r, L.n_rain, L.rain_default, t1, rj);
    for (int k = 0; k < 4; ++k) nx->tw[k] = oi + k < L.n_times ? L.times[oi + k] : HUGE_VAL;
    nx->prev_dt = g.dt;
    nx->prev_dtrate = g.dtrate_f;
    nx->steps = pl.steps + 1;
    nx->oi = oi;
    nx->done = (oi >= L.n_times || pl.steps + 1 >= L.max_steps) ? 1 : 0;
    if (L.dt_log && pl.steps < L.dt_log_cap) L.dt_log[pl.steps] = g.dt;
  }
}

// Account the step before `pl` (its partials at [first, first + count)) and
// record the loop's position at the start of `pl` (every thread of one
// workgroup).  m: CFL maximum of the state entering `pl`.
template <int UNROLL>
__device__ inline void account_before(LoopControl* __restrict__ ctl, const LoopConst& L,
                                      const Partials& P, int first, int count,
                                      const StepPlan& pl, float m, bool nonfinite) {
  const Totals tot = reduce_partials<kBlock, UNROLL, false>(P, first, count);
  if (threadIdx.x == 0) {
    account_step(ctl, L, tot, pl.prev_dt, pl.prev_dtrate);
    ctl->nonfinite |= nonfinite ? 1 : 0;
    ctl->max_inv = (double)m;
    ctl->steps = pl.steps;
    ctl->t = pl.t;
    ctl->out_idx = pl.oi;
    ctl->acc_steps = pl.steps;
  }
}

// Series a finish kernel stages in LDS when they are this short or shorter.
constexpr int kLdsSeries = 512;

// Finish of one step by one workgroup of BLOCK threads: reduce the partials
// [first, first + count) of the step kernel that just ended (visible at the
// kernel boundary), close the step in FP64 and plan the next one.  Every
// global load is issued up front, in one round trip: the control block and
// the rain and output-time series go to LDS while the partials are in
// flight, so the single-thread bookkeeping afterwards reads only LDS.
template <int BLOCK, int UNROLL, class Parts>
__device__ void finish_block(LoopControl* __restrict__ ctl, const LoopConst& L, const Parts& P,
                             int first, int count) {
  __shared__ LoopControl s_ctl;
  __shared__ double s_rt[kLdsSeries], s_rr[kLdsSeries], s_tm[kLdsSeries];
  static_assert(sizeof(LoopControl) % sizeof(unsigned) == 0, "LoopControl is copied by words");
  constexpr int kWords = (int)(sizeof(LoopControl) / sizeof(unsigned));
  static_assert(kWords <= BLOCK, "one word per thread");
  const int tid = threadIdx.x;
  if (tid < kWords) reinterpret_cast<unsigned*>(&s_ctl)[tid] = reinterpret_cast<const unsigned*>(ctl)[tid];
  const bool rain_lds = L.n_rain <= kLdsSeries, times_lds = L.n_times <= kLdsSeries;
  if (rain_lds)
    for (int k = tid; k < L.n_rain; k += BLOCK) {
      s_rt[k] = L.rain_t[k];
      s_rr[k] = L.rain_r[k];
    }
  if (times_lds)
    for (int k = tid; k < L.n_times; k += BLOCK) s_tm[k] = L.times[k];
  const Totals tot = reduce_partials<BLOCK, UNROLL, true>(P, first, count);  // syncs the LDS copies
  if (s_ctl.done) return;  // block-uniform
  LoopConst Ls = L;
  if (rain_lds) {
    Ls.rain_t = s_rt;
    Ls.rain_r = s_rr;
  }
  if (times_lds) Ls.times = s_tm;
  const double t_new = s_ctl.t + s_ctl.dt;
  const int rj = rain_bracket(Ls, t_new);
  if (tid == 0) finish_step(&s_ctl, Ls, tot, t_new, rj);
  __syncthreads();
  if (tid < kWords) reinterpret_cast<unsigned*>(ctl)[tid] = reinterpret_cast<const unsigned*>(&s_ctl)[tid];
}

}  // namespace swe
