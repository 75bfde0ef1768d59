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

// Device-resident state of the time loop.  The first fields are what every
// step kernel reads on entry.
struct LoopControl {
  float dt_f, dtrate_f, dtg_f;  // float32(dt), float32(dt*rate), float32(dt*g)
  int done;
  int rec_lo, rec_hi;  // output slots the pending step writes
  int ck_slot;         // checkpoint slot the pending step writes (-1: none)
  int steps, out_idx;
  double t, dt, rate;
  double vol0, vol, rain, bnd, clip;
  double max_inv;
  int nonfinite;
  unsigned counter;  // workgroups finished in the current launch
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
};

struct Partials {
  float* inv;
  double* vol;
  double* clip;
  double* bnd;
  int* nf;
};

// MODE_STEP: caller-given dt, no epilogue.  MODE_LOOP: device loop, the last
// workgroup finishes the step.  MODE_LOOP_SPLIT: device loop, every
// workgroup stores its partials and a one-workgroup kernel finishes the step.
enum { MODE_STEP = 0, MODE_LOOP = 1, MODE_LOOP_SPLIT = 2 };
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
__device__ Totals block_reduce(float inv, double vol, double clip, double bnd, int nf) {
  __shared__ float s_inv[WAVES];
  __shared__ double s_vol[WAVES], s_clip[WAVES], s_bnd[WAVES];
  __shared__ int s_nf[WAVES];
  inv = wave_max(inv);
  vol = wave_sum(vol);
  clip = wave_sum(clip);
  bnd = wave_sum(bnd);
  nf = wave_or(nf);
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

// Close the step just taken (single thread): account its rain, open-boundary
// and clip volumes in FP64, advance the clock and step count, move past the
// output times it reached (numerics.py:398-412), then stop or plan the next
// step.  Resets the workgroup counter for the next launch.
__device__ inline void finish_step(LoopControl* ctl, const LoopConst& L, const Totals& tot,
                                   double t_new, int rain_j) {
  ctl->rain += (double)ctl->dtrate_f * *L.sum_mult * L.cell_area;
  ctl->bnd += ctl->dt * tot.bnd;
  ctl->clip += tot.clip * L.cell_area;
  ctl->vol = tot.vol * L.cell_area;
  ctl->nonfinite |= tot.nf;
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
      pi[u] = in ? P.inv[q] : 0.0f;
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
  const Totals tot = block_reduce<BLOCK / kWave>(inv, vol, clip, bnd, nf);  // syncs the LDS copies
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
