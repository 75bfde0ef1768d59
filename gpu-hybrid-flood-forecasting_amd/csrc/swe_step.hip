// libswe_hip: fused shallow-water step, CFL reduction and device time loop
// for MI355X (gfx950), behind the C ABI declared in include/swe_hip.h.
//
// Kernels
//   k_step<M, MODE>    one fused pass per step: ghost cells, wall mirroring,
//                      hydrostatic reconstruction, HLL on every face once,
//                      divergence, rain, positivity clip, semi-implicit
//                      Manning friction and wet/dry masking
//                      (jax_solver/numerics.py:312-369).  In MODE_LOOP the
//                      epilogue also reduces the CFL maximum of the new state
//                      and the FP64 volume, clip and open-boundary partials;
//                      the last workgroup to finish sets the next dt, time,
//                      output slots and stop flag (numerics.py:398-418).
//   k_reduce<M, MODE>  CFL maximum + volume of a state (numerics.py:297-309);
//                      MODE_INIT also primes the loop control block.
//   k_volume_packed    volume of a packed loop's final state.
//   k_cfl_field<M>     per-cell inverse CFL time.
//   k_sum_mult         FP64 sum of the rain multiplier over active cells.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "swe_device.hpp"
#include "swe_strip.hpp"
#include "swe_fast.hpp"
#include "swe_loop.hpp"
#include "swe_host.hpp"
#include "../../include/swe_hip.h"

namespace swe {

// ---------------------------------------------------------------------------
// The fused step
// ---------------------------------------------------------------------------
// Partials in flight per thread when workgroup 0 of a fused-loop step sums
// the step before's (few: it shares the step kernel's register budget).
constexpr int kFusedUnroll = 2;

template <class M, int MODE, bool PACKED = false>
__global__ __launch_bounds__(kBlock) void k_step(
    Grid G, const float* __restrict__ H, const float* __restrict__ HU,
    const float* __restrict__ HV, float* __restrict__ Ho, float* __restrict__ HUo,
    float* __restrict__ HVo, float dt_arg, float dtx_arg, float dtrate_arg, float dtg_arg, int rows,
    int segs_x, int n_waves, int rev, int par, LoopControl* __restrict__ ctl, LoopConst L,
    Partials P, float* __restrict__ cube_h, float* __restrict__ cube_hu,
    float* __restrict__ cube_hv) {
  float dt = dt_arg, dtx = dtx_arg, dtrate = dtrate_arg, dtg = dtg_arg;
  int rec_lo = 0, rec_hi = 0;
  int steps_before = 0;  // fused loop: index of this step
  int sr_step = -1;      // device loop, rain rounded stochastically: index of this step
  int lay = kStatic12;   // packed loop: static record (L.sum_mult[3], set at the loop start)
  float cls_n2 = 0.0f;   // kStaticCls: n^2 of class threadIdx.x (negative: inactive)
  if constexpr (PACKED) {
    static_assert(kBlock == 256, "one class-table entry per thread");
    lay = (int)L.sum_mult[3];
    if (lay == kStaticCls) cls_n2 = G.cls_tab[threadIdx.x];
  }
  if constexpr (MODE == MODE_LOOP_FUSED) {
    // par: parity of this step.  Every workgroup derives the step from its
    // plan and the CFL maximum that the previous step left in the slots.
    if (ctl->done) return;
    // Wavefront 0 derives the step (float64 arithmetic every lane would
    // repeat) and hands it to the others through LDS.
    __shared__ StepGo s_go;
    __shared__ int s_steps;
    if (wave_in_block() == 0) {
      const StepPlan pl = ctl->plan[par];
      const StepGo g0 = step_go(pl, cfl_bits(P.cfl_slots + kCflSlots * par, pl.steps), L);
      if (threadIdx.x == 0) {
        s_go = g0;
        s_steps = pl.steps;
      }
    }
    __syncthreads();
    StepGo go = s_go;
    // What comes from LDS is a vector value to the compiler.  The step's
    // control values are uniform: read into scalar registers they give scalar
    // branches, not lane masks kept (and spilled) through the whole march.
    go.stop = __builtin_amdgcn_readfirstlane((int)go.stop) != 0;
    go.rec_lo = __builtin_amdgcn_readfirstlane(go.rec_lo);
    go.rec_hi = __builtin_amdgcn_readfirstlane(go.rec_hi);
    go.ck_slot = __builtin_amdgcn_readfirstlane(go.ck_slot);
    steps_before = __builtin_amdgcn_readfirstlane(s_steps);
    if (go.stop) {
      // Every workgroup sees the same plan and slots, so all stop here;
      // workgroup 0 accounts the last step taken and marks the loop done.
      if (blockIdx.x == 0) {
        const StepPlan pl = ctl->plan[par];
        if (pl.steps >= 1 && ctl->acc_steps < pl.steps)
          account_before<kFusedUnroll>(ctl, L, P, (1 - par) * P.stride, (int)gridDim.x, pl,
                                       go.m, go.nonfinite);
        if (threadIdx.x == 0) {
          ctl->nonfinite |= go.nonfinite ? 1 : 0;
          ctl->done = 1;
        }
      }
      return;
    }
    dt = go.dt_f;
    dtx = go.dtx_f;
    dtrate = go.dtrate_f;
    dtg = go.dtg_f;
    rec_lo = go.rec_lo;
    rec_hi = go.rec_hi;
    if (L.rain_sr && dtrate != 0.0f) sr_step = steps_before;
    if (go.ck_slot >= 0) {
      cube_h = L.ck_h;
      cube_hu = L.ck_hu;
      cube_hv = L.ck_hv;
      rec_lo = go.ck_slot;
      rec_hi = go.ck_slot + 1;
    }
  } else if constexpr (MODE != MODE_STEP) {
    if (ctl->done) return;
    dt = ctl->dt_f;
    dtx = ctl->dtx_f;
    dtrate = ctl->dtrate_f;
    dtg = ctl->dtg_f;
    rec_lo = ctl->rec_lo;
    rec_hi = ctl->rec_hi;
    if (L.rain_sr && dtrate != 0.0f) sr_step = ctl->steps;
    // A checkpoint step (never also an output step) records into the
    // checkpoint slots through the same stores.
    const int ck = ctl->ck_slot;
    if (ck >= 0) {
      cube_h = L.ck_h;
      cube_hu = L.ck_hu;
      cube_hv = L.ck_hv;
      rec_lo = ck;
      rec_hi = ck + 1;
    }
  }
  static_assert(!PACKED || (M::kFast && MODE != MODE_STEP), "packed layout: fast device loop only");
  if constexpr (PACKED) {
    // The class table holds dtg n^2, the product the friction factor starts
    // from (rounded as the march would round it), so the march takes it as
    // is.  Negative entries (inactive) keep their sign and value.
    if (lay == kStaticCls)
      class_table()[threadIdx.x] = signbit(cls_n2) ? cls_n2 : dtg * cls_n2;
    __syncthreads();
  }
  int gw = xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + wave_in_block();
  // rev: this step sweeps the strips in the opposite order, so that its
  // first strips are the ones the previous step touched last and still has
  // in the memory-side cache.  Strip order does not change any cell's
  // arithmetic.
  if (rev && gw < n_waves) gw = n_waves - 1 - gw;
  StripAcc acc;
  if (gw < n_waves) {
    if constexpr (M::kFast && PACKED) {
      // The static record chosen at the loop start; the uniform multiplier
      // is L.sum_mult[2].
      const float mc = (float)L.sum_mult[2];
      if (lay == kStaticCls)
        march_strip_fast<true, true, kStaticCls, false>(G, H, HU, HV, Ho, HUo, HVo, gw % segs_x,
                                                 gw / segs_x, rows, dtx, dtrate, dtg, L.dxd,
                                                 L.dyd, rec_lo, rec_hi, cube_h, cube_hu, cube_hv,
                                                 acc, mc, class_table(), sr_step);
      else if (lay == kStatic8)
        march_strip_fast<true, true, kStatic8, false>(G, H, HU, HV, Ho, HUo, HVo, gw % segs_x,
                                               gw / segs_x, rows, dtx, dtrate, dtg, L.dxd, L.dyd,
                                               rec_lo, rec_hi, cube_h, cube_hu, cube_hv, acc, mc,
                                               nullptr, sr_step);
      else
        march_strip_fast<true, true, kStatic12, false>(G, H, HU, HV, Ho, HUo, HVo, gw % segs_x,
                                                gw / segs_x, rows, dtx, dtrate, dtg, L.dxd, L.dyd,
                                                rec_lo, rec_hi, cube_h, cube_hu, cube_hv, acc, 0.0f,
                                                nullptr, sr_step);
    } else if constexpr (M::kFast)
      march_strip_fast<MODE != MODE_STEP, false>(G, H, HU, HV, Ho, HUo, HVo, gw % segs_x,
                                                 gw / segs_x, rows, dtx, dtrate, dtg, L.dxd, L.dyd,
                                                 rec_lo, rec_hi, cube_h, cube_hu, cube_hv, acc,
                                                 0.0f, nullptr, sr_step);
    else
      march_strip<M, MODE != MODE_STEP>(G, H, HU, HV, Ho, HUo, HVo, gw % segs_x, gw / segs_x,
                                        rows, dt, dtrate, dtg, L.dxd, L.dyd, rec_lo, rec_hi,
                                        cube_h, cube_hu, cube_hv, acc, sr_step);
  }

  if constexpr (MODE == MODE_LOOP_SPLIT) {
    const Totals mine = block_reduce(acc.inv, acc.vol, acc.clip, acc.bnd, acc.nf);
    if (threadIdx.x == 0) {
      const int b = blockIdx.x;
      P.inv[b] = mine.inv;
      P.vol[b] = mine.vol;
      P.clip[b] = mine.clip;
      P.bnd[b] = mine.bnd;
      P.nf[b] = mine.nf;
    }
  }
  if constexpr (MODE == MODE_LOOP_FUSED) {
    const Totals mine = block_reduce(acc.inv, acc.vol, acc.clip, acc.bnd, acc.nf);
    if (threadIdx.x == 0) {
      const int q = par * P.stride + (int)blockIdx.x;
      P.vol[q] = mine.vol;
      P.clip[q] = mine.clip;
      P.bnd[q] = mine.bnd;
      P.nf[q] = mine.nf;
      const unsigned bits = mine.nf ? kNonfiniteBits : __float_as_uint(mine.inv);
      const unsigned long long tag = (unsigned long long)(unsigned)(steps_before + 1) << 32;
      __hip_atomic_fetch_max(P.cfl_slots + kCflSlots * (1 - par) + (blockIdx.x & (kCflSlots - 1)),
                             tag | bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (blockIdx.x == 0) {
      // Workgroup 0: the plan of the next step, and the FP64 accounting of
      // the step before this one (its partials are complete: that kernel
      // ended before this one started).  The plan is read and the step
      // derived again rather than kept in registers through the march (the
      // memory clobber forces the reload; nothing this kernel writes changes
      // them).
      asm volatile("" ::: "memory");
      const StepPlan pl = ctl->plan[par];
      const StepGo go = step_go(pl, cfl_bits(P.cfl_slots + kCflSlots * par, pl.steps), L);
      write_next_plan(&ctl->plan[1 - par], L, pl, go);
      if (pl.steps >= 1)
        account_before<kFusedUnroll>(ctl, L, P, (1 - par) * P.stride, (int)gridDim.x, pl, go.m,
                                     false);
    }
  }
  if constexpr (MODE == MODE_LOOP) {
    const Totals mine = block_reduce(acc.inv, acc.vol, acc.clip, acc.bnd, acc.nf);
    if (!publish_partials(P, ctl, mine)) return;
    // Last workgroup: finish the step on the device.
    const Totals tot = gather_partials(P, gridDim.x);
    const double t_new = ctl->t + ctl->dt;
    const int rj = rain_bracket(L, t_new);
    if (threadIdx.x == 0) finish_step(ctl, L, tot, t_new, rj);
  }
}

// ---------------------------------------------------------------------------
// CFL maximum / volume of a state; MODE RED_INIT primes the loop.
// ---------------------------------------------------------------------------
// Finish of MODE_LOOP_SPLIT: one workgroup reduces the step kernel's
// partials and plans the next step.  It is on the critical path of every
// step, so it is wide, loads kFinishUnroll partials per thread per round
// trip, and issues all of its global loads at once (finish_block).
constexpr int kFinishBlock = 256;
constexpr int kFinishUnroll = 16;

__global__ __launch_bounds__(kFinishBlock) void k_finish(LoopControl* __restrict__ ctl,
                                                         LoopConst L, Partials P, int nparts) {
  finish_block<kFinishBlock, kFinishUnroll>(ctl, L, P, 0, nparts);
}

// End of a fused device loop (one workgroup): account the last step taken if
// no later step kernel did (the loop ran out of launches before it stopped
// itself), and leave the clock, step count and output cursor after it.
__global__ __launch_bounds__(kBlock) void k_close(LoopControl* __restrict__ ctl, LoopConst L,
                                                  Partials P, int nparts) {
  const int executed = max(ctl->plan[0].steps, ctl->plan[1].steps);
  if (executed < 1 || ctl->acc_steps >= executed) return;  // block-uniform
  const int p = executed & 1;  // the plan after the last step taken
  const StepPlan pl = ctl->plan[p];
  const unsigned bits = cfl_bits(P.cfl_slots + kCflSlots * p, executed);
  const bool nonfinite = bits == kNonfiniteBits;
  account_before<kFinishUnroll>(ctl, L, P, (1 - p) * P.stride, nparts, pl,
                                __uint_as_float(nonfinite ? 0u : bits), nonfinite);
  if (threadIdx.x == 0 && (pl.done || nonfinite)) ctl->done = 1;
}

// Volume of the final state of a packed loop, whose step kernels do not sum
// it: the active cells' depths in float64, with the cell-to-thread mapping,
// launch shape and partial sums of k_reduce, so that the final volume is
// added up in the same order as the initial one.  The state is work buffer a
// or b by the parity of the loop's step count.  Sets ctl->vol.
__global__ __launch_bounds__(kBlock) void k_volume_packed(Grid G, const float* __restrict__ a,
                                                         const float* __restrict__ b,
                                                         LoopControl* __restrict__ ctl,
                                                         LoopConst L, Partials P) {
  const float* S = (ctl->steps & 1) ? b : a;
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  double vol = 0.0;
  for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < N;
       k += (size_t)gridDim.x * kBlock)
    vol += G.mask[k] != 0 ? (double)S[3 * k] : 0.0;
  const Totals mine = block_reduce(0.0f, vol, 0.0, 0.0, 0);
  if (!publish_partials(P, ctl, mine)) return;
  const Totals tot = gather_partials(P, gridDim.x);
  if (threadIdx.x == 0) {
    ctl->vol = tot.vol * L.cell_area;
    __hip_atomic_store(&ctl->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <class M, int MODE>
__global__ __launch_bounds__(kBlock) void k_reduce(Grid G, const float* __restrict__ H,
                                                  const float* __restrict__ HU,
                                                  const float* __restrict__ HV,
                                                  LoopControl* __restrict__ ctl, LoopConst L,
                                                  Partials P) {
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  float inv = 0.0f;
  double vol = 0.0;
  int nf = 0;
  for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < N;
       k += (size_t)gridDim.x * kBlock) {
    const float h = H[k], hu = HU[k], hv = HV[k];
    const int a = G.mask[k] != 0;
    inv = fmaxf(inv, cfl_inverse<M>(G, h, hu, hv, a));
    vol += a ? (double)h : 0.0;
    nf |= !(isfinite(h) && isfinite(hu) && isfinite(hv));
  }
  const Totals mine = block_reduce(inv, vol, 0.0, 0.0, nf);
  if (!publish_partials(P, ctl, mine)) return;
  const Totals tot = gather_partials(P, gridDim.x);
  double t0 = 0.0;
  if (MODE == RED_INIT) t0 = L.times[0];
  const int rj = rain_bracket(L, t0);
  if (MODE == RED_INIT) {
    // Slots of the fused loop: the CFL maximum of the initial state, tagged
    // step 0, and the other parity's cleared (tags of an earlier loop would
    // outrank this one's).
    __shared__ unsigned s_bits;
    if (threadIdx.x == 0) s_bits = tot.nf ? kNonfiniteBits : __float_as_uint(tot.inv);
    __syncthreads();
    if (threadIdx.x < 2 * kCflSlots)
      __hip_atomic_store(P.cfl_slots + threadIdx.x,
                         threadIdx.x < kCflSlots ? (unsigned long long)s_bits : 0ull,
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (threadIdx.x == 0) {
    ctl->max_inv = (double)tot.inv;
    if (MODE == RED_INIT) {
      ctl->t = t0;
      ctl->dt = 0.0;
      ctl->steps = 0;
      ctl->vol0 = ctl->vol = tot.vol * L.cell_area;
      ctl->rain = ctl->bnd = ctl->clip = 0.0;
      ctl->nonfinite = tot.nf;
      int oi = 1;
      while (oi < L.n_times && t0 >= L.times[oi] - kOutputSlack) ++oi;
      ctl->out_idx = oi;
      if (oi >= L.n_times || L.max_steps <= 0 || tot.nf) {
        ctl->done = 1;
        ctl->rec_lo = ctl->rec_hi = 0;
        ctl->ck_slot = -1;
      } else {
        ctl->done = 0;
        plan_next_step(ctl, L, tot.inv, t0, oi, rj);
      }
      // Plan of step 0 of the fused loop (the same clock, cursor and rate).
      StepPlan& p0 = ctl->plan[0];
      p0.t = t0;
      p0.rate = interp_at(L.rain_t, L.rain_r, L.n_rain, L.rain_default, t0, rj);
      for (int k = 0; k < 4; ++k) p0.tw[k] = oi + k < L.n_times ? L.times[oi + k] : HUGE_VAL;
      p0.prev_dt = 0.0;
      p0.prev_dtrate = 0.0f;
      p0.steps = 0;
      p0.oi = oi;
      p0.done = ctl->done;
      ctl->plan[1].steps = -1;
      ctl->acc_steps = 0;
    }
    __hip_atomic_store(&ctl->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <class M>
__global__ __launch_bounds__(kBlock) void k_cfl_field(Grid G, const float* __restrict__ H,
                                                     const float* __restrict__ HU,
                                                     const float* __restrict__ HV,
                                                     float* __restrict__ out) {
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < N;
       k += (size_t)gridDim.x * kBlock) {
    out[k] = cfl_inverse<M>(G, H[k], HU[k], HV[k], G.mask[k] != 0);
  }
}

// Per workgroup: the FP64 sum of the rain multiplier over active cells and
// the smallest and largest bit pattern it takes there (equal bit patterns on
// every active cell make the multiplier uniform).
__global__ __launch_bounds__(kBlock) void k_sum_mult(Grid G, double* __restrict__ sum_part,
                                                     unsigned* __restrict__ bits_part) {
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  double s = 0.0;
  unsigned lo = 0xFFFFFFFFu, hi = 0u;
  for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < N;
       k += (size_t)gridDim.x * kBlock) {
    if (G.mask[k]) {
      const float m = G.mult[k];
      const unsigned b = __float_as_uint(m);
      s += (double)m;
      lo = min(lo, b);
      hi = max(hi, b);
    }
  }
  __shared__ unsigned s_lo[kBlock], s_hi[kBlock];
  s_lo[threadIdx.x] = lo;
  s_hi[threadIdx.x] = hi;
  const Totals t = block_reduce(0.0f, s, 0.0, 0.0, 0);  // synchronises
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBlock; ++k) {
      lo = min(lo, s_lo[k]);
      hi = max(hi, s_hi[k]);
    }
    sum_part[blockIdx.x] = t.vol;
    bits_part[2 * blockIdx.x] = lo;
    bits_part[2 * blockIdx.x + 1] = hi;
  }
}

// One workgroup: out[0] = the sum of sum_part[0 .. n) in a fixed order;
// out[1] = 1 if the multiplier has one bit pattern on every active cell (or
// there is no active cell), else 0; out[2] = that value (0 without active
// cells); out[3] = the packed loop's static record without classes
// (kStatic8 for a uniform multiplier, else kStatic12).
__global__ __launch_bounds__(kBlock) void k_sum_partials(const double* __restrict__ sum_part,
                                                         const unsigned* __restrict__ bits_part,
                                                         int n, double* __restrict__ out) {
  double s = 0.0;
  unsigned lo = 0xFFFFFFFFu, hi = 0u;
  for (int k = threadIdx.x; k < n; k += kBlock) {
    s += sum_part[k];
    lo = min(lo, bits_part[2 * k]);
    hi = max(hi, bits_part[2 * k + 1]);
  }
  __shared__ unsigned s_lo[kBlock], s_hi[kBlock];
  s_lo[threadIdx.x] = lo;
  s_hi[threadIdx.x] = hi;
  const Totals t = block_reduce(0.0f, s, 0.0, 0.0, 0);  // synchronises
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBlock; ++k) {
      lo = min(lo, s_lo[k]);
      hi = max(hi, s_hi[k]);
    }
    const bool none = lo > hi;  // no active cell
    out[0] = t.vol;
    out[1] = (none || lo == hi) ? 1.0 : 0.0;
    out[2] = none ? 0.0 : (double)__uint_as_float(lo);
    out[3] = (none || lo == hi) ? (double)kStatic8 : (double)kStatic12;
  }
}

// Copy the loop's final state (buffer A or B by the parity of the step count)
// to `dst`, so the host need not know how many steps the device took.
__global__ __launch_bounds__(kBlock) void k_fill3(float* __restrict__ a, float* __restrict__ b,
                                                 float* __restrict__ c, size_t n, float v) {
  for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < n;
       k += (size_t)gridDim.x * kBlock) {
    a[k] = v;
    b[k] = v;
    c[k] = v;
  }
}

// ---- Class-indexed static record (swe_fast.hpp, kStaticCls) ---------------
// At a loop start the Manning values of the active cells are gathered into a
// 512-slot open-addressing table of their bit patterns; with at most 255 of
// them (and a uniform rain multiplier) every active cell is stored as its
// value's class and the step reads n^2 from a 255-entry table.
constexpr int kClsSlots = 512;

struct ClassBuf {
  unsigned long long* keys;  // kClsSlots: 0 = empty, else (1 << 32) | float bits of n
  float* vals;               // 256: n^2 of class c
  unsigned char* slot_cls;   // kClsSlots: class of the value in slot s
  int* overflow;             // the table filled up
};

__device__ __forceinline__ unsigned cls_hash(unsigned long long key) {
  return (unsigned)((key * 0x9E3779B97F4A7C15ull) >> 55);  // 9 bits: kClsSlots
}

__device__ __forceinline__ unsigned long long cls_key(float n) {
  return (1ull << 32) | (unsigned long long)__float_as_uint(n);
}

__global__ __launch_bounds__(kClsSlots) void k_cls_clear(ClassBuf B) {
  B.keys[threadIdx.x] = 0ull;
  if (threadIdx.x == 0) *B.overflow = 0;
}

// Insert the Manning value of every active cell (the first of each run of
// equal values along a wavefront: cell k - 1 is the west lane's cell).
__global__ __launch_bounds__(kBlock) void k_cls_insert(Grid G, ClassBuf B) {
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  const int lane = threadIdx.x & (kWave - 1);
  for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < N;
       k += (size_t)gridDim.x * kBlock) {
    const int a = G.mask[k] != 0;
    const unsigned bits = __float_as_uint(G.n[k]);
    const int wa = from_west(a);
    const unsigned wbits = (unsigned)from_west((int)bits);
    if (!a || (lane > 0 && wa && wbits == bits)) continue;
    const unsigned long long key = (1ull << 32) | bits;
    unsigned h = cls_hash(key);
    bool placed = false;
    for (int p = 0; p < kClsSlots && !placed; ++p) {
      // A plain read first: after its first insert a value is found without
      // an atomic (a few values over millions of cells would otherwise
      // serialise on their slots); a stale empty slot falls through to the
      // compare-and-swap, which sees the current key.
      unsigned long long old = B.keys[h];
      if (old == 0ull) old = atomicCAS(B.keys + h, 0ull, key);
      placed = old == 0ull || old == key;
      h = (h + 1) & (kClsSlots - 1);
    }
    if (!placed) atomicOr(B.overflow, 1);
  }
}

// One workgroup of kClsSlots threads: number the used slots in slot order,
// fill the class table, and record the static layout of this loop in sm[3]
// (kStaticCls with at most 255 classes, a uniform multiplier and `allow`;
// else kStatic8 with a uniform multiplier; else kStatic12).
__global__ __launch_bounds__(kClsSlots) void k_cls_finish(ClassBuf B, double* __restrict__ sm,
                                                          int allow) {
  __shared__ int s_scan[kClsSlots];
  const int t = threadIdx.x;
  const unsigned long long key = B.keys[t];
  const int used = key != 0ull;
  s_scan[t] = used;
  __syncthreads();
  for (int d = 1; d < kClsSlots; d <<= 1) {  // inclusive scan
    const int v = t >= d ? s_scan[t - d] : 0;
    __syncthreads();
    s_scan[t] += v;
    __syncthreads();
  }
  const int count = s_scan[kClsSlots - 1];
  const bool uni = sm[1] != 0.0;
  const bool ok = allow && uni && *B.overflow == 0 && count <= (int)kClsInactive;
  if (t < 256) B.vals[t] = t == (int)kClsInactive ? -1.0f : 0.0f;  // inactive: n' = -1
  __syncthreads();
  if (ok && used) {
    const int c = s_scan[t] - 1;
    const float n = __uint_as_float((unsigned)key);
    B.vals[c] = n * n;  // as the 8-byte record's n'
    B.slot_cls[t] = (unsigned char)c;
  }
  if (t == 0) sm[3] = ok ? (double)kStaticCls : (uni ? (double)kStatic8 : (double)kStatic12);
}

// Packed layout of the fast device loop (see swe_fast.hpp): static fields and
// state in, state out.  The static record is the one k_cls_finish chose in
// sm[3]: the bed and the class (5 bytes, in blocks of 64 columns), (bed, n')
// with a uniform multiplier, else (bed, n', multiplier).
__global__ __launch_bounds__(kBlock) void k_pack_static(Grid G, float* __restrict__ pk, size_t N,
                                                        const double* __restrict__ sm,
                                                        ClassBuf B) {
  const int lay = (int)sm[3];
  const size_t pitch = cls_row_bytes(G.nx);
  for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < N;
       k += (size_t)gridDim.x * kBlock) {
    const float b = G.bed[k], n2 = G.mask[k] ? G.n[k] * G.n[k] : -1.0f;
    if (lay == kStaticCls) {
      unsigned c = kClsInactive;
      if (G.mask[k]) {
        const unsigned long long key = cls_key(G.n[k]);
        unsigned h = cls_hash(key);
        int p = 0;  // the key is in the table (k_cls_insert); the bound only ends the search
        for (; p < kClsSlots && B.keys[h] != key; ++p) h = (h + 1) & (kClsSlots - 1);
        if (p < kClsSlots) c = B.slot_cls[h];
      }
      const size_t j = k / (size_t)G.nx, i = k % (size_t)G.nx;
      unsigned char* blk = reinterpret_cast<unsigned char*>(pk) + j * pitch + i / 64 * kClsBlockBytes;
      reinterpret_cast<float*>(blk)[i % 64] = b;
      blk[256 + i % 64] = (unsigned char)c;
    } else if (lay == kStatic8) {
      pk[2 * k] = b;
      pk[2 * k + 1] = n2;
    } else {
      pk[3 * k] = b;
      pk[3 * k + 1] = n2;
      pk[3 * k + 2] = G.mult[k];
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_pack_state(const float* __restrict__ h,
                                                      const float* __restrict__ hu,
                                                      const float* __restrict__ hv,
                                                      float* __restrict__ S, size_t N) {
  for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < N;
       k += (size_t)gridDim.x * kBlock) {
    S[3 * k] = h[k];
    S[3 * k + 1] = hu[k];
    S[3 * k + 2] = hv[k];
  }
}

// ctl non-null: the source is a or b by the parity of the loop's step count.
__global__ __launch_bounds__(kBlock) void k_unpack_state(const LoopControl* __restrict__ ctl,
                                                        const float* __restrict__ a,
                                                        const float* __restrict__ b,
                                                        float* __restrict__ h,
                                                        float* __restrict__ hu,
                                                        float* __restrict__ hv, size_t N) {
  const float* S = (ctl && (ctl->steps & 1)) ? b : a;
  for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < N;
       k += (size_t)gridDim.x * kBlock) {
    h[k] = S[3 * k];
    hu[k] = S[3 * k + 1];
    hv[k] = S[3 * k + 2];
  }
}

__global__ __launch_bounds__(kBlock) void k_copy_final(const LoopControl* __restrict__ ctl,
                                                      const float* __restrict__ a,
                                                      const float* __restrict__ b,
                                                      float* __restrict__ d0,
                                                      float* __restrict__ d1,
                                                      float* __restrict__ d2, size_t N) {
  const float* src = (ctl->steps & 1) ? b : a;
  for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < N;
       k += (size_t)gridDim.x * kBlock) {
    d0[k] = src[k];
    d1[k] = src[N + k];
    d2[k] = src[2 * N + k];
  }
}

// ---- The static record of a packed loop, for any grid ------------------------
size_t class_buf_bytes() {
  return sizeof(unsigned long long) * kClsSlots + sizeof(float) * 256 + kClsSlots + sizeof(int);
}

const float* class_values(const void* cls_mem) {
  return reinterpret_cast<const float*>(static_cast<const char*>(cls_mem) +
                                        sizeof(unsigned long long) * kClsSlots);
}

ClassBuf class_buf_at(void* mem) {
  char* m = static_cast<char*>(mem);
  constexpr size_t kKeys = sizeof(unsigned long long) * kClsSlots, kVals = sizeof(float) * 256;
  return ClassBuf{reinterpret_cast<unsigned long long*>(m), reinterpret_cast<float*>(m + kKeys),
                  reinterpret_cast<unsigned char*>(m + kKeys + kVals),
                  reinterpret_cast<int*>(m + kKeys + kVals + kClsSlots)};
}

int prepare_static_record(const Grid& G, void* cls_mem, double* sm, int allow_classes,
                          double* sum_part, unsigned* bits_part, int max_blocks, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  const int nb = (int)std::max<size_t>(
      1, std::min<size_t>((N + (size_t)kBlock * 8 - 1) / ((size_t)kBlock * 8), (size_t)max_blocks));
  hipLaunchKernelGGL(k_sum_mult, dim3(nb), dim3(kBlock), 0, s, G, sum_part, bits_part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, s, sum_part, bits_part, nb, sm);
  if (cls_mem) {
    const ClassBuf B = class_buf_at(cls_mem);
    const int np = (int)std::min<size_t>((N + kBlock - 1) / kBlock, 8192);
    hipLaunchKernelGGL(k_cls_clear, dim3(1), dim3(kClsSlots), 0, s, B);
    hipLaunchKernelGGL(k_cls_insert, dim3(np), dim3(kBlock), 0, s, G, B);
    hipLaunchKernelGGL(k_cls_finish, dim3(1), dim3(kClsSlots), 0, s, B, sm, allow_classes);
    hipLaunchKernelGGL(k_pack_static, dim3(np), dim3(kBlock), 0, s, G, const_cast<float*>(G.pk), N,
                       sm, B);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(SWE_EHIP, std::string("prepare_static_record: ") + hipGetErrorString(e));
  return SWE_OK;
}

}  // namespace swe

// ===========================================================================
// Host side
// ===========================================================================
using namespace swe;

namespace {
thread_local std::string g_last_error;

}  // namespace

int swe::set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

namespace {
int fail(int code, const std::string& msg) { return swe::set_error(code, msg); }

#define SWE_HIP_TRY(expr)                                                              \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(SWE_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
  } while (0)

}  // namespace

struct SweContext {
  int device = 0;
  Grid G{};
  int variant = SWE_VARIANT_FAST;
  int rows = kDefaultRows;
  int graph_chunk = 64;
  double dx = 1, dy = 1, gravity = 9.81, cfl = 0.45, dry = 1e-4, max_dt = 10, rain_default = 0;
  std::vector<double> rain_t, rain_r;
  // [FP64 rain-multiplier sum over active cells, 1 if the multiplier is
  // uniform there, its value, static record of the packed loop]
  double* d_sum_mult = nullptr;
  double* d_rain = nullptr;  // [rain_t | rain_r]
  double* d_times = nullptr;
  int times_cap = 0;
  LoopControl* d_ctl = nullptr;
  LoopControl* h_ctl = nullptr;  // pinned mirror
  Partials P{};
  int parts_cap = 0;
  float* work = nullptr;  // 2 x 3 x N ping-pong state (SoA, or packed: see ws())
  float* pk = nullptr;    // packed static fields of the fast device loop (largest record)
  void* cls_mem = nullptr;   // the Manning class table of the packed loop (ClassBuf)
  ClassBuf cls{};
  int static_classes = 1;    // the packed loop may store the Manning value as a class
  bool loop_packed = false;  // layout of the work state of the current device loop
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t graph = nullptr;
  std::vector<const void*> graph_key;
  // How a device-loop step closes: 1 = fused (no finish kernel), 2 = a
  // one-workgroup finish kernel after each step, 0 = the step's last workgroup.
  int split_finish = 1;
  int alt_sweep = 1;         // odd device-loop steps sweep the strips in reverse order
  int rain_sr = 0;           // device loops add the rain with stochastic rounding
  int sweep_rev = 0;         // reverse flag of the step launch being issued
  int step_par = 0;          // parity of the step launch being issued (fused loop)
  float step_dtx = 0.0f;     // float32(dt/dx) of a caller-given step (swe_step, replays)
  double* dt_log = nullptr;  // caller's device buffer for the dt schedule
  int dt_log_cap = 0;
  float *ck_h = nullptr, *ck_hu = nullptr, *ck_hv = nullptr;  // caller's checkpoint cube
  int ck_every = 0, ck_cap = 0;
  // swe_loop_begin/advance/end state
  double loop_times[2] = {0.0, 0.0};
  LoopConst loop_L{};
  bool loop_active = false;
  int loop_launch_par = 0;  // parity of the next step launch of swe_loop_advance
  double* aux = nullptr;  // scratch of the coupled-forecast kernels
  size_t aux_cap = 0;
  int lds_per_cu = 0;        // bytes of LDS per compute unit
  int adjoint_rows = 0;      // strip height of the adjoint kernel (0 = default_rows)
  unsigned step_lds = 0;     // dynamic LDS reserved per step workgroup (occupancy cap)

  size_t cells() const { return (size_t)G.nx * (size_t)G.ny; }
  float* buf(int which, int field) { return work + ((size_t)which * 3 + field) * cells(); }
  // Packed work state `which`: (h, hu, hv) interleaved, in the same memory.
  float* ws(int which) { return work + (size_t)which * 3 * cells(); }
  // The fast variant's device loop runs on the packed layout.
  bool packed_variant() const { return variant == SWE_VARIANT_FAST; }
  int segs_x() const { return (G.nx + kStripCols - 1) / kStripCols; }
  int n_waves() const { return segs_x() * ((G.ny + rows - 1) / rows); }
  int step_blocks() const { return (n_waves() + kWavesPerBlock - 1) / kWavesPerBlock; }
  int reduce_blocks() const {
    const size_t b = (cells() + (size_t)kBlock * 8 - 1) / ((size_t)kBlock * 8);
    return (int)std::max<size_t>(1, std::min<size_t>(b, 2048));
  }
  LoopConst loop_const(int n_times, int max_steps) const {
    LoopConst L{};
    L.cfl = cfl;
    L.max_dt = max_dt;
    L.gravity = gravity;
    L.cell_area = dx * dy;
    L.sum_mult = d_sum_mult;
    L.rain_default = rain_default;
    L.dxd = dx;
    L.dyd = dy;
    L.rain_t = d_rain;
    L.rain_r = d_rain + rain_t.size();
    L.times = d_times;
    L.dt_log = dt_log;
    L.dt_log_cap = dt_log_cap;
    L.ck_h = ck_h;
    L.ck_hu = ck_hu;
    L.ck_hv = ck_hv;
    L.ck_every = ck_every;
    L.ck_cap = ck_cap;
    L.n_rain = (int)rain_t.size();
    L.n_times = n_times;
    L.max_steps = max_steps;
    L.record = 1;
    L.rain_sr = rain_sr;
    return L;
  }
};

namespace {

int ensure_partials(SweContext* c) {
  const int need = std::max(c->step_blocks(), c->reduce_blocks());
  if (need <= c->parts_cap) return SWE_OK;
  if (c->P.cfl_slots) {
    SWE_HIP_TRY(hipFree(c->P.cfl_slots));
    c->P = Partials{};
    c->parts_cap = 0;
  }
  // One allocation: the fused loop's 2 x 64 CFL slots, then two parities of
  // need partials each: vol | clip | bnd (double) | inv (float) | nf (int).
  const size_t n2 = 2 * (size_t)need;
  const size_t bytes = 2 * kCflSlots * sizeof(unsigned long long) +
                       n2 * (3 * sizeof(double) + sizeof(float) + sizeof(int));
  char* base = nullptr;
  SWE_HIP_TRY(hipMalloc(&base, bytes));
  SWE_HIP_TRY(hipMemset(base, 0, bytes));
  c->P.cfl_slots = reinterpret_cast<unsigned long long*>(base);
  c->P.vol = reinterpret_cast<double*>(c->P.cfl_slots + 2 * kCflSlots);
  c->P.clip = c->P.vol + n2;
  c->P.bnd = c->P.clip + n2;
  c->P.inv = reinterpret_cast<float*>(c->P.bnd + n2);
  c->P.nf = reinterpret_cast<int*>(c->P.inv + n2);
  c->P.stride = need;
  c->parts_cap = need;
  return SWE_OK;
}

int ensure_times(SweContext* c, int n) {
  if (n <= c->times_cap) return SWE_OK;
  if (c->d_times) SWE_HIP_TRY(hipFree(c->d_times));
  c->d_times = nullptr;
  SWE_HIP_TRY(hipMalloc(&c->d_times, sizeof(double) * (size_t)n));
  c->times_cap = n;
  return SWE_OK;
}

int copy_state(SweContext* c, float* dh, float* dhu, float* dhv, const float* sh,
               const float* shu, const float* shv, hipStream_t s);

int ensure_work(SweContext* c) {
  if (!c->work) SWE_HIP_TRY(hipMalloc(&c->work, sizeof(float) * 6 * c->cells()));
  if (!c->pk) {
    // the largest static record: 3 floats per cell, or the class blocks
    const size_t bytes = std::max(sizeof(float) * 3 * c->cells(), cls_row_bytes(c->G.nx) * (size_t)c->G.ny);
    SWE_HIP_TRY(hipMalloc(&c->pk, bytes));
    c->G.pk = c->pk;
  }
  if (!c->cls_mem) {
    SWE_HIP_TRY(hipMalloc(&c->cls_mem, class_buf_bytes()));
    c->cls = class_buf_at(c->cls_mem);
    c->G.cls_tab = c->cls.vals;
  }
  return SWE_OK;
}

// Start a device loop on `src`: pick the work layout for the variant and fill
// work state 0 (for the packed layout also the packed static fields, from the
// caller's current arrays).
// The FP64 sum of the rain multiplier over active cells into d_sum_mult (the
// fields may have been changed in place since the last run).  Uses the
// partials as scratch.
// The partials' vol and clip arrays (reduce_blocks() doubles each) are the
// scratch of the multiplier sum.
int refresh_sum_mult(SweContext* c, hipStream_t s) {
  return prepare_static_record(c->G, nullptr, c->d_sum_mult, 0, c->P.vol,
                               reinterpret_cast<unsigned*>(c->P.clip), c->reduce_blocks(), s);
}

int stage_in(SweContext* c, const SweState& src, hipStream_t s) {
  c->loop_packed = c->packed_variant();
  if (!c->loop_packed) {
    if (int rc = refresh_sum_mult(c, s)) return rc;
    return copy_state(c, c->buf(0, 0), c->buf(0, 1), c->buf(0, 2), src.h, src.hu, src.hv, s);
  }
  if (int rc = prepare_static_record(c->G, c->cls_mem, c->d_sum_mult, c->static_classes, c->P.vol,
                                     reinterpret_cast<unsigned*>(c->P.clip), c->reduce_blocks(), s))
    return rc;
  const size_t N = c->cells();
  const int nb = (int)std::min<size_t>((N + kBlock - 1) / kBlock, 8192);
  hipLaunchKernelGGL(k_pack_state, dim3(nb), dim3(kBlock), 0, s, src.h, src.hu, src.hv, c->ws(0), N);
  SWE_HIP_TRY(hipGetLastError());
  return SWE_OK;
}

// Copy work state `which` of the current loop to the caller's arrays.
int stage_out(SweContext* c, int which, float* dh, float* dhu, float* dhv, hipStream_t s) {
  if (!c->loop_packed) return copy_state(c, dh, dhu, dhv, c->buf(which, 0), c->buf(which, 1), c->buf(which, 2), s);
  const size_t N = c->cells();
  const int nb = (int)std::min<size_t>((N + kBlock - 1) / kBlock, 8192);
  hipLaunchKernelGGL(k_unpack_state, dim3(nb), dim3(kBlock), 0, s, nullptr, c->ws(which), c->ws(which), dh, dhu, dhv, N);
  SWE_HIP_TRY(hipGetLastError());
  return SWE_OK;
}

bool valid_state(const SweState& s) { return s.h && s.hu && s.hv; }

template <class M, int MODE, bool PACKED = false>
void launch_step(SweContext* c, const float* h, const float* hu, const float* hv, float* ho,
                 float* huo, float* hvo, float dt, float dtrate, float dtg, const LoopConst& L,
                 float* ch, float* chu, float* chv, hipStream_t s) {
  hipLaunchKernelGGL((k_step<M, MODE, PACKED>), dim3(c->step_blocks()), dim3(kBlock), c->step_lds, s, c->G, h, hu,
                     hv, ho, huo, hvo, dt, c->step_dtx, dtrate, dtg, c->rows, c->segs_x(), c->n_waves(),
                     c->sweep_rev, c->step_par, c->d_ctl, L, c->P, ch, chu, chv);
}

void launch_step_any(SweContext* c, int mode, const float* h, const float* hu, const float* hv,
                     float* ho, float* huo, float* hvo, float dt, float dtrate, float dtg,
                     const LoopConst& L, float* ch, float* chu, float* chv, hipStream_t s) {
  if (mode == MODE_LOOP && c->split_finish == 1) {
    if (c->loop_packed)
      launch_step<FastMath, MODE_LOOP_FUSED, true>(c, h, hu, hv, ho, huo, hvo, dt, dtrate, dtg, L,
                                                   ch, chu, chv, s);
    else if (c->variant == SWE_VARIANT_REFERENCE)
      launch_step<RefMath, MODE_LOOP_FUSED>(c, h, hu, hv, ho, huo, hvo, dt, dtrate, dtg, L, ch,
                                            chu, chv, s);
    else
      launch_step<FastMath, MODE_LOOP_FUSED>(c, h, hu, hv, ho, huo, hvo, dt, dtrate, dtg, L, ch,
                                             chu, chv, s);
    return;
  }
  if (mode == MODE_LOOP && c->split_finish == 2) {
    if (c->loop_packed)
      launch_step<FastMath, MODE_LOOP_SPLIT, true>(c, h, hu, hv, ho, huo, hvo, dt, dtrate, dtg, L,
                                                   ch, chu, chv, s);
    else if (c->variant == SWE_VARIANT_REFERENCE)
      launch_step<RefMath, MODE_LOOP_SPLIT>(c, h, hu, hv, ho, huo, hvo, dt, dtrate, dtg, L, ch,
                                            chu, chv, s);
    else
      launch_step<FastMath, MODE_LOOP_SPLIT>(c, h, hu, hv, ho, huo, hvo, dt, dtrate, dtg, L, ch,
                                             chu, chv, s);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kFinishBlock), 0, s, c->d_ctl, L, c->P,
                       c->step_blocks());
    return;
  }
  if (c->variant == SWE_VARIANT_REFERENCE) {
    if (mode == MODE_LOOP)
      launch_step<RefMath, MODE_LOOP>(c, h, hu, hv, ho, huo, hvo, dt, dtrate, dtg, L, ch, chu,
                                      chv, s);
    else
      launch_step<RefMath, MODE_STEP>(c, h, hu, hv, ho, huo, hvo, dt, dtrate, dtg, L, ch, chu,
                                      chv, s);
  } else {
    if (mode == MODE_LOOP && c->loop_packed)
      launch_step<FastMath, MODE_LOOP, true>(c, h, hu, hv, ho, huo, hvo, dt, dtrate, dtg, L, ch,
                                             chu, chv, s);
    else if (mode == MODE_LOOP)
      launch_step<FastMath, MODE_LOOP>(c, h, hu, hv, ho, huo, hvo, dt, dtrate, dtg, L, ch, chu,
                                       chv, s);
    else
      launch_step<FastMath, MODE_STEP>(c, h, hu, hv, ho, huo, hvo, dt, dtrate, dtg, L, ch, chu,
                                       chv, s);
  }
}

void launch_reduce(SweContext* c, int mode, const float* h, const float* hu, const float* hv,
                   const LoopConst& L, hipStream_t s) {
  const dim3 grid(c->reduce_blocks()), block(kBlock);
  if (c->variant == SWE_VARIANT_REFERENCE) {
    if (mode == RED_INIT)
      hipLaunchKernelGGL((k_reduce<RefMath, RED_INIT>), grid, block, 0, s, c->G, h, hu, hv,
                         c->d_ctl, L, c->P);
    else
      hipLaunchKernelGGL((k_reduce<RefMath, RED_CFL>), grid, block, 0, s, c->G, h, hu, hv,
                         c->d_ctl, L, c->P);
  } else {
    if (mode == RED_INIT)
      hipLaunchKernelGGL((k_reduce<FastMath, RED_INIT>), grid, block, 0, s, c->G, h, hu, hv,
                         c->d_ctl, L, c->P);
    else
      hipLaunchKernelGGL((k_reduce<FastMath, RED_CFL>), grid, block, 0, s, c->G, h, hu, hv,
                         c->d_ctl, L, c->P);
  }
}

double host_interp(const SweContext* c, double t) {
  const int n = (int)c->rain_t.size();
  int j = 0;
  if (n > 1) {
    // largest j with x[j] <= t (only meaningful inside [x0, x_{n-1}))
    const auto it = std::upper_bound(c->rain_t.begin(), c->rain_t.end(), t);
    j = (int)(it - c->rain_t.begin()) - 1;
    j = std::min(std::max(j, 0), n - 2);
  }
  return interp_at(c->rain_t.data(), c->rain_r.data(), n, c->rain_default, t, j);
}

// One device-loop step reading work state a and writing 1 - a, in the layout
// of the current loop.
void launch_loop_step(SweContext* c, int a, const LoopConst& L, float* ch, float* chu, float* chv,
                      hipStream_t s) {
  c->sweep_rev = c->alt_sweep ? a : 0;
  c->step_par = a;
  struct Reset {
    SweContext* c;
    ~Reset() {
      c->sweep_rev = 0;
      c->step_par = 0;
    }
  } reset{c};
  if (c->loop_packed)
    launch_step_any(c, MODE_LOOP, c->ws(a), nullptr, nullptr, c->ws(1 - a), nullptr, nullptr, 0.f,
                    0.f, 0.f, L, ch, chu, chv, s);
  else
    launch_step_any(c, MODE_LOOP, c->buf(a, 0), c->buf(a, 1), c->buf(a, 2), c->buf(1 - a, 0),
                    c->buf(1 - a, 1), c->buf(1 - a, 2), 0.f, 0.f, 0.f, L, ch, chu, chv, s);
}

// Launch `count` loop steps starting at launch parity 0 (buffer A -> B first).
int launch_loop_chunk(SweContext* c, int count, const LoopConst& L, float* ch, float* chu,
                      float* chv, hipStream_t s) {
  if (c->graph_chunk > 0) {
    std::vector<const void*> key = {c->work, ch, chu, chv, (const void*)(intptr_t)c->variant,
                                    (const void*)(intptr_t)c->rows,
                                    (const void*)(intptr_t)c->graph_chunk, c->d_times,
                                    (const void*)(intptr_t)L.n_times,
                                    (const void*)(intptr_t)L.max_steps,
                                    (const void*)(intptr_t)c->split_finish,
                                    (const void*)(intptr_t)c->alt_sweep,
                                    (const void*)(intptr_t)c->step_lds,
                                    (const void*)(intptr_t)c->loop_packed};
    // The loop constants are baked into the captured kernel arguments.
    uintptr_t words[(sizeof(LoopConst) + sizeof(uintptr_t) - 1) / sizeof(uintptr_t)] = {};
    std::memcpy(words, &L, sizeof(LoopConst));
    for (uintptr_t w : words) key.push_back((const void*)w);
    if (!c->graph || key != c->graph_key) {
      if (c->graph) {
        SWE_HIP_TRY(hipGraphExecDestroy(c->graph));
        c->graph = nullptr;
      }
      if (!c->cap_stream) SWE_HIP_TRY(hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
      SWE_HIP_TRY(hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < c->graph_chunk; ++k) launch_loop_step(c, k & 1, L, ch, chu, chv, c->cap_stream);
      hipGraph_t g = nullptr;
      SWE_HIP_TRY(hipStreamEndCapture(c->cap_stream, &g));
      hipError_t e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return fail(SWE_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      c->graph_key = key;
    }
    for (int done = 0; done < count; done += c->graph_chunk) SWE_HIP_TRY(hipGraphLaunch(c->graph, s));
    return SWE_OK;
  }
  for (int k = 0; k < count; ++k) launch_loop_step(c, k & 1, L, ch, chu, chv, s);
  SWE_HIP_TRY(hipGetLastError());
  return SWE_OK;
}

// After the last step kernel of a loop.  Fused loop: account the last step
// taken (k_close); the control block then holds the loop's final clock, step
// count, output cursor and FP64 balance, as after a finish kernel.  Packed
// loop: sum the final state's volume (k_volume_packed).
int close_loop(SweContext* c, const LoopConst& L, hipStream_t s) {
  if (c->split_finish == 1)
    hipLaunchKernelGGL(k_close, dim3(1), dim3(kBlock), 0, s, c->d_ctl, L, c->P, c->step_blocks());
  if (c->loop_packed)
    hipLaunchKernelGGL(k_volume_packed, dim3(c->reduce_blocks()), dim3(kBlock), 0, s, c->G,
                       c->ws(0), c->ws(1), c->d_ctl, L, c->P);
  SWE_HIP_TRY(hipGetLastError());
  return SWE_OK;
}

// Run the device loop until the control block says done.  Chunks of steps
// are queued two deep; after each chunk the control block is copied to a
// pinned slot and the older chunk's copy decides whether to continue.  On
// return c->h_ctl[0] holds the final control block.
int drive_loop(SweContext* c, const LoopConst& L, float* ch, float* chu, float* chv,
               hipStream_t s, int max_steps) {
  const int chunk = c->graph_chunk > 0 ? c->graph_chunk : 64;
  hipEvent_t ev[2] = {nullptr, nullptr};
  int rc = SWE_OK;
  LoopControl* hc = c->h_ctl;
  long launched = 0;
  bool have_prev = false;
  int slot = 0;
  hipError_t e = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
  if (e != hipSuccess) rc = fail(SWE_EHIP, std::string("loop events: ") + hipGetErrorString(e));
  while (rc == SWE_OK) {
    rc = launch_loop_chunk(c, chunk, L, ch, chu, chv, s);
    if (rc) break;
    launched += chunk;
    e = hipMemcpyAsync(hc + slot, c->d_ctl, sizeof(LoopControl), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipEventRecord(ev[slot], s);
    if (e != hipSuccess) {
      rc = fail(SWE_EHIP, std::string("loop poll: ") + hipGetErrorString(e));
      break;
    }
    if (have_prev) {
      const int prev = 1 - slot;
      e = hipEventSynchronize(ev[prev]);
      if (e != hipSuccess) {
        rc = fail(SWE_EHIP, std::string("loop sync: ") + hipGetErrorString(e));
        break;
      }
      if (hc[prev].done) break;
    }
    have_prev = true;
    slot = 1 - slot;
    // The device stops itself at max_steps; this bound only guards the host.
    if (launched > (long)max_steps + 4L * chunk) {
      rc = fail(SWE_EHIP, "device loop did not stop at max_steps");
      break;
    }
  }
  if (rc == SWE_OK) rc = close_loop(c, L, s);
  e = hipStreamSynchronize(s);
  if (e != hipSuccess && rc == SWE_OK) rc = fail(SWE_EHIP, std::string("loop end: ") + hipGetErrorString(e));
  if (rc == SWE_OK) {
    e = hipMemcpy(hc, c->d_ctl, sizeof(LoopControl), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = fail(SWE_EHIP, std::string("loop read: ") + hipGetErrorString(e));
  }
  if (ev[0]) (void)hipEventDestroy(ev[0]);
  if (ev[1]) (void)hipEventDestroy(ev[1]);
  return rc;
}

int copy_final(SweContext* c, SweState dst, hipStream_t s) {
  const size_t N = c->cells();
  const int nb = (int)std::min<size_t>((N + kBlock - 1) / kBlock, 4096);
  if (c->loop_packed)
    hipLaunchKernelGGL(k_unpack_state, dim3(nb), dim3(kBlock), 0, s, c->d_ctl, c->ws(0), c->ws(1),
                       dst.h, dst.hu, dst.hv, N);
  else
    hipLaunchKernelGGL(k_copy_final, dim3(nb), dim3(kBlock), 0, s, c->d_ctl, c->buf(0, 0),
                       c->buf(1, 0), dst.h, dst.hu, dst.hv, N);
  SWE_HIP_TRY(hipGetLastError());
  return SWE_OK;
}

void fill_info(const SweContext* c, const LoopControl& k, SweRunInfo* info) {
  (void)c;
  if (!info) return;
  info->outputs_written = k.out_idx;
  info->steps_used = k.steps;
  info->final_time_s = k.t;
  info->volume_initial = k.vol0;
  info->volume_final = k.vol;
  info->rain_volume = k.rain;
  info->boundary_volume = k.bnd;
  info->clip_volume = k.clip;
  info->mass_balance_error = k.vol - k.vol0 - k.rain - k.bnd - k.clip;
  const double scale = std::max(std::fabs(k.vol0) + std::fabs(k.rain) + std::fabs(k.bnd) +
                                    std::fabs(k.clip),
                                1.0);
  info->mass_balance_rel = info->mass_balance_error / scale;
  info->nonfinite = k.nonfinite;
  info->_pad = 0;
}

int copy_state(SweContext* c, float* dh, float* dhu, float* dhv, const float* sh,
               const float* shu, const float* shv, hipStream_t s) {
  const size_t bytes = sizeof(float) * c->cells();
  SWE_HIP_TRY(hipMemcpyAsync(dh, sh, bytes, hipMemcpyDeviceToDevice, s));
  SWE_HIP_TRY(hipMemcpyAsync(dhu, shu, bytes, hipMemcpyDeviceToDevice, s));
  SWE_HIP_TRY(hipMemcpyAsync(dhv, shv, bytes, hipMemcpyDeviceToDevice, s));
  return SWE_OK;
}

}  // namespace

// Accessors for the other translation units (the adjoint).
const Grid* swe::context_grid(const SweContext* c) { return &c->G; }
int swe::context_variant(const SweContext* c) { return c->variant; }
double swe::context_gravity(const SweContext* c) { return c->gravity; }
double swe::context_rate(const SweContext* c, double t) { return host_interp(c, t); }
int swe::context_reduce_blocks(const SweContext* c) { return c->reduce_blocks(); }
int swe::context_device(const SweContext* c) { return c->device; }
int swe::context_adjoint_rows(const SweContext* c) {
  return c->adjoint_rows > 0 ? c->adjoint_rows : default_rows(c->G.nx, c->G.ny);
}
double* swe::context_scratch(SweContext* c, size_t n_doubles) {
  if (n_doubles > c->aux_cap) {
    if (c->aux && hipFree(c->aux) != hipSuccess) return nullptr;
    c->aux = nullptr;
    c->aux_cap = 0;
    if (hipMalloc(&c->aux, sizeof(double) * n_doubles) != hipSuccess) return nullptr;
    c->aux_cap = n_doubles;
  }
  return c->aux;
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* swe_version(void) { return "swe_hip 0.1 (gfx950)"; }
const char* swe_last_error(void) { return g_last_error.c_str(); }

int swe_interp_rate(const float* times_s, const float* rates_m_s, int32_t n, double dflt,
                    double time_s, double* rate_out) {
  if (!times_s || !rates_m_s || n < 1 || !rate_out) return fail(SWE_EINVAL, "swe_interp_rate: bad arguments");
  std::vector<double> x(times_s, times_s + n), y(rates_m_s, rates_m_s + n);
  int j = 0;
  if (n > 1) {
    const auto it = std::upper_bound(x.begin(), x.end(), time_s);
    j = std::min(std::max((int)(it - x.begin()) - 1, 0), n - 2);
  }
  *rate_out = interp_at(x.data(), y.data(), n, dflt, time_s, j);
  return SWE_OK;
}

int swe_create(const SweParamsDesc* p, int32_t variant, SweContext** out) {
  if (!p || !out) return fail(SWE_EINVAL, "swe_create: null argument");
  *out = nullptr;
  if (p->nx < 1 || p->ny < 1) return fail(SWE_EINVAL, "swe_create: grid must be at least 1x1");
  if ((int64_t)p->nx * (int64_t)p->ny > (int64_t)1 << 31)
    return fail(SWE_EINVAL, "swe_create: grid larger than 2^31 cells");
  if (!p->bed || !p->manning_n || !p->domain_mask || !p->rainfall_multiplier)
    return fail(SWE_EINVAL, "swe_create: a field pointer is null");
  if (p->n_rainfall < 1 || !p->rainfall_times_s || !p->rainfall_rates_m_s)
    return fail(SWE_EINVAL, "swe_create: at least one rainfall sample is required");
  for (int k = 1; k < p->n_rainfall; ++k)
    if (!(p->rainfall_times_s[k] > p->rainfall_times_s[k - 1]))
      return fail(SWE_EINVAL, "swe_create: rainfall times must increase strictly");
  for (int k = 0; k < 4; ++k)
    if (p->boundary_types[k] != SWE_REFLECTIVE && p->boundary_types[k] != SWE_OPEN)
      return fail(SWE_EINVAL, "swe_create: boundary types must be 0 (reflective) or 1 (open)");
  if (!(p->dx > 0) || !(p->dy > 0) || !(p->gravity > 0) || !(p->cfl > 0) ||
      !(p->dry_tolerance > 0) || !(p->max_dt > 0))
    return fail(SWE_EINVAL, "swe_create: dx, dy, gravity, cfl, dry_tolerance, max_dt must be > 0");
  if (variant != SWE_VARIANT_REFERENCE && variant != SWE_VARIANT_FAST)
    return fail(SWE_EINVAL, "swe_create: unknown kernel variant");

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(SWE_ENODEV, "swe_create: no HIP device is visible");
  int dev = 0;
  SWE_HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  SWE_HIP_TRY(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(SWE_ENODEV, std::string("swe_create: device is ") + prop.gcnArchName +
                                ", this library is built for gfx950");

  SweContext* c = new SweContext();
  c->device = dev;
  c->lds_per_cu = (int)prop.maxSharedMemoryPerMultiProcessor;
  c->variant = variant;
  c->dx = p->dx;
  c->dy = p->dy;
  c->gravity = p->gravity;
  c->cfl = p->cfl;
  c->dry = p->dry_tolerance;
  c->max_dt = p->max_dt;
  c->rain_default = p->default_rainfall_m_s;
  c->rain_t.assign(p->rainfall_times_s, p->rainfall_times_s + p->n_rainfall);
  c->rain_r.assign(p->rainfall_rates_m_s, p->rainfall_rates_m_s + p->n_rainfall);
  Grid& G = c->G;
  G.nx = p->nx;
  G.ny = p->ny;
  for (int k = 0; k < 4; ++k) G.bt[k] = p->boundary_types[k];
  // Array-side scalars are float32 (jax_solver/shallow_water_2d.py:246-251);
  // 0.5*g is the float64 product rounded once (numpy_solver/shallow_water_2d.py:84).
  G.g = (float)p->gravity;
  G.half_g = (float)(0.5 * p->gravity);
  G.dry = (float)p->dry_tolerance;
  G.dx = (float)p->dx;
  G.dy = (float)p->dy;
  G.rdx = 1.0f / G.dx;
  G.rdy = 1.0f / G.dy;
  G.rho = (float)(p->dy / p->dx);
  G.rhox = (float)(p->dx / p->dy);
  G.bed = p->bed;
  G.n = p->manning_n;
  G.mask = p->domain_mask;
  G.mult = p->rainfall_multiplier;

  auto cleanup = [&](int rc) {
    swe_destroy(c);
    return rc;
  };
  if (hipMalloc(&c->d_rain, sizeof(double) * 2 * c->rain_t.size()) != hipSuccess)
    return cleanup(fail(SWE_EHIP, "swe_create: hipMalloc rain"));
  std::vector<double> rr(c->rain_t);
  rr.insert(rr.end(), c->rain_r.begin(), c->rain_r.end());
  if (hipMemcpy(c->d_rain, rr.data(), sizeof(double) * rr.size(), hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(SWE_EHIP, "swe_create: upload rain"));
  if (hipMalloc(&c->d_ctl, sizeof(LoopControl)) != hipSuccess)
    return cleanup(fail(SWE_EHIP, "swe_create: hipMalloc control"));
  if (hipMemset(c->d_ctl, 0, sizeof(LoopControl)) != hipSuccess)
    return cleanup(fail(SWE_EHIP, "swe_create: clear control"));
  if (hipHostMalloc(&c->h_ctl, 2 * sizeof(LoopControl), hipHostMallocDefault) != hipSuccess)
    return cleanup(fail(SWE_EHIP, "swe_create: hipHostMalloc control"));
  c->rows = default_rows(G.nx, G.ny);
  int rc = ensure_partials(c);
  if (rc) return cleanup(rc);
  // Static scratch of the allocation-free residual correction, resolved once
  // per device while this device is current.
  if ((rc = noalloc_scratch_init(dev))) return cleanup(rc);
  // Sum of the rain multiplier over active cells, for the FP64 rain volume
  // (refreshed at every loop start).
  if (hipMalloc(&c->d_sum_mult, 4 * sizeof(double)) != hipSuccess)
    return cleanup(fail(SWE_EHIP, "swe_create: hipMalloc rain multiplier sum"));
  if ((rc = refresh_sum_mult(c, nullptr))) return cleanup(rc);
  if (hipDeviceSynchronize() != hipSuccess)
    return cleanup(fail(SWE_EHIP, "swe_create: rain multiplier sum"));
  *out = c;
  return SWE_OK;
}

int swe_destroy(SweContext* c) {
  if (!c) return SWE_OK;
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  if (c->cap_stream) (void)hipStreamDestroy(c->cap_stream);
  if (c->work) (void)hipFree(c->work);
  if (c->pk) (void)hipFree(c->pk);
  if (c->cls_mem) (void)hipFree(c->cls_mem);
  if (c->P.cfl_slots) (void)hipFree(c->P.cfl_slots);
  if (c->d_times) (void)hipFree(c->d_times);
  if (c->d_ctl) (void)hipFree(c->d_ctl);
  if (c->d_sum_mult) (void)hipFree(c->d_sum_mult);
  if (c->h_ctl) (void)hipHostFree(c->h_ctl);
  if (c->d_rain) (void)hipFree(c->d_rain);
  if (c->aux) (void)hipFree(c->aux);
  delete c;
  return SWE_OK;
}

int swe_set_variant(SweContext* c, int32_t variant) {
  if (!c) return fail(SWE_EINVAL, "swe_set_variant: null context");
  if (variant != SWE_VARIANT_REFERENCE && variant != SWE_VARIANT_FAST)
    return fail(SWE_EINVAL, "swe_set_variant: unknown kernel variant");
  if (c->loop_active) return fail(SWE_EINVAL, "swe_set_variant: a device loop is active");
  c->variant = variant;
  return SWE_OK;
}

int swe_set_tile_rows(SweContext* c, int32_t rows) {
  if (!c) return fail(SWE_EINVAL, "swe_set_tile_rows: null context");
  if (rows < 0 || rows > 4096) return fail(SWE_EINVAL, "swe_set_tile_rows: rows must be in [0, 4096]");
  c->rows = rows == 0 ? default_rows(c->G.nx, c->G.ny) : rows;
  return ensure_partials(c);
}

int swe_set_adjoint_rows(SweContext* c, int32_t rows) {
  if (!c) return fail(SWE_EINVAL, "swe_set_adjoint_rows: null context");
  if (rows < 0 || rows > 4096) return fail(SWE_EINVAL, "swe_set_adjoint_rows: rows must be in [0, 4096]");
  c->adjoint_rows = rows;
  return SWE_OK;
}

int swe_set_waves_per_cu(SweContext* c, int32_t waves) {
  if (!c) return fail(SWE_EINVAL, "swe_set_waves_per_cu: null context");
  if (waves < 0 || (waves > 0 && (waves % kWavesPerBlock != 0 || waves < 2 * kWavesPerBlock)))
    return fail(SWE_EINVAL, "swe_set_waves_per_cu: waves must be 0 or a multiple of 4 >= 8");
  if (waves == 0 || c->lds_per_cu <= 0) {
    c->step_lds = 0;
    return SWE_OK;
  }
  // Reserve enough LDS per workgroup that at most waves/4 workgroups fit on a
  // compute unit (one more would exceed its LDS).
  const int blocks = waves / kWavesPerBlock;
  const unsigned lds = (unsigned)(c->lds_per_cu / (blocks + 1) + 512);
  if ((long)lds * blocks > (long)c->lds_per_cu - 1024 || lds > 65536)
    return fail(SWE_EINVAL, "swe_set_waves_per_cu: cap not reachable with this LDS size");
  c->step_lds = lds;
  return SWE_OK;
}

int swe_set_graph_chunk(SweContext* c, int32_t chunk) {
  if (!c) return fail(SWE_EINVAL, "swe_set_graph_chunk: null context");
  if (chunk < 0 || chunk > 4096 || (chunk & 1))
    return fail(SWE_EINVAL, "swe_set_graph_chunk: chunk must be an even number in [0, 4096]");
  c->graph_chunk = chunk;
  return SWE_OK;
}

int swe_set_loop_finish(SweContext* c, int32_t split) {
  if (!c || split < 0 || split > 2) return fail(SWE_EINVAL, "swe_set_loop_finish: mode must be 0, 1 or 2");
  if (c->loop_active) return fail(SWE_EINVAL, "swe_set_loop_finish: a device loop is active");
  c->split_finish = split;
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
    c->graph_key.clear();
  }
  return SWE_OK;
}

int swe_set_static_classes(SweContext* c, int32_t enable) {
  if (!c || (enable != 0 && enable != 1))
    return fail(SWE_EINVAL, "swe_set_static_classes: enable must be 0 or 1");
  if (c->loop_active) return fail(SWE_EINVAL, "swe_set_static_classes: a device loop is active");
  c->static_classes = enable;
  return SWE_OK;
}

int swe_get_static_layout(SweContext* c, int32_t* layout) {
  if (!c || !layout) return fail(SWE_EINVAL, "swe_get_static_layout: bad arguments");
  double v = 0.0;
  SWE_HIP_TRY(hipSetDevice(c->device));
  // the loop start may have been queued on any stream of the device
  SWE_HIP_TRY(hipDeviceSynchronize());
  SWE_HIP_TRY(hipMemcpy(&v, c->d_sum_mult + 3, sizeof(double), hipMemcpyDeviceToHost));
  *layout = (int32_t)v;
  return SWE_OK;
}

int swe_set_rain_rounding(SweContext* c, int32_t mode) {
  if (!c || (mode != 0 && mode != 1))
    return fail(SWE_EINVAL, "swe_set_rain_rounding: mode must be 0 (nearest) or 1 (stochastic)");
  if (mode == 1 && c->ck_every > 0)
    return fail(SWE_EINVAL, "swe_set_rain_rounding: not while checkpoints are set");
  if (c->loop_active) return fail(SWE_EINVAL, "swe_set_rain_rounding: not while a loop is active");
  c->rain_sr = mode;
  return SWE_OK;
}

int swe_set_sweep_order(SweContext* c, int32_t alternate) {
  if (!c || (alternate != 0 && alternate != 1))
    return fail(SWE_EINVAL, "swe_set_sweep_order: alternate must be 0 or 1");
  c->alt_sweep = alternate;
  return SWE_OK;
}

int swe_set_checkpoints(SweContext* c, float* h, float* hu, float* hv, int32_t every,
                        int32_t capacity) {
  if (!c || every < 0 || capacity < 0 || ((every > 0 && capacity > 0) && (!h || !hu || !hv)))
    return fail(SWE_EINVAL, "swe_set_checkpoints: bad arguments");
  const bool on = every > 0 && capacity > 0;
  if (on && c->rain_sr)
    return fail(SWE_EINVAL,
                "swe_set_checkpoints: the replays of an adjoint round the rain to nearest; "
                "set swe_set_rain_rounding(ctx, 0) first");
  c->ck_h = on ? h : nullptr;
  c->ck_hu = on ? hu : nullptr;
  c->ck_hv = on ? hv : nullptr;
  c->ck_every = on ? every : 0;
  c->ck_cap = on ? capacity : 0;
  return SWE_OK;
}

int swe_set_dt_log(SweContext* c, double* dt_log, int32_t capacity) {
  if (!c || capacity < 0 || (capacity > 0 && !dt_log))
    return fail(SWE_EINVAL, "swe_set_dt_log: bad arguments");
  c->dt_log = capacity > 0 ? dt_log : nullptr;
  c->dt_log_cap = capacity;
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
    c->graph_key.clear();
  }
  return SWE_OK;
}

int swe_rainfall_rate(const SweContext* c, double time_s, double* rate_out) {
  if (!c || !rate_out) return fail(SWE_EINVAL, "swe_rainfall_rate: null argument");
  *rate_out = host_interp(c, time_s);
  return SWE_OK;
}

int swe_step(SweContext* c, SweState in, SweState out, double dt, double time_s, void* stream) {
  if (!c) return fail(SWE_EINVAL, "swe_step: null context");
  if (!valid_state(in) || !valid_state(out)) return fail(SWE_EINVAL, "swe_step: null state pointer");
  if (in.h == out.h || in.hu == out.hu || in.hv == out.hv || in.h == out.hu || in.h == out.hv ||
      in.hu == out.h || in.hu == out.hv || in.hv == out.h || in.hv == out.hu)
    return fail(SWE_EINVAL, "swe_step: input and output state must not alias");
  const double rate = host_interp(c, time_s);
  const float dt_f = (float)dt, dtrate_f = (float)(dt * rate), dtg_f = (float)(dt * c->gravity);
  const LoopConst L = c->loop_const(0, 0);
  c->step_dtx = (float)(dt / c->dx);
  launch_step_any(c, MODE_STEP, in.h, in.hu, in.hv, out.h, out.hu, out.hv, dt_f, dtrate_f, dtg_f,
                  L, nullptr, nullptr, nullptr, (hipStream_t)stream);
  SWE_HIP_TRY(hipGetLastError());
  return SWE_OK;
}

int swe_cfl_field(SweContext* c, SweState in, float* inv_out, void* stream) {
  if (!c || !valid_state(in) || !inv_out) return fail(SWE_EINVAL, "swe_cfl_field: null argument");
  const size_t N = c->cells();
  const int nb = (int)std::min<size_t>((N + kBlock - 1) / kBlock, 4096);
  if (c->variant == SWE_VARIANT_REFERENCE)
    hipLaunchKernelGGL(k_cfl_field<RefMath>, dim3(nb), dim3(kBlock), 0, (hipStream_t)stream, c->G,
                       in.h, in.hu, in.hv, inv_out);
  else
    hipLaunchKernelGGL(k_cfl_field<FastMath>, dim3(nb), dim3(kBlock), 0, (hipStream_t)stream, c->G,
                       in.h, in.hu, in.hv, inv_out);
  SWE_HIP_TRY(hipGetLastError());
  return SWE_OK;
}

int swe_cfl_dt(SweContext* c, SweState in, double* dt_out, void* stream) {
  if (!c || !valid_state(in) || !dt_out) return fail(SWE_EINVAL, "swe_cfl_dt: null argument");
  hipStream_t s = (hipStream_t)stream;
  int rc = ensure_partials(c);
  if (rc) return rc;
  launch_reduce(c, RED_CFL, in.h, in.hu, in.hv, c->loop_const(0, 0), s);
  SWE_HIP_TRY(hipGetLastError());
  SWE_HIP_TRY(hipMemcpyAsync(c->h_ctl, c->d_ctl, sizeof(LoopControl), hipMemcpyDeviceToHost, s));
  SWE_HIP_TRY(hipStreamSynchronize(s));
  const double m = c->h_ctl->max_inv;
  *dt_out = m > 0.0 ? std::min(c->cfl / m, c->max_dt) : c->max_dt;
  return SWE_OK;
}

int swe_run_to_outputs(SweContext* c, SweState initial, const double* times, int32_t n_times,
                       int32_t max_steps, float* out_h, float* out_hu, float* out_hv,
                       SweState final_state, SweRunInfo* info, void* stream) {
  if (!c || !valid_state(initial)) return fail(SWE_EINVAL, "swe_run_to_outputs: null argument");
  if (!times || n_times < 1)
    return fail(SWE_EINVAL, "At least one one-dimensional output time is required.");
  for (int k = 0; k < n_times; ++k)
    if (!std::isfinite(times[k])) return fail(SWE_EINVAL, "Output times must be finite.");
  if (!out_h || !out_hu || !out_hv) return fail(SWE_EINVAL, "swe_run_to_outputs: null output cube");
  if (max_steps < 0) return fail(SWE_EINVAL, "swe_run_to_outputs: max_steps must be >= 0");
  hipStream_t s = (hipStream_t)stream;
  int rc;
  if ((rc = ensure_work(c)) || (rc = ensure_partials(c)) || (rc = ensure_times(c, n_times))) return rc;
  SWE_HIP_TRY(hipMemcpyAsync(c->d_times, times, sizeof(double) * n_times, hipMemcpyHostToDevice, s));
  const size_t N = c->cells();
  if ((rc = stage_in(c, initial, s))) return rc;
  if ((rc = copy_state(c, out_h, out_hu, out_hv, initial.h, initial.hu, initial.hv, s))) return rc;
  const LoopConst L = c->loop_const(n_times, max_steps);
  launch_reduce(c, RED_INIT, initial.h, initial.hu, initial.hv, L, s);
  SWE_HIP_TRY(hipGetLastError());
  // Targets already met at t0 hold the initial state (no step is taken for them).
  int pre = 1;
  while (pre < n_times && times[0] >= times[pre] - kOutputSlack) {
    if ((rc = copy_state(c, out_h + (size_t)pre * N, out_hu + (size_t)pre * N,
                         out_hv + (size_t)pre * N, initial.h, initial.hu, initial.hv, s)))
      return rc;
    ++pre;
  }
  if ((rc = drive_loop(c, L, out_h, out_hu, out_hv, s, max_steps))) return rc;
  const LoopControl k = c->h_ctl[0];
  const int fin = k.steps & 1;  // launch n reads buffer n&1 and writes the other
  for (int slot = k.out_idx; slot < n_times; ++slot) {
    if ((rc = stage_out(c, fin, out_h + (size_t)slot * N, out_hu + (size_t)slot * N,
                        out_hv + (size_t)slot * N, s)))
      return rc;
  }
  if (valid_state(final_state)) {
    if ((rc = stage_out(c, fin, final_state.h, final_state.hu, final_state.hv, s))) return rc;
  }
  SWE_HIP_TRY(hipStreamSynchronize(s));
  fill_info(c, k, info);
  if (k.nonfinite) return fail(SWE_ENONFINITE, "a non-finite value appeared in the state");
  if (k.out_idx != n_times) {
    char msg[256];
    std::snprintf(msg, sizeof msg,
                  "max_steps=%d was insufficient: wrote %d/%d outputs and reached t=%.3fs.",
                  max_steps, k.out_idx, n_times, k.t);
    return fail(SWE_EMAXSTEPS, msg);
  }
  return SWE_OK;
}

int swe_advance_to_time(SweContext* c, SweState in, double t_start, double t_target,
                        int32_t max_substeps, SweState out, int32_t* steps_used, double* t_final,
                        void* stream) {
  if (!c || !valid_state(in) || !valid_state(out))
    return fail(SWE_EINVAL, "swe_advance_to_time: null argument");
  if (!std::isfinite(t_start) || !std::isfinite(t_target))
    return fail(SWE_EINVAL, "swe_advance_to_time: times must be finite");
  if (max_substeps < 0) return fail(SWE_EINVAL, "swe_advance_to_time: max_substeps must be >= 0");
  if (c->loop_active) return fail(SWE_EINVAL, "swe_advance_to_time: a device loop is active");
  hipStream_t s = (hipStream_t)stream;
  int rc;
  if ((rc = ensure_work(c)) || (rc = ensure_partials(c)) || (rc = ensure_times(c, 2))) return rc;
  double times[2] = {t_start, t_target};
  SWE_HIP_TRY(hipMemcpyAsync(c->d_times, times, sizeof times, hipMemcpyHostToDevice, s));
  if ((rc = stage_in(c, in, s))) return rc;
  LoopConst L = c->loop_const(2, max_substeps);
  L.record = 0;
  launch_reduce(c, RED_INIT, in.h, in.hu, in.hv, L, s);
  SWE_HIP_TRY(hipGetLastError());
  // The time arguments stay alive until the copy above has run.
  SWE_HIP_TRY(hipStreamSynchronize(s));
  if ((rc = drive_loop(c, L, nullptr, nullptr, nullptr, s, max_substeps))) return rc;
  const LoopControl k = c->h_ctl[0];
  const int fin = k.steps & 1;  // launch n reads buffer n&1 and writes the other
  if ((rc = stage_out(c, fin, out.h, out.hu, out.hv, s))) return rc;
  // coupled_forecast.py:152-156: a target not reached within 1e-6 s gives NaN.
  if (!(k.t >= t_target - 1.0e-6)) {
    const size_t N = c->cells();
    const int nb = (int)std::min<size_t>((N + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(k_fill3, dim3(nb), dim3(kBlock), 0, s, out.h, out.hu, out.hv, N,
                       __builtin_nanf(""));
    SWE_HIP_TRY(hipGetLastError());
  }
  if (steps_used) *steps_used = k.steps;
  if (t_final) *t_final = k.t;
  SWE_HIP_TRY(hipStreamSynchronize(s));
  return SWE_OK;
}

int swe_replay(SweContext* c, SweState in, const double* dts, const double* ts, int32_t n,
               SweState out, void* stream) {
  if (!c || !valid_state(in) || !valid_state(out) || (n > 0 && (!dts || !ts)) || n < 0)
    return fail(SWE_EINVAL, "swe_replay: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  int rc;
  if ((rc = ensure_work(c))) return rc;
  if (n == 0) return copy_state(c, out.h, out.hu, out.hv, in.h, in.hu, in.hv, s);
  const LoopConst L = c->loop_const(0, 0);
  const float* sh = in.h;
  const float* shu = in.hu;
  const float* shv = in.hv;
  for (int k = 0; k < n; ++k) {
    const int d = k & 1;  // work buffer written by step k (never the input of step k)
    float* dh = c->buf(d, 0);
    float* dhu = c->buf(d, 1);
    float* dhv = c->buf(d, 2);
    const double rate = host_interp(c, ts[k]);
    c->step_dtx = (float)(dts[k] / c->dx);
    launch_step_any(c, MODE_STEP, sh, shu, shv, dh, dhu, dhv, (float)dts[k],
                    (float)(dts[k] * rate), (float)(dts[k] * c->gravity), L, nullptr, nullptr,
                    nullptr, s);
    sh = dh;
    shu = dhu;
    shv = dhv;
  }
  SWE_HIP_TRY(hipGetLastError());
  return copy_state(c, out.h, out.hu, out.hv, sh, shu, shv, s);
}

int swe_replay_store(SweContext* c, SweState in, const double* dts, const double* ts, int32_t n,
                     float* cube_h, float* cube_hu, float* cube_hv, void* stream) {
  if (!c || !valid_state(in) || n < 0 || (n > 0 && (!dts || !ts || !cube_h || !cube_hu || !cube_hv)))
    return fail(SWE_EINVAL, "swe_replay_store: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const size_t N = c->cells();
  const LoopConst L = c->loop_const(0, 0);
  const float* sh = in.h;
  const float* shu = in.hu;
  const float* shv = in.hv;
  for (int k = 0; k < n; ++k) {
    float* dh = cube_h + (size_t)k * N;
    float* dhu = cube_hu + (size_t)k * N;
    float* dhv = cube_hv + (size_t)k * N;
    const double rate = host_interp(c, ts[k]);
    c->step_dtx = (float)(dts[k] / c->dx);
    launch_step_any(c, MODE_STEP, sh, shu, shv, dh, dhu, dhv, (float)dts[k],
                    (float)(dts[k] * rate), (float)(dts[k] * c->gravity), L, nullptr, nullptr,
                    nullptr, s);
    sh = dh;
    shu = dhu;
    shv = dhv;
  }
  SWE_HIP_TRY(hipGetLastError());
  return SWE_OK;
}

int swe_loop_begin(SweContext* c, SweState state, double t0, int32_t max_steps, void* stream) {
  if (!c || !valid_state(state)) return fail(SWE_EINVAL, "swe_loop_begin: null argument");
  if (max_steps < 0) return fail(SWE_EINVAL, "swe_loop_begin: max_steps must be >= 0");
  if (!std::isfinite(t0)) return fail(SWE_EINVAL, "swe_loop_begin: t0 must be finite");
  hipStream_t s = (hipStream_t)stream;
  int rc;
  if ((rc = ensure_work(c)) || (rc = ensure_partials(c)) || (rc = ensure_times(c, 2))) return rc;
  // No output target: the second "output time" is +inf, so dt = CFL dt.
  c->loop_times[0] = t0;
  c->loop_times[1] = HUGE_VAL;
  SWE_HIP_TRY(hipMemcpyAsync(c->d_times, c->loop_times, sizeof c->loop_times,
                             hipMemcpyHostToDevice, s));
  if ((rc = stage_in(c, state, s))) return rc;
  c->loop_L = c->loop_const(2, max_steps);
  launch_reduce(c, RED_INIT, state.h, state.hu, state.hv, c->loop_L, s);
  SWE_HIP_TRY(hipGetLastError());
  c->loop_active = true;
  c->loop_launch_par = 0;
  return SWE_OK;
}

int swe_loop_advance(SweContext* c, int32_t n_launches, void* stream) {
  if (!c || !c->loop_active) return fail(SWE_EINVAL, "swe_loop_advance: call swe_loop_begin first");
  if (n_launches < 0) return fail(SWE_EINVAL, "swe_loop_advance: n_launches must be >= 0");
  if (n_launches == 0) return SWE_OK;
  hipStream_t s = (hipStream_t)stream;
  // Exactly n_launches step kernels.  Launch k reads work state k & 1, so a
  // call that starts on an odd launch issues one plain launch first; whole
  // graph chunks (which start on an even launch) follow, then plain launches
  // for the remainder.
  int left = n_launches;
  if (c->loop_launch_par == 1) {
    launch_loop_step(c, 1, c->loop_L, nullptr, nullptr, nullptr, s);
    c->loop_launch_par = 0;
    --left;
  }
  if (c->graph_chunk > 0 && left >= c->graph_chunk) {
    const int full = (left / c->graph_chunk) * c->graph_chunk;
    if (int rc = launch_loop_chunk(c, full, c->loop_L, nullptr, nullptr, nullptr, s)) return rc;
    left -= full;
  }
  for (int k = 0; k < left; ++k) launch_loop_step(c, k & 1, c->loop_L, nullptr, nullptr, nullptr, s);
  c->loop_launch_par = left & 1;
  SWE_HIP_TRY(hipGetLastError());
  return SWE_OK;
}

int swe_loop_end(SweContext* c, SweState state, SweRunInfo* info, void* stream) {
  if (!c || !c->loop_active || !valid_state(state))
    return fail(SWE_EINVAL, "swe_loop_end: call swe_loop_begin first");
  hipStream_t s = (hipStream_t)stream;
  int rc;
  c->loop_active = false;
  if ((rc = close_loop(c, c->loop_L, s))) return rc;
  if ((rc = copy_final(c, state, s))) return rc;
  if (info) {
    SWE_HIP_TRY(hipMemcpyAsync(c->h_ctl, c->d_ctl, sizeof(LoopControl), hipMemcpyDeviceToHost, s));
    SWE_HIP_TRY(hipStreamSynchronize(s));
    fill_info(c, c->h_ctl[0], info);
    if (c->h_ctl[0].nonfinite) return fail(SWE_ENONFINITE, "a non-finite value appeared in the state");
  }
  return SWE_OK;
}

int swe_run_steps(SweContext* c, SweState state, int32_t n_steps, double t0, SweRunInfo* info,
                  void* stream) {
  int rc;
  if ((rc = swe_loop_begin(c, state, t0, n_steps, stream))) return rc;
  if ((rc = swe_loop_advance(c, n_steps, stream))) return rc;
  return swe_loop_end(c, state, info, stream);
}

}  // extern "C"
