// Batched ensemble of independent shallow-water domains of uneven size.
//
// Every launch of k_ensemble_step advances each unfinished member by one of
// its own adaptive steps.  One launch covers the flattened list of all
// members' strips (62 columns x `rows` rows each), one wavefront per strip,
// so members of 256^2 and 2048^2 cells share one launch and the hardware
// dispatcher balances them.  (A persistent grid sized to the resident
// capacity, each wavefront walking several strips, measured 18 % slower on
// MI355X: every strip then starts with a serial chain of member lookup and
// control loads that no other wavefront hides.)  Like the single-domain step,
// the launch reserves LDS to hold 16 wavefronts per compute unit.
// Each member keeps its own control block (dt, clock, step count) and its
// own per-wavefront partials, written with plain stores; after the launch,
// k_ensemble_finish (one workgroup per running member) reduces each
// member's CFL maximum and FP64 mass terms and plans that member's next
// step, so members advance at their own dt.  (A wavefront-level "last one
// closes the member" handshake needs an agent-scope release per wavefront,
// which on gfx950 writes back the XCD's L2 each time; the kernel boundary
// gives the same visibility once per launch.)  The host captures chunks of
// launches in a HIP graph and polls the members' stop flags between chunks.
//
// The arithmetic per member is exactly the single-domain step
// (swe_strip.hpp / swe_device.hpp), i.e. jax_solver/numerics.py:312-369
// with the device loop of numerics.py:372-434.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/swe_hip.h"
#include "swe_device.hpp"
#include "swe_loop.hpp"
#include "swe_strip.hpp"
#include "swe_fast.hpp"
#include "swe_host.hpp"

namespace swe {

struct Member {
  Grid G;
  LoopConst L;
  LoopControl* ctl;
  float* buf[2];  // two work states of 3N floats: [h | hu | hv], or packed when G.pk is set
  int segs_x, rows, n_waves, wave_base;
};

struct EnsPartials {
  float* inv;
  double* vol;
  double* clip;
  double* bnd;
  int* nf;
};

__device__ __forceinline__ int wave_max_int(int x) {
  for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o, kWave));
  return x;
}

// Member m with wave_base[m] <= w < wave_base[m+1] (wave-uniform).
__device__ __forceinline__ int find_member(const Member* __restrict__ mem, int n, int w) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (mem[mid].wave_base <= w)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// One class table per wavefront of a workgroup (members differ between the
// wavefronts of one workgroup).
__device__ __forceinline__ float* ens_class_tables() {
  __shared__ float t[kWavesPerBlock * 256];
  return t;
}

// PACKED (fast variant): members' work states and static fields in the packed
// layout of the single-domain fast loop (swe_fast.hpp), with the static record
// each member's start chose (the class-indexed one, the (bed, n') one or the
// full one), same arithmetic.
template <class M, bool PACKED = false>
__global__ __launch_bounds__(256) void k_ensemble_step(const Member* __restrict__ mem,
                                                      int n_members, int total_waves, int rev,
                                                      EnsPartials P) {
  const int lane = threadIdx.x & (kWave - 1);
  // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs, so the
  // workgroups of XCD x (blockIdx % 8 == x) take the x-th contiguous eighth
  // of the flattened strip list.  Neighbouring strips then run on one XCD and
  // share their halo rows through its L2.  The host makes gridDim a multiple
  // of 8 and large enough for one strip per wavefront; the stride loop keeps
  // any smaller grid correct.  Every strip belongs to exactly one chunk.
  constexpr int kXcd = 8;
  const int wpb = blockDim.x / kWave;
  const int per_xcd = gridDim.x / kXcd;
  const int chunk = (total_waves + kXcd - 1) / kXcd;
  const int w_lo = (blockIdx.x % kXcd) * chunk;
  const int w_hi = min(w_lo + chunk, total_waves);
  const int stride = per_xcd * wpb;
  for (int wd = w_lo + (blockIdx.x / kXcd) * wpb + wave_in_block(); wd < w_hi; wd += stride) {
    // rev: odd launches sweep the strip list backwards, so each launch starts
    // on the strips the previous one touched last (still in the memory-side
    // cache).  The order does not change any cell's arithmetic.
    const int w = rev ? total_waves - 1 - wd : wd;
    const int m = find_member(mem, n_members, w);
    const Member& mb = mem[m];
    const LoopControl* ctl = mb.ctl;
    if (ctl->done) continue;
    const int local = w - mb.wave_base;
    const size_t N = (size_t)mb.G.nx * (size_t)mb.G.ny;
    const int par = ctl->steps & 1;
    const float* src = mb.buf[par];
    float* dst = mb.buf[par ^ 1];
    // the member's rain added with stochastic rounding (drawn for its step index)
    const int sr_step = (mb.L.rain_sr && ctl->dtrate_f != 0.0f) ? ctl->steps : -1;
    StripAcc acc;
    if constexpr (M::kFast && PACKED) {
      // The member's static record, chosen at its start (swe_ensemble_begin):
      // sum_mult[3] is the record, sum_mult[2] the uniform multiplier.
      const int lay = __builtin_amdgcn_readfirstlane((int)mb.L.sum_mult[3]);
      const float mc = (float)mb.L.sum_mult[2];
      const float dtg = ctl->dtg_f;
      const int sx = local % mb.segs_x, sy = local / mb.segs_x;
      if (lay == kStaticCls) {
        // This wavefront's class table: dtg n^2 of the member's classes,
        // rounded as the single-domain step rounds them.  The wavefront's
        // own LDS accesses complete in order, so the march reads what these
        // stores wrote, and a later strip's refill follows its reads.
        float* tab = ens_class_tables() + wave_in_block() * 256;
        for (int k = lane; k < 256; k += kWave) {
          const float v = mb.G.cls_tab[k];
          tab[k] = signbit(v) ? v : dtg * v;
        }
        march_strip_fast<true, true, kStaticCls>(mb.G, src, nullptr, nullptr, dst, nullptr,
                                                 nullptr, sx, sy, mb.rows, ctl->dtx_f,
                                                 ctl->dtrate_f, dtg, mb.L.dxd, mb.L.dyd, 0, 0,
                                                 nullptr, nullptr, nullptr, acc, mc, tab, sr_step);
      } else if (lay == kStatic8) {
        march_strip_fast<true, true, kStatic8>(mb.G, src, nullptr, nullptr, dst, nullptr, nullptr,
                                               sx, sy, mb.rows, ctl->dtx_f, ctl->dtrate_f, dtg,
                                               mb.L.dxd, mb.L.dyd, 0, 0, nullptr, nullptr, nullptr,
                                               acc, mc, nullptr, sr_step);
      } else {
        march_strip_fast<true, true, kStatic12>(mb.G, src, nullptr, nullptr, dst, nullptr,
                                                nullptr, sx, sy, mb.rows, ctl->dtx_f,
                                                ctl->dtrate_f, dtg, mb.L.dxd, mb.L.dyd, 0, 0,
                                                nullptr, nullptr, nullptr, acc, 0.0f, nullptr,
                                                sr_step);
      }
    } else if constexpr (M::kFast)
      march_strip_fast<true, false>(mb.G, src, src + N, src + 2 * N, dst, dst + N, dst + 2 * N,
                                    local % mb.segs_x, local / mb.segs_x, mb.rows, ctl->dtx_f,
                                    ctl->dtrate_f, ctl->dtg_f, mb.L.dxd, mb.L.dyd, 0, 0, nullptr,
                                    nullptr, nullptr, acc, 0.0f, nullptr, sr_step);
    else
      march_strip<M, true>(mb.G, src, src + N, src + 2 * N, dst, dst + N, dst + 2 * N,
                           local % mb.segs_x, local / mb.segs_x, mb.rows, ctl->dt_f, ctl->dtrate_f,
                           ctl->dtg_f, mb.L.dxd, mb.L.dyd, 0, 0, nullptr, nullptr, nullptr, acc,
                           sr_step);
    // Plain stores: the finish kernel reads them after the kernel boundary.
    const float inv = wave_max(acc.inv);
    const double vol = wave_sum(acc.vol);
    const double clip = wave_sum(acc.clip);
    const double bnd = wave_sum(acc.bnd);
    const int nf = wave_any(acc.nf);
    if (lane == 0) {
      P.inv[w] = inv;
      P.vol[w] = vol;
      P.clip[w] = clip;
      P.bnd[w] = bnd;
      P.nf[w] = nf;
    }
  }
}

// One workgroup per running member: reduce the member's wavefront partials of
// the launch that just ended, close its step in FP64 and plan its next one.
__global__ __launch_bounds__(256) void k_ensemble_finish(const Member* __restrict__ mem,
                                                        EnsPartials P) {
  const Member& mb = mem[blockIdx.x];
  finish_block<256, 4>(mb.ctl, mb.L, P, mb.wave_base, mb.n_waves);
}

}  // namespace swe

// ===========================================================================
// Host side
// ===========================================================================
using namespace swe;

namespace {
int efail(int code, const std::string& m) { return swe::set_error(code, m); }
#define ENS_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return efail(SWE_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)
}  // namespace

struct SweEnsemble {
  int variant = SWE_VARIANT_FAST;
  int rows = 16;
  int graph_chunk = 32;
  std::vector<SweMemberDesc> desc;
  std::vector<Member> host;  // every member, in the caller's order
  std::vector<Member> act;   // the device table: running members, compacted
  int act_rows = 16;
  std::vector<std::vector<double>> rain;  // per member [t | r]
  Member* d_mem = nullptr;
  LoopControl* d_ctl = nullptr;
  LoopControl* h_ctl = nullptr;  // pinned, n members
  double* d_rain = nullptr;      // all members' series
  double* d_times = nullptr;     // per member {0, t_end}
  float* d_work = nullptr;       // all members' A|B states
  char* d_pk = nullptr;          // fast variant: all members' packed static records
  char* d_cls = nullptr;         // fast variant: each member's Manning classes (class_buf_bytes)
  double* d_sm = nullptr;        // per member: multiplier sum, uniform flag and value, record
  double* d_scratch = nullptr;   // 2 x kSumBlocks: scratch of the multiplier sums
  EnsPartials P{};
  int parts_cap = 0;
  int total_waves = 0;
  int grid = 0;
  unsigned step_lds = 0;  // dynamic LDS per step workgroup: caps residency at 16 waves/CU
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t graph = nullptr;
  double times_host[2] = {0.0, 0.0};
  bool active = false;
  std::vector<int> steps_at_begin;
};

namespace {

// Blocks of the multiplier sum: as a single-domain context's reduction
// (at most 2048), so a member's rain volume has the same bits as its solo run.
constexpr int kSumBlocks = 2048;

// Lay out the strips of the members flagged in `running` (all when null)
// into the device member table.  When few cells remain, strips are made
// shorter so that the launch still holds enough wavefronts to hide memory
// latency (at least ~2 per resident slot), down to 4 rows.
int layout(SweEnsemble* e, const std::vector<char>* running = nullptr) {
  int dev = 0, cus = 0, per_cu = 0;
  ENS_TRY(hipGetDevice(&dev));
  ENS_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  if (e->variant == SWE_VARIANT_REFERENCE)
    ENS_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ensemble_step<RefMath>, 256, 0));
  else
    ENS_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ensemble_step<FastMath, true>, 256, 0));
  // Resident wavefronts: at most 4 workgroups (16 wavefronts) per CU under the LDS cap.
  const long slots = (long)cus * std::min(std::max(per_cu, 1), 4) * 4;
  auto waves_for = [&](int rows) {
    long w = 0;
    for (size_t m = 0; m < e->host.size(); ++m) {
      if (running && !(*running)[m]) continue;
      const Member& mb = e->host[m];
      w += (long)((mb.G.nx + kStripCols - 1) / kStripCols) * ((mb.G.ny + rows - 1) / rows);
    }
    return w;
  };
  int rows = e->rows;
  while (rows > 4 && waves_for(rows) < 2 * slots) rows = std::max(4, rows / 2);
  e->act.clear();
  int base = 0;
  for (size_t m = 0; m < e->host.size(); ++m) {
    if (running && !(*running)[m]) continue;
    Member mb = e->host[m];
    mb.rows = rows;
    mb.segs_x = (mb.G.nx + kStripCols - 1) / kStripCols;
    mb.n_waves = mb.segs_x * ((mb.G.ny + rows - 1) / rows);
    mb.wave_base = base;
    base += mb.n_waves;
    e->act.push_back(mb);
  }
  e->act_rows = rows;
  e->total_waves = base;
  if (base > e->parts_cap) {
    if (e->P.vol) ENS_TRY(hipFree(e->P.vol));
    e->P = EnsPartials{};
    e->parts_cap = 0;
    const size_t bytes = (size_t)base * (3 * sizeof(double) + sizeof(float) + sizeof(int));
    char* p = nullptr;
    ENS_TRY(hipMalloc(&p, bytes));
    e->P.vol = reinterpret_cast<double*>(p);
    e->P.clip = e->P.vol + base;
    e->P.bnd = e->P.clip + base;
    e->P.inv = reinterpret_cast<float*>(e->P.bnd + base);
    e->P.nf = reinterpret_cast<int*>(e->P.inv + base);
    e->parts_cap = base;
  }
  if (!e->act.empty())
    ENS_TRY(hipMemcpy(e->d_mem, e->act.data(), sizeof(Member) * e->act.size(), hipMemcpyHostToDevice));
  // One wavefront per strip, in a multiple of 8 workgroups (one share per
  // XCD, see k_ensemble_step).
  const long want = ((long)base + 3) / 4;
  e->grid = (int)std::max<long>(8, (want + 7) / 8 * 8);
  int lds_cu = 0;
  ENS_TRY(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
  const unsigned lds = lds_cu > 0 ? (unsigned)(lds_cu / 5 + 512) : 0u;  // 4 workgroups fit, 5 do not
  e->step_lds = (lds > 0 && lds <= 65536 && 4L * lds <= (long)lds_cu - 1024) ? lds : 0u;
  if (e->graph) {
    (void)hipGraphExecDestroy(e->graph);
    e->graph = nullptr;
  }
  return SWE_OK;
}

void launch_ens(SweEnsemble* e, int rev, hipStream_t s) {
  if (e->variant == SWE_VARIANT_REFERENCE)
    hipLaunchKernelGGL(k_ensemble_step<RefMath>, dim3(e->grid), dim3(256), e->step_lds, s, e->d_mem,
                       (int)e->act.size(), e->total_waves, rev, e->P);
  else
    hipLaunchKernelGGL((k_ensemble_step<FastMath, true>), dim3(e->grid), dim3(256), e->step_lds, s,
                       e->d_mem, (int)e->act.size(), e->total_waves, rev, e->P);
  hipLaunchKernelGGL(k_ensemble_finish, dim3((unsigned)e->act.size()), dim3(256), 0, s, e->d_mem,
                     e->P);
}

int launch_chunks(SweEnsemble* e, int launches, hipStream_t s) {
  if (e->graph_chunk <= 0) {
    for (int k = 0; k < launches; ++k) launch_ens(e, k & 1, s);
    ENS_TRY(hipGetLastError());
    return SWE_OK;
  }
  if (!e->graph) {
    if (!e->cap_stream) ENS_TRY(hipStreamCreateWithFlags(&e->cap_stream, hipStreamNonBlocking));
    ENS_TRY(hipStreamBeginCapture(e->cap_stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < e->graph_chunk; ++k) launch_ens(e, k & 1, e->cap_stream);
    hipGraph_t g = nullptr;
    ENS_TRY(hipStreamEndCapture(e->cap_stream, &g));
    hipError_t err = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (err != hipSuccess) return efail(SWE_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(err));
  }
  for (int k = 0; k < launches; k += e->graph_chunk) ENS_TRY(hipGraphLaunch(e->graph, s));
  return SWE_OK;
}

// Launch per-member init reductions (CFL, volume) that prime each member's control.
// Start of a run, first half: many workgroups per member copy the caller's
// state into the member's work buffer A and write per-workgroup partials of
// the CFL maximum, the FP64 volume and the non-finite flag (slots
// wave_base .. wave_base + gridDim.x - 1 of the partials, gridDim.x <= n_waves).
// With G.pk set (fast variant) the state goes in packed and the member's
// static fields are packed from its current arrays.
template <class MM>
__global__ __launch_bounds__(256) void k_ens_load(const Member* __restrict__ mem, int m,
                                                  const float* __restrict__ h,
                                                  const float* __restrict__ hu,
                                                  const float* __restrict__ hv, EnsPartials P) {
  const Member& mb = mem[m];
  const Grid& G = mb.G;
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  float* A = mb.buf[0];
  float inv = 0.0f;
  double vol = 0.0;
  int nf = 0;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < N;
       k += (size_t)gridDim.x * blockDim.x) {
    const float a0 = h[k], a1 = hu[k], a2 = hv[k];
    const int a = G.mask[k] != 0;
    if (G.pk) {
      A[3 * k] = a0;
      A[3 * k + 1] = a1;
      A[3 * k + 2] = a2;
    } else {
      A[k] = a0;
      A[N + k] = a1;
      A[2 * N + k] = a2;
    }
    inv = fmaxf(inv, cfl_inverse<MM>(G, a0, a1, a2, a));
    vol += a ? (double)a0 : 0.0;
    nf |= !(isfinite(a0) && isfinite(a1) && isfinite(a2));
  }
  const Totals tot = block_reduce(inv, vol, 0.0, 0.0, nf);
  if (threadIdx.x == 0) {
    const int q = mb.wave_base + blockIdx.x;
    P.inv[q] = tot.inv;
    P.vol[q] = tot.vol;
    P.nf[q] = tot.nf;
  }
}

// Second half, one workgroup: reduce the `parts` partials and plan the
// member's first step.
__global__ __launch_bounds__(256) void k_ens_init(const Member* __restrict__ mem, int m,
                                                  int parts, EnsPartials P) {
  const Member& mb = mem[m];
  float inv = 0.0f;
  double vol = 0.0;
  int nf = 0;
  for (int k = threadIdx.x; k < parts; k += blockDim.x) {
    const int q = mb.wave_base + k;
    inv = fmaxf(inv, P.inv[q]);
    vol += P.vol[q];
    nf |= P.nf[q];
  }
  const Totals tot = block_reduce(inv, vol, 0.0, 0.0, nf);
  const double t0 = mb.L.times[0];
  __shared__ int s_j;
  if (threadIdx.x == 0) s_j = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < mb.L.n_rain - 1; k += blockDim.x)
    if (mb.L.rain_t[k] <= t0 && t0 < mb.L.rain_t[k + 1]) s_j = k;
  __syncthreads();
  if (threadIdx.x == 0) {
    LoopControl* ctl = mb.ctl;
    ctl->t = t0;
    ctl->dt = 0.0;
    ctl->steps = 0;
    ctl->vol0 = ctl->vol = tot.vol * mb.L.cell_area;
    ctl->rain = ctl->bnd = ctl->clip = 0.0;
    ctl->nonfinite = tot.nf;
    ctl->max_inv = (double)tot.inv;
    ctl->counter = 0;
    ctl->out_idx = 1;
    if (mb.L.n_times < 2 || mb.L.max_steps <= 0 || tot.nf || t0 >= mb.L.times[1] - kOutputSlack) {
      ctl->done = 1;
      ctl->rec_lo = ctl->rec_hi = 0;
      ctl->ck_slot = -1;
    } else {
      ctl->done = 0;
      plan_next_step(ctl, mb.L, tot.inv, t0, 1, s_j);
    }
  }
}

// Copy a member's final work state ([h | hu | hv], or packed) out to the
// caller's arrays.
__global__ void k_ens_copy(const float* __restrict__ src, float* __restrict__ h,
                           float* __restrict__ hu, float* __restrict__ hv, size_t N, int packed) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < N;
       k += (size_t)gridDim.x * blockDim.x) {
    h[k] = packed ? src[3 * k] : src[k];
    hu[k] = packed ? src[3 * k + 1] : src[N + k];
    hv[k] = packed ? src[3 * k + 2] : src[2 * N + k];
  }
}

}  // namespace

extern "C" {

int swe_ensemble_create(const SweMemberDesc* members, int32_t n, int32_t variant, SweEnsemble** out) {
  if (!members || !out || n < 1) return efail(SWE_EINVAL, "swe_ensemble_create: need >= 1 member");
  *out = nullptr;
  if (variant != SWE_VARIANT_REFERENCE && variant != SWE_VARIANT_FAST)
    return efail(SWE_EINVAL, "swe_ensemble_create: unknown kernel variant");
  for (int m = 0; m < n; ++m) {
    const SweParamsDesc& p = members[m].params;
    if (p.nx < 1 || p.ny < 1 || !p.bed || !p.manning_n || !p.domain_mask || !p.rainfall_multiplier ||
        p.n_rainfall < 1 || !p.rainfall_times_s || !p.rainfall_rates_m_s || !(p.dx > 0) ||
        !(p.dy > 0) || !(p.gravity > 0) || !(p.cfl > 0) || !(p.dry_tolerance > 0) || !(p.max_dt > 0))
      return efail(SWE_EINVAL, "swe_ensemble_create: member " + std::to_string(m) + " has invalid parameters");
    for (int k = 0; k < 4; ++k)
      if (p.boundary_types[k] != SWE_REFLECTIVE && p.boundary_types[k] != SWE_OPEN)
        return efail(SWE_EINVAL, "swe_ensemble_create: bad boundary type");
    for (int k = 1; k < p.n_rainfall; ++k)
      if (!(p.rainfall_times_s[k] > p.rainfall_times_s[k - 1]))
        return efail(SWE_EINVAL, "swe_ensemble_create: rainfall times must increase strictly");
    const SweState& s = members[m].state;
    if (!s.h || !s.hu || !s.hv) return efail(SWE_EINVAL, "swe_ensemble_create: null member state");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return efail(SWE_ENODEV, "swe_ensemble_create: no HIP device is visible");
  SweEnsemble* e = new SweEnsemble();
  e->variant = variant;
  e->desc.assign(members, members + n);
  e->host.resize(n);
  auto bail = [&](int rc) {
    swe_ensemble_destroy(e);
    return rc;
  };
  size_t rain_total = 0, cells_total = 0, pk_total = 0;
  // A member's packed static record takes at most 12 bytes per cell, or the
  // class blocks of a narrow grid; each starts on a 256-byte boundary.
  auto pk_bytes = [](const SweParamsDesc& p) {
    const size_t b = std::max((size_t)12 * p.nx * p.ny, cls_row_bytes(p.nx) * (size_t)p.ny);
    return (b + 255) / 256 * 256;
  };
  for (int m = 0; m < n; ++m) {
    rain_total += 2 * (size_t)members[m].params.n_rainfall;
    cells_total += (size_t)members[m].params.nx * members[m].params.ny;
    pk_total += pk_bytes(members[m].params);
  }
  const size_t cls_bytes = (class_buf_bytes() + 255) / 256 * 256;
  if (hipMalloc(&e->d_rain, sizeof(double) * rain_total) != hipSuccess ||
      hipMalloc(&e->d_sm, sizeof(double) * 4 * n) != hipSuccess ||
      hipMalloc(&e->d_scratch, sizeof(double) * 2 * kSumBlocks) != hipSuccess ||
      hipMalloc(&e->d_times, sizeof(double) * 2 * n) != hipSuccess ||
      hipMalloc(&e->d_ctl, sizeof(LoopControl) * n) != hipSuccess ||
      hipHostMalloc(&e->h_ctl, sizeof(LoopControl) * n, hipHostMallocDefault) != hipSuccess ||
      hipMalloc(&e->d_work, sizeof(float) * 6 * cells_total) != hipSuccess ||
      (variant == SWE_VARIANT_FAST && (hipMalloc(&e->d_pk, pk_total) != hipSuccess ||
                                       hipMalloc(&e->d_cls, cls_bytes * n) != hipSuccess)) ||
      hipMalloc(&e->d_mem, sizeof(Member) * n) != hipSuccess)
    return bail(efail(SWE_EHIP, "swe_ensemble_create: device allocation failed"));
  if (hipMemset(e->d_sm, 0, sizeof(double) * 4 * n) != hipSuccess)
    return bail(efail(SWE_EHIP, "swe_ensemble_create: clear the record slots"));
  if (hipMemset(e->d_ctl, 0, sizeof(LoopControl) * n) != hipSuccess)
    return bail(efail(SWE_EHIP, "swe_ensemble_create: clear control"));
  std::vector<double> rain_all;
  rain_all.reserve(rain_total);
  size_t rain_off = 0, cell_off = 0, pk_off = 0;
  for (int m = 0; m < n; ++m) {
    const SweParamsDesc& p = members[m].params;
    Member& mb = e->host[m];
    Grid& G = mb.G;
    G.nx = p.nx;
    G.ny = p.ny;
    for (int k = 0; k < 4; ++k) G.bt[k] = p.boundary_types[k];
    G.g = (float)p.gravity;
    G.half_g = (float)(0.5 * p.gravity);
    G.dry = (float)p.dry_tolerance;
    G.dx = (float)p.dx;
    G.dy = (float)p.dy;
    G.rdx = 1.0f / G.dx;
    G.rdy = 1.0f / G.dy;
    G.rho = (float)(p.dy / p.dx);
    G.rhox = (float)(p.dx / p.dy);
    G.bed = p.bed;
    G.n = p.manning_n;
    G.mask = p.domain_mask;
    G.mult = p.rainfall_multiplier;
    LoopConst& L = mb.L;
    L.cfl = p.cfl;
    L.max_dt = p.max_dt;
    L.gravity = p.gravity;
    L.cell_area = p.dx * p.dy;
    L.rain_default = p.default_rainfall_m_s;
    L.dxd = p.dx;
    L.dyd = p.dy;
    L.rain_t = e->d_rain + rain_off;
    L.rain_r = e->d_rain + rain_off + p.n_rainfall;
    L.times = e->d_times + 2 * m;
    L.n_rain = p.n_rainfall;
    L.n_times = 2;
    L.record = 0;
    L.rain_sr = 0;
    for (int k = 0; k < p.n_rainfall; ++k) rain_all.push_back(p.rainfall_times_s[k]);
    for (int k = 0; k < p.n_rainfall; ++k) rain_all.push_back(p.rainfall_rates_m_s[k]);
    rain_off += 2 * (size_t)p.n_rainfall;
    mb.ctl = e->d_ctl + m;
    const size_t N = (size_t)p.nx * p.ny;
    mb.buf[0] = e->d_work + cell_off;
    mb.buf[1] = e->d_work + cell_off + 3 * N;
    G.pk = e->d_pk ? reinterpret_cast<const float*>(e->d_pk + pk_off) : nullptr;
    G.cls_tab = e->d_cls ? class_values(e->d_cls + cls_bytes * m) : nullptr;
    pk_off += pk_bytes(p);
    cell_off += 6 * N;
    L.sum_mult = e->d_sm + 4 * m;
  }
  if (hipMemcpy(e->d_rain, rain_all.data(), sizeof(double) * rain_all.size(), hipMemcpyHostToDevice) != hipSuccess)
    return bail(efail(SWE_EHIP, "swe_ensemble_create: upload rain"));
  // Each member's FP64 rain-multiplier sum over active cells and its static
  // record live on the device; swe_ensemble_begin takes them from the
  // member's current arrays.
  int rc = layout(e);
  if (rc) return bail(rc);
  *out = e;
  return SWE_OK;
}

int swe_ensemble_destroy(SweEnsemble* e) {
  if (!e) return SWE_OK;
  if (e->graph) (void)hipGraphExecDestroy(e->graph);
  if (e->cap_stream) (void)hipStreamDestroy(e->cap_stream);
  if (e->P.vol) (void)hipFree(e->P.vol);
  if (e->d_mem) (void)hipFree(e->d_mem);
  if (e->d_work) (void)hipFree(e->d_work);
  if (e->d_pk) (void)hipFree(e->d_pk);
  if (e->d_cls) (void)hipFree(e->d_cls);
  if (e->d_sm) (void)hipFree(e->d_sm);
  if (e->d_scratch) (void)hipFree(e->d_scratch);
  if (e->h_ctl) (void)hipHostFree(e->h_ctl);
  if (e->d_ctl) (void)hipFree(e->d_ctl);
  if (e->d_times) (void)hipFree(e->d_times);
  if (e->d_rain) (void)hipFree(e->d_rain);
  delete e;
  return SWE_OK;
}

int swe_ensemble_set_tile_rows(SweEnsemble* e, int32_t rows) {
  if (!e || rows < 1 || rows > 4096) return efail(SWE_EINVAL, "swe_ensemble_set_tile_rows: rows in [1, 4096]");
  if (e->active) return efail(SWE_EINVAL, "swe_ensemble_set_tile_rows: not while a run is active");
  e->rows = rows;
  return layout(e);
}

int swe_ensemble_set_rain_rounding(SweEnsemble* e, int32_t mode) {
  if (!e || (mode != 0 && mode != 1))
    return efail(SWE_EINVAL, "swe_ensemble_set_rain_rounding: mode must be 0 (nearest) or 1 (stochastic)");
  if (e->active) return efail(SWE_EINVAL, "swe_ensemble_set_rain_rounding: not while a run is active");
  for (Member& mb : e->host) mb.L.rain_sr = mode;
  return layout(e);  // the device member table carries the flag
}

int swe_ensemble_set_graph_chunk(SweEnsemble* e, int32_t chunk) {
  if (!e || chunk < 0 || chunk > 4096) return efail(SWE_EINVAL, "swe_ensemble_set_graph_chunk: chunk in [0, 4096]");
  e->graph_chunk = chunk;
  if (e->graph) {
    (void)hipGraphExecDestroy(e->graph);
    e->graph = nullptr;
  }
  return SWE_OK;
}

int swe_ensemble_begin(SweEnsemble* e, double t_end, int32_t max_steps, void* stream) {
  if (!e) return efail(SWE_EINVAL, "swe_ensemble_begin: null ensemble");
  if (!(t_end > 0.0) || !std::isfinite(t_end) || max_steps < 0)
    return efail(SWE_EINVAL, "swe_ensemble_begin: need t_end > 0 and max_steps >= 0");
  hipStream_t s = (hipStream_t)stream;
  const int n = (int)e->host.size();
  std::vector<double> times(2 * n);
  for (int m = 0; m < n; ++m) {
    times[2 * m] = 0.0;
    times[2 * m + 1] = t_end;
    e->host[m].L.max_steps = max_steps;
  }
  ENS_TRY(hipMemcpy(e->d_times, times.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice));
  int lrc = layout(e);  // every member, in order: table index m == member m
  if (lrc) return lrc;
  const size_t cls_bytes = (class_buf_bytes() + 255) / 256 * 256;
  for (int m = 0; m < n; ++m) {
    const SweState& st = e->desc[m].state;
    const size_t N = (size_t)e->host[m].G.nx * e->host[m].G.ny;
    // The member's multiplier sum (its fields may have changed in place since
    // the last run, and its rain volume must follow them) and, in the fast
    // variant, its static record: the same passes, with the same block count,
    // as a single-domain loop start on the member's grid.
    int prc = prepare_static_record(e->host[m].G, e->d_cls ? e->d_cls + cls_bytes * m : nullptr,
                                    e->d_sm + 4 * m, 1, e->d_scratch,
                                    reinterpret_cast<unsigned*>(e->d_scratch + kSumBlocks),
                                    kSumBlocks, s);
    if (prc) return prc;
    // Workgroups of the load pass: ~8 cells per thread, at most one per
    // partial slot the member owns.
    const int nb = (int)std::max<size_t>(
        1, std::min<size_t>((N + 2047) / 2048, (size_t)e->act[m].n_waves));
    if (e->variant == SWE_VARIANT_REFERENCE)
      hipLaunchKernelGGL(k_ens_load<RefMath>, dim3(nb), dim3(256), 0, s, e->d_mem, m, st.h, st.hu,
                         st.hv, e->P);
    else
      hipLaunchKernelGGL(k_ens_load<FastMath>, dim3(nb), dim3(256), 0, s, e->d_mem, m, st.h, st.hu,
                         st.hv, e->P);
    hipLaunchKernelGGL(k_ens_init, dim3(1), dim3(256), 0, s, e->d_mem, m, nb, e->P);
  }
  ENS_TRY(hipGetLastError());
  e->steps_at_begin.assign(n, 0);
  e->active = true;
  return SWE_OK;
}

int swe_ensemble_advance(SweEnsemble* e, int32_t n_launches, void* stream) {
  if (!e || !e->active) return efail(SWE_EINVAL, "swe_ensemble_advance: call swe_ensemble_begin first");
  if (n_launches < 0) return efail(SWE_EINVAL, "swe_ensemble_advance: n_launches must be >= 0");
  if (n_launches == 0) return SWE_OK;
  const int chunk = e->graph_chunk > 0 ? e->graph_chunk : 1;
  return launch_chunks(e, ((n_launches + chunk - 1) / chunk) * chunk, (hipStream_t)stream);
}

// Synchronises; returns the number of members still running.
int swe_ensemble_poll(SweEnsemble* e, int32_t* running, void* stream) {
  if (!e || !running) return efail(SWE_EINVAL, "swe_ensemble_poll: null argument");
  hipStream_t s = (hipStream_t)stream;
  const int n = (int)e->host.size();
  ENS_TRY(hipMemcpyAsync(e->h_ctl, e->d_ctl, sizeof(LoopControl) * n, hipMemcpyDeviceToHost, s));
  ENS_TRY(hipStreamSynchronize(s));
  int r = 0;
  for (int m = 0; m < n; ++m) r += e->h_ctl[m].done ? 0 : 1;
  *running = r;
  return SWE_OK;
}

// Rebuild the launch over the members still running (host-synchronised: call
// after swe_ensemble_poll, with no launch in flight).
int swe_ensemble_compact(SweEnsemble* e, void* stream) {
  if (!e || !e->active) return efail(SWE_EINVAL, "swe_ensemble_compact: no active run");
  int32_t running = 0;
  int rc = swe_ensemble_poll(e, &running, stream);
  if (rc) return rc;
  std::vector<char> flags(e->host.size());
  for (size_t m = 0; m < flags.size(); ++m) flags[m] = e->h_ctl[m].done ? 0 : 1;
  if (running == 0) return SWE_OK;
  return layout(e, &flags);
}

int swe_ensemble_end(SweEnsemble* e, SweRunInfo* infos, void* stream) {
  if (!e || !e->active) return efail(SWE_EINVAL, "swe_ensemble_end: call swe_ensemble_begin first");
  hipStream_t s = (hipStream_t)stream;
  const int n = (int)e->host.size();
  ENS_TRY(hipMemcpyAsync(e->h_ctl, e->d_ctl, sizeof(LoopControl) * n, hipMemcpyDeviceToHost, s));
  ENS_TRY(hipStreamSynchronize(s));
  int rc = SWE_OK;
  for (int m = 0; m < n; ++m) {
    const LoopControl& k = e->h_ctl[m];
    const SweState& st = e->desc[m].state;
    const size_t N = (size_t)e->host[m].G.nx * e->host[m].G.ny;
    const int nb = (int)std::min<size_t>((N + 255) / 256, 4096);
    hipLaunchKernelGGL(k_ens_copy, dim3(nb), dim3(256), 0, s, e->host[m].buf[k.steps & 1], st.h,
                       st.hu, st.hv, N, e->d_pk ? 1 : 0);
    if (infos) {
      SweRunInfo& I = infos[m];
      I.outputs_written = k.out_idx;
      I.steps_used = k.steps;
      I.final_time_s = k.t;
      I.volume_initial = k.vol0;
      I.volume_final = k.vol;
      I.rain_volume = k.rain;
      I.boundary_volume = k.bnd;
      I.clip_volume = k.clip;
      I.mass_balance_error = k.vol - k.vol0 - k.rain - k.bnd - k.clip;
      const double sc = std::max(std::fabs(k.vol0) + std::fabs(k.rain) + std::fabs(k.bnd) + std::fabs(k.clip), 1.0);
      I.mass_balance_rel = I.mass_balance_error / sc;
      I.nonfinite = k.nonfinite;
      I._pad = 0;
    }
    if (k.nonfinite) rc = efail(SWE_ENONFINITE, "member " + std::to_string(m) + " went non-finite");
    else if (k.out_idx < 2 && rc == SWE_OK)
      rc = efail(SWE_EMAXSTEPS, "member " + std::to_string(m) + " did not reach t_end");
  }
  ENS_TRY(hipGetLastError());
  ENS_TRY(hipStreamSynchronize(s));
  e->active = false;
  return rc;
}

int swe_ensemble_run(SweEnsemble* e, double t_end, int32_t max_steps, SweRunInfo* infos, void* stream) {
  int rc = swe_ensemble_begin(e, t_end, max_steps, stream);
  if (rc) return rc;
  const int chunk = e->graph_chunk > 0 ? e->graph_chunk : 16;
  long launched = 0;
  int last_running = (int)e->host.size();
  while (true) {
    if ((rc = swe_ensemble_advance(e, chunk, stream))) return rc;
    launched += chunk;
    int running = 0;
    if ((rc = swe_ensemble_poll(e, &running, stream))) return rc;
    if (running == 0) break;
    if (running < last_running) {  // drop finished members from the launch
      if ((rc = swe_ensemble_compact(e, stream))) return rc;
      last_running = running;
    }
    if (launched > (long)max_steps + 2L * chunk)
      return efail(SWE_EHIP, "swe_ensemble_run: members did not stop at max_steps");
  }
  return swe_ensemble_end(e, infos, stream);
}

int swe_ensemble_get_static_layout(SweEnsemble* e, int32_t member, int32_t* layout) {
  if (!e || !layout || member < 0 || member >= (int32_t)e->host.size())
    return efail(SWE_EINVAL, "swe_ensemble_get_static_layout: bad arguments");
  // the member's start may have been queued on any stream of the device
  ENS_TRY(hipDeviceSynchronize());
  double v = 0.0;
  ENS_TRY(hipMemcpy(&v, e->d_sm + 4 * member + 3, sizeof(double), hipMemcpyDeviceToHost));
  *layout = (int32_t)v;
  return SWE_OK;
}

int swe_ensemble_total_waves(const SweEnsemble* e, int32_t* waves, int32_t* grid) {
  if (!e || !waves || !grid) return efail(SWE_EINVAL, "swe_ensemble_total_waves: null argument");
  *waves = e->total_waves;
  *grid = e->grid;
  return SWE_OK;
}

}  // extern "C"
