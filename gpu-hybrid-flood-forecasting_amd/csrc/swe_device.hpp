// Device building blocks of the fused shallow-water step for gfx950.
//
// One wavefront owns a strip of 64 consecutive columns (62 updated, one halo
// column each side) and marches through a band of rows (north; in the fast
// variant, odd bands march south).  The cell to the west comes from the
// neighbouring lane by a DPP wave shift, the next row is the next row's load,
// so every face flux is evaluated once per strip and no LDS or barrier is
// needed inside the march.
//
// The arithmetic is the reference step (jax_solver/numerics.py:58-369) in
// face-normal coordinates.  Two policies evaluate it:
//   RefMath  - the reference evaluation order with correctly rounded division
//              and square root (the build uses -ffp-contract=off), so depth
//              and the wet mask are bit-identical with the FP32 oracle;
//   FastMath - one hardware reciprocal per quotient, hardware sqrt/log2/exp2,
//              and the HLL velocities reused instead of re-divided
//              ((hr*u)/hr == u up to one rounding).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swe {

constexpr int kWave = 64;
constexpr int kStripCols = kWave - 2;  // columns updated per wavefront
constexpr float kPow73 = 2.33333325386047363281f;        // float32(7/3)
constexpr float kPow73MinusTwo = 0.33333325386047363281f; // float32(7/3) - 2, exact
constexpr double kOutputSlack = 1.0e-7;  // numerics.py:401

// Static per-grid constants, passed by value to every kernel.
struct Grid {
  int nx, ny;
  int bt[4];  // left, right, bottom, top; 0 reflective, 1 open
  float g, half_g, dry, dx, dy, rdx, rdy;
  float rho;   // dy / dx (the fast variant's CFL form)
  float rhox;  // dx / dy (the fast variant's divergence form)
  const float* __restrict__ bed;
  const float* __restrict__ n;
  const uint8_t* __restrict__ mask;
  const float* __restrict__ mult;
  // Packed static fields of the fast device loop: per cell (bed, n', rain
  // multiplier), (bed, n') or the bed and a one-byte class (swe_fast.hpp), n'
  // = n^2 for an active cell and -1 for an inactive one (null where unused).
  const float* __restrict__ pk;
  // Class table of the class-indexed static record (256 floats: n^2 of each
  // Manning value of the active cells; null where unused).
  const float* __restrict__ cls_tab;
};

// ---------------------------------------------------------------------------
// Arithmetic policies
// ---------------------------------------------------------------------------
struct RefMath {
  static constexpr bool kFast = false;
  __device__ static __forceinline__ float sqrt(float x) { return __builtin_sqrtf(x); }
  // h > dry ? q / max(h, dry) : 0   (numerics.py:58-59)
  __device__ static __forceinline__ void velocities(float h, float qx, float qy, float dry,
                                                    float& u, float& v) {
    const float d = fmaxf(h, dry);
    const bool wet = h > dry;
    u = wet ? qx / d : 0.0f;
    v = wet ? qy / d : 0.0f;
  }
  // HLL re-derives the velocity from the reconstructed momentum (numerics.py:72-75).
  __device__ static __forceinline__ float hll_velocity(float hr, float q, float /*u*/, float dry) {
    return hr > dry ? q / fmaxf(hr, dry) : 0.0f;
  }
  __device__ static __forceinline__ float quot(float a, float b, float /*rb*/) { return a / b; }
  __device__ static __forceinline__ float recip(float /*b*/) { return 0.0f; }
  __device__ static __forceinline__ float per_dx(float a, float dx, float /*rdx*/) { return a / dx; }
  // max(h, dry) ** float32(7/3)  (numerics.py:361)
  __device__ static __forceinline__ float pow73(float x) { return powf(x, kPow73); }
};

struct FastMath {
  static constexpr bool kFast = true;
  __device__ static __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
  __device__ static __forceinline__ void velocities(float h, float qx, float qy, float dry,
                                                    float& u, float& v) {
    const float r = __builtin_amdgcn_rcpf(fmaxf(h, dry));
    const bool wet = h > dry;
    u = wet ? qx * r : 0.0f;
    v = wet ? qy * r : 0.0f;
  }
  __device__ static __forceinline__ float hll_velocity(float hr, float /*q*/, float u, float dry) {
    return hr > dry ? u : 0.0f;
  }
  __device__ static __forceinline__ float quot(float a, float /*b*/, float rb) { return a * rb; }
  __device__ static __forceinline__ float recip(float b) { return __builtin_amdgcn_rcpf(b); }
  __device__ static __forceinline__ float per_dx(float a, float /*dx*/, float rdx) { return a * rdx; }
  // x^(7/3) = x*x * 2^(log2(x) * (7/3 - 2)); x >= dry > 0 is a normal float.
  __device__ static __forceinline__ float pow73(float x) {
    return (x * x) * __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * kPow73MinusTwo);
  }
};

// ---------------------------------------------------------------------------
// Cells with ghost semantics
// ---------------------------------------------------------------------------
struct Cell {
  float h, hu, hv, b, u, v;
  int a;
};

// Load cell (j, i) where j in [-1, ny] and i in [-1, nx] index the ghosted
// grid: ghost cells copy the edge cell and negate the normal momentum on
// reflective sides (boundary_conditions.py:36-51; bed and mask are edge-padded,
// numerics.py:325-326).  Indices outside that range are clamped; their values
// are finite and unused.
template <class M>
__device__ __forceinline__ Cell load_cell(const Grid& G, const float* __restrict__ H,
                                          const float* __restrict__ HU,
                                          const float* __restrict__ HV, int j, int i) {
  const int jj = min(max(j, 0), G.ny - 1);
  const int ii = min(max(i, 0), G.nx - 1);
  const size_t idx = (size_t)jj * (size_t)G.nx + (size_t)ii;
  Cell c;
  c.h = H[idx];
  c.hu = HU[idx];
  c.hv = HV[idx];
  c.b = G.bed[idx];
  c.a = G.mask[idx] != 0;
  if ((i < 0 && G.bt[0] == 0) || (i >= G.nx && G.bt[1] == 0)) c.hu = -c.hu;
  if ((j < 0 && G.bt[2] == 0) || (j >= G.ny && G.bt[3] == 0)) c.hv = -c.hv;
  M::velocities(c.h, c.hu, c.hv, G.dry, c.u, c.v);
  return c;
}

// ---------------------------------------------------------------------------
// One interface: wall mirroring, hydrostatic reconstruction, HLL
// ---------------------------------------------------------------------------
struct FaceFlux {
  float f0, fn, ft;  // mass, normal-momentum and tangential-momentum flux
  float cA, cB;      // pressure corrections of the A-side and B-side cell
};

// Side A is west (x faces) or south (y faces).  un/ut are the normal and
// tangential velocities of the original depth of each side.  XDIR selects the
// x-face tangential flux q_n * v_t' (numerics.py:84) instead of the y-face
// q_t * v_n' (numerics.py:120).  Follows numerics.py:162-224 / 227-283.
// MASKED = false is the same arithmetic for a face whose two cells are both
// active (no wall mirroring, always relevant); callers use it when a whole
// wavefront's faces are interior to the domain.
template <class M, bool XDIR, bool MASKED = true>
__device__ __forceinline__ FaceFlux face_flux(const Grid& G, float hA, float unA, float utA,
                                              float bA, int aA, float hB, float unB,
                                              float utB, float bB, int aB) {
  if (MASKED) {
    // An active/inactive face is a wall: the inactive side becomes the mirror
    // image of the active one (numerics.py:183-192).
    const bool wallA = aA && !aB;
    const bool wallB = !aA && aB;
    hB = wallA ? hA : hB;
    unB = wallA ? -unA : unB;
    utB = wallA ? utA : utB;
    bB = wallA ? bA : bB;
    hA = wallB ? hB : hA;
    unA = wallB ? -unB : unA;
    utA = wallB ? utB : utA;
    bA = wallB ? bB : bA;
  }

  const float dry = G.dry;
  const float bf = fmaxf(bA, bB);
  const float hrA = fmaxf((hA + bA) - bf, 0.0f);
  const float hrB = fmaxf((hB + bB) - bf, 0.0f);
  const float QnA = hrA * unA, QtA = hrA * utA;
  const float QnB = hrB * unB, QtB = hrB * utB;

  const float vnA = M::hll_velocity(hrA, QnA, unA, dry);
  const float vnB = M::hll_velocity(hrB, QnB, unB, dry);
  const float cA = M::sqrt(G.g * hrA);
  const float cB = M::sqrt(G.g * hrB);
  const float sA = fminf(vnA - cA, vnB - cB);
  const float sB = fmaxf(vnA + cA, vnB + cB);

  float FtA, FtB;
  if (XDIR) {
    FtA = QnA * M::hll_velocity(hrA, QtA, utA, dry);
    FtB = QnB * M::hll_velocity(hrB, QtB, utB, dry);
  } else {
    FtA = QtA * vnA;
    FtB = QtB * vnB;
  }
  const float FnA = QnA * vnA + G.half_g * (hrA * hrA);
  const float FnB = QnB * vnB + G.half_g * (hrB * hrB);

  const float spread = sB - sA;
  const float den = spread > 1.0e-12f ? spread : 1.0f;
  const float rden = M::recip(den);
  const float sAB = sA * sB;
  const float m0 = M::quot((sB * QnA - sA * QnB) + sAB * (hrB - hrA), den, rden);
  const float mn = M::quot((sB * FnA - sA * FnB) + sAB * (QnB - QnA), den, rden);
  const float mt = M::quot((sB * FtA - sA * FtB) + sAB * (QtB - QtA), den, rden);
  const bool upA = sA >= 0.0f;
  const bool upB = sB <= 0.0f;
  const bool rel = !MASKED || aA || aB;
  FaceFlux F;
  F.f0 = rel ? (upA ? QnA : (upB ? QnB : m0)) : 0.0f;
  F.fn = rel ? (upA ? FnA : (upB ? FnB : mn)) : 0.0f;
  F.ft = rel ? (upA ? FtA : (upB ? FtB : mt)) : 0.0f;
  F.cA = rel ? G.half_g * (hA * hA - hrA * hrA) : 0.0f;
  F.cB = rel ? G.half_g * (hB * hB - hrB * hrB) : 0.0f;
  return F;
}

// ---------------------------------------------------------------------------
// Stochastic rounding of the rain source (opt-in, no reference counterpart)
// ---------------------------------------------------------------------------
// The rain of one step adds the same increment to every ponded cell, so a
// round-to-nearest addition makes the same error on every cell of a binade
// and the errors add up over the grid instead of averaging out: this is the
// float32 state's drift of the FP64 mass balance.  Stochastic rounding
// rounds a + r to one of its two float neighbours with the probabilities that
// make the expected result exact, drawing from a hash of the cell's index and
// the step's index: deterministic, and the same for every tile, sweep order
// and launch.  The hash is "lowbias32" (C. Wellons, "Prospecting for hash
// functions", 2018).
__host__ __device__ __forceinline__ unsigned sr_hash(unsigned x) {
  x ^= x >> 16;
  x *= 0x7FEB352Du;
  x ^= x >> 15;
  x *= 0x846CA68Bu;
  x ^= x >> 16;
  return x;
}

// Seed of step `step` (0-based index of the step in its run).
__host__ __device__ __forceinline__ unsigned sr_seed(int step) {
  return sr_hash((unsigned)step * 0x9E3779B9u + 0x632BE5ABu);
}

// a + r rounded stochastically with the 32 random bits `bits`.  The exact
// error e of the nearest sum s (TwoSum, exact in float32 without overflow)
// decides the direction; s moves one unit in the last place toward a + r
// when u * gap < |e|, u uniform in [0, 1) on 24 bits: probability |e| / gap.
__device__ __forceinline__ float sr_add(float a, float r, unsigned bits) {
  const float s = a + r;
  const float bb = s - a;
  const float e = (a - (s - bb)) + (r - bb);
  // away from zero when e has the sign of s, else toward it (s = 0 gives e = 0)
  const int dir = (e > 0.0f) == (s >= 0.0f) ? 1 : -1;
  const float t = __int_as_float(__float_as_int(s) + dir);
  const float gap = fabsf(t - s);
  const float u = (float)(bits >> 8) * 5.9604644775390625e-8f;  // 2^-24
  return (e != 0.0f && u * gap < fabsf(e)) ? t : s;
}

// Pointwise part of the update (numerics.py:350-369) for one cell, given the
// flux divergences.  Returns the new state; h_star (before the clip) is kept
// for the mass balance.  sr: the rain is added with sr_add and the random
// bits `srb` (the reference adds it rounded to nearest).
struct CellUpdate {
  float h, hu, hv, h_star;
  bool wet;
};

template <class M>
__device__ __forceinline__ CellUpdate cell_update(const Grid& G, float h, float hu, float hv,
                                                  float div0, float div1, float div2, float n,
                                                  float mult, int active, float dt,
                                                  float dtrate, float dtg, bool sr = false,
                                                  unsigned srb = 0u) {
  CellUpdate r;
  float hs = h - dt * div0;
  const float hus = hu - dt * div1;
  const float hvs = hv - dt * div2;
  hs = sr ? sr_add(hs, dtrate * mult, srb) : hs + dtrate * mult;
  const float hn = fmaxf(hs, 0.0f);
  const float q = M::sqrt(hus * hus + hvs * hvs);
  const float num = (dtg * (n * n)) * q;
  const float den = M::pow73(fmaxf(hn, G.dry));
  float D;
  if (M::kFast) {
    D = 1.0f + num * __builtin_amdgcn_rcpf(den);
  } else {
    D = 1.0f + num / den;
  }
  const float rD = M::recip(D);
  const float hu_n = M::quot(hus, D, rD);
  const float hv_n = M::quot(hvs, D, rD);
  const bool wet = active && (hn > G.dry);
  r.h = active ? hn : 0.0f;
  r.hu = wet ? hu_n : 0.0f;
  r.hv = wet ? hv_n : 0.0f;
  r.h_star = hs;
  r.wet = wet;
  return r;
}

// Inverse CFL time of one cell (numerics.py:299-303); 0 outside the mask.
// The fast variant evaluates ((|u| + c) dy/dx + (|v| + c)) / dy with |u| =
// |hu| / max(h, dry) as one fused multiply-add on the reciprocal, the form
// the fast step's epilogue takes for the state it writes, so both give the
// same bits for the same state.
template <class M>
__device__ __forceinline__ float cfl_inverse(const Grid& G, float h, float hu, float hv,
                                             int active) {
  if constexpr (M::kFast) {
    const float r = __builtin_amdgcn_rcpf(fmaxf(h, G.dry));
    const bool wet = h > G.dry;
    const float w = M::sqrt(G.g * fmaxf(h, 0.0f));
    const float su = fmaf(fabsf(wet ? hu : 0.0f), r, w);
    const float sv = fmaf(fabsf(wet ? hv : 0.0f), r, w);
    return active ? fmaf(su, G.rho, sv) * G.rdy : 0.0f;
  } else {
    float u, v;
    M::velocities(h, hu, hv, G.dry, u, v);
    const float w = M::sqrt(G.g * fmaxf(h, 0.0f));
    const float inv = M::per_dx(fabsf(u) + w, G.dx, G.rdx) + M::per_dx(fabsf(v) + w, G.dy, G.rdy);
    return active ? inv : 0.0f;
  }
}

// ---------------------------------------------------------------------------
// Cross-lane moves inside a wavefront (gfx9 DPP wave shifts)
// ---------------------------------------------------------------------------
// Lane l receives lane l-1's value (lane 0 receives 0).  bound_ctrl makes
// the lanes without a source read zero, so no old value has to be set up.
__device__ __forceinline__ int from_west(int x) {
  return __builtin_amdgcn_mov_dpp(x, 0x138, 0xF, 0xF, true);
}
__device__ __forceinline__ float from_west(float x) {
  return __builtin_bit_cast(float, from_west(__builtin_bit_cast(int, x)));
}
// Lane l receives lane l+1's value (lane 63 receives 0).
__device__ __forceinline__ float from_east(float x) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0x130, 0xF, 0xF, true));
}

// Workgroups are dealt round-robin to the 8 XCDs, each with its own L2.
// Renumbering them so that every XCD sweeps one contiguous band of the
// launch keeps neighbouring strips, which share halo rows and the cache
// lines at their edges, in one L2.  A bijection on [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb) {
  constexpr int kXcd = 8;
  const int per = nb / kXcd, rem = nb % kXcd;
  const int x = b % kXcd, k = b / kXcd;
  return x * per + min(x, rem) + k;
}

// Doubles move as two 32-bit halves.
__device__ __forceinline__ double from_west(double x) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = (unsigned)from_west((int)(unsigned)b);
  const unsigned hi = (unsigned)from_west((int)(unsigned)(b >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double from_east(double x) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, 0x130, 0xF, 0xF, true);
  const unsigned hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0x130, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Index of the calling wavefront within its workgroup, as a wave-uniform
// (scalar) value so that everything derived from it stays in SGPRs.
__device__ __forceinline__ int wave_in_block() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
}

// Wavefront reductions (all lanes receive the result).
//
// Maximum of non-negative floats (the CFL inverses; no NaN), taken on their
// bit patterns, which order as the values do: DPP row shifts by 1, 2, 4 and 8
// leave each 16-lane row's maximum in its lane 15, two row broadcasts carry
// it to lane 63.  A lane without a source reads 0, the identity here.
__device__ __forceinline__ float wave_max(float x) {
  unsigned u = __float_as_uint(x);
  u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x111, 0xF, 0xF, true));  // row_shr:1
  u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x112, 0xF, 0xF, true));  // row_shr:2
  u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x114, 0xF, 0xF, true));  // row_shr:4
  u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x118, 0xF, 0xF, true));  // row_shr:8
  u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x142, 0xA, 0xF, false));  // row_bcast:15
  u = max(u, (unsigned)__builtin_amdgcn_update_dpp(0, (int)u, 0x143, 0xC, 0xF, false));  // row_bcast:31
  return __uint_as_float(__builtin_amdgcn_readlane(u, kWave - 1));
}
__device__ __forceinline__ double wave_sum(double x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, kWave);
  return x;
}
// 1 if x is nonzero on any active lane, else 0 (the flags reduced here are 0 or 1).
__device__ __forceinline__ int wave_any(int x) {
  return __builtin_amdgcn_ballot_w64(x != 0) != 0 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Rainfall interpolation shared by host and device
// ---------------------------------------------------------------------------
// np.interp(t, x, y, left=d, right=d) for one point: the float64 rain rate of
// numpy_solver/shallow_water_2d.py:254-264 given the bracketing index j with
// x[j] <= t < x[j+1] (found by the caller).  Same formula as NumPy's
// compiled interpolation so host, device and NumPy agree bit for bit.
__host__ __device__ inline double interp_at(const double* x, const double* y, int n, double d,
                                            double t, int j) {
  if (!(t >= x[0])) return (t < x[0]) ? d : t;  // NaN propagates
  if (t > x[n - 1]) return d;
  if (t == x[n - 1]) return y[n - 1];
  if (x[j] == t) return y[j];
  const double slope = (y[j + 1] - y[j]) / (x[j + 1] - x[j]);
  double r = slope * (t - x[j]) + y[j];
  if (r != r) {
    r = slope * (t - x[j + 1]) + y[j + 1];
    if (r != r && y[j] == y[j + 1]) r = y[j];
  }
  return r;
}

}  // namespace swe
