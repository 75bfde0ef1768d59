// Throughput form of the fused step (SWE_VARIANT_FAST).
//
// Same discretisation as the reference-order march (swe_strip.hpp), with the
// arithmetic arranged for the VALU of gfx950 rather than for bit parity:
//   - per cell, once: eta = h + b, h^2 and the velocities from one reciprocal;
//   - HLL in its clamped-speed form F = (sR' FA - sL' FB + sL' sR' dU)/(sR' - sL')
//     with sL' = min(sL, 0), sR' = max(sR, 0), which equals the upwind choice
//     of numerics.py:107-111 and needs no selects;
//   - fused multiply-adds everywhere except the mass flux, whose two products
//     stay separately rounded so that a wall (mirrored) face carries exactly
//     zero mass: a*q + b*(-q) with a == b;
//   - one reciprocal of max(h, dry) for the CFL epilogue; x^(-7/3) of the
//     friction law as 2^(-7/3 log2 x) on the hardware log/exp;
//   - depth summed in float64 per cell, so the volume does not depend on the
//     strip shape; the (small) clip depth per lane in float32.
// Results stay within the stated tolerance of the FP64 oracle; the lake at
// rest, wall mass flux and open-boundary accounting keep their exact forms.
#pragma once

#include <type_traits>

#include "swe_device.hpp"
#include "swe_strip.hpp"

namespace swe {

struct SideF {
  float h, eta, b, u, v, h2;
  int a;
};

__device__ __forceinline__ SideF make_side(float h, float hu, float hv, float b, int a,
                                           float dry) {
  SideF s;
  const float r = __builtin_amdgcn_rcpf(fmaxf(h, dry));
  const bool wet = h > dry;
  s.h = h;
  s.b = b;
  s.eta = h + b;
  s.h2 = h * h;
  s.u = wet ? hu * r : 0.0f;
  s.v = wet ? hv * r : 0.0f;
  s.a = a;
  return s;
}

// un/ut of each side are (u, v) for x faces and (v, u) for y faces.
template <bool XDIR, bool MASKED>
__device__ __forceinline__ FaceFlux face_fast(const Grid& G, float hA, float etaA, float bA,
                                              float unA, float utA, float h2A, int aA, float hB,
                                              float etaB, float bB, float unB, float utB,
                                              float h2B, int aB) {
  if (MASKED) {
    const bool wallA = aA && !aB;
    const bool wallB = !aA && aB;
    hB = wallA ? hA : hB;
    etaB = wallA ? etaA : etaB;
    bB = wallA ? bA : bB;
    unB = wallA ? -unA : unB;
    utB = wallA ? utA : utB;
    h2B = wallA ? h2A : h2B;
    hA = wallB ? hB : hA;
    etaA = wallB ? etaB : etaA;
    bA = wallB ? bB : bA;
    unA = wallB ? -unB : unA;
    utA = wallB ? utB : utA;
    h2A = wallB ? h2B : h2A;
  }
  const float dry = G.dry;
  const float bf = fmaxf(bA, bB);
  const float hrA = fmaxf(etaA - bf, 0.0f);
  const float hrB = fmaxf(etaB - bf, 0.0f);
  const bool wA = hrA > dry, wB = hrB > dry;
  const float vnA = wA ? unA : 0.0f, vnB = wB ? unB : 0.0f;
  const float QnA = hrA * unA, QnB = hrB * unB;
  const float cA = __builtin_amdgcn_sqrtf(G.g * hrA);
  const float cB = __builtin_amdgcn_sqrtf(G.g * hrB);
  const float sL = fminf(fminf(vnA - cA, vnB - cB), 0.0f);
  const float sR = fmaxf(fmaxf(vnA + cA, vnB + cB), 0.0f);
  const float hr2A = hrA * hrA, hr2B = hrB * hrB;
  const float FnA = fmaf(QnA, vnA, G.half_g * hr2A);
  const float FnB = fmaf(QnB, vnB, G.half_g * hr2B);
  float FtA, FtB, dQt;
  if (XDIR) {
    const float vtA = wA ? utA : 0.0f, vtB = wB ? utB : 0.0f;
    FtA = QnA * vtA;
    FtB = QnB * vtB;
    dQt = fmaf(hrB, utB, -hrA * utA);
  } else {
    const float QtA = hrA * utA, QtB = hrB * utB;
    FtA = QtA * vnA;
    FtB = QtB * vnB;
    dQt = QtB - QtA;
  }
  const float spread = sR - sL;
  const float r = __builtin_amdgcn_rcpf(spread > 1.0e-12f ? spread : 1.0f);
  const float ka = sR * r;          // weight of the A-side flux
  const float kb = -sL * r;         // weight of the B-side flux
  const float kc = (sL * sR) * r;   // weight of the state jump
  FaceFlux F;
  F.f0 = fmaf(kc, hrB - hrA, ka * QnA + kb * QnB);
  F.fn = fmaf(kc, QnB - QnA, fmaf(ka, FnA, kb * FnB));
  F.ft = fmaf(kc, dQt, fmaf(ka, FtA, kb * FtB));
  F.cA = G.half_g * (h2A - hr2A);
  F.cB = G.half_g * (h2B - hr2B);
  if (MASKED) {
    const bool rel = aA || aB;
    F.f0 = rel ? F.f0 : 0.0f;
    F.fn = rel ? F.fn : 0.0f;
    F.ft = rel ? F.ft : 0.0f;
    F.cA = rel ? F.cA : 0.0f;
    F.cB = rel ? F.cB : 0.0f;
  }
  return F;
}

template <bool XDIR>
__device__ __forceinline__ FaceFlux face_fast_any(const Grid& G, const SideF& A, const SideF& B,
                                                  bool all_active) {
  const float unA = XDIR ? A.u : A.v, utA = XDIR ? A.v : A.u;
  const float unB = XDIR ? B.u : B.v, utB = XDIR ? B.v : B.u;
  if (all_active)
    return face_fast<XDIR, false>(G, A.h, A.eta, A.b, unA, utA, A.h2, A.a, B.h, B.eta, B.b, unB,
                                  utB, B.h2, B.a);
  return face_fast<XDIR, true>(G, A.h, A.eta, A.b, unA, utA, A.h2, A.a, B.h, B.eta, B.b, unB, utB,
                               B.h2, B.a);
}

// Raw inputs of one lane's cell in one row, as loaded (nothing is applied to
// them until they are used, so the loads stay in flight).
struct RawCell {
  float h, hu, hv, b;
  unsigned a;
  bool neg_hv;  // row beyond a reflective south/north edge
  float n, m;   // Manning n and rain multiplier (packed layout only)
};

// Per-strip addressing through raw buffer resources.  Each array gets a
// resource based at the strip's lowest row; a row is selected by a
// wave-uniform byte offset (scalar register) and each lane adds its fixed
// column offset, so no vector address arithmetic runs per row.  Loads past a
// resource's end read 0 and stores past it are dropped.
using Rsrc = __amdgpu_buffer_rsrc_t;

__device__ __forceinline__ Rsrc make_rsrc(const void* base, size_t bytes) {
  const int records = bytes > 0xFFFFFFFFull ? -1 : (int)(unsigned)bytes;  // uint32 field
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, records,
                                           0x00020000);
}

struct StripIO {
  Rsrc h, hu, hv, bed, n, mult, mask;
  unsigned off4, off1;  // this lane's column, in bytes, in a float / uint8 row
  unsigned pitch;       // bytes per float row
  int jlo;              // row the resources start at

  __device__ __forceinline__ unsigned row4(int jj) const { return (unsigned)(jj - jlo) * pitch; }
  __device__ __forceinline__ float ld(Rsrc r, int jj) const {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off4, row4(jj), 0));
  }
  __device__ __forceinline__ unsigned ld_mask(int jj) const {
    return __builtin_amdgcn_raw_buffer_load_b8(mask, off1, (unsigned)(jj - jlo) * (pitch >> 2), 0);
  }
};

__device__ __forceinline__ StripIO strip_io(const Grid& G, const float* H, const float* HU,
                                            const float* HV, int jlo, int ic) {
  StripIO io;
  const size_t row0 = (size_t)jlo * (size_t)G.nx;
  const size_t cells = (size_t)(G.ny - jlo) * (size_t)G.nx;
  io.h = make_rsrc(H + row0, cells * 4);
  io.hu = make_rsrc(HU + row0, cells * 4);
  io.hv = make_rsrc(HV + row0, cells * 4);
  io.bed = make_rsrc(G.bed + row0, cells * 4);
  io.n = make_rsrc(G.n + row0, cells * 4);
  io.mult = make_rsrc(G.mult + row0, cells * 4);
  io.mask = make_rsrc(G.mask + row0, cells);
  io.off4 = (unsigned)ic * 4u;
  io.off1 = (unsigned)ic;
  io.pitch = (unsigned)G.nx * 4u;
  io.jlo = jlo;
  return io;
}

// Packed layout of the fast device loop: 12 bytes per cell for the work
// state (h, hu, hv) and 12 for the static fields (bed, n', multiplier), where
// n' = |n| on an active cell and -1 on an inactive one -- the friction law
// uses n^2 only, and the sign carries the domain mask.  A row costs two
// 96-bit loads per lane instead of seven scalar ones, and the step streams
// three arrays through HBM instead of ten.
// The packed rows go through plain global pointers: three adjacent floats per
// lane compile to one global_load_dwordx3 / global_store_dwordx3.  (The raw
// buffer forms are not usable here with the ROCm 7.2 compiler: element reads
// after __builtin_bit_cast of the 96-bit load builtin are miscompiled into
// one 32-bit load standing for all three elements; the whole-vector form is
// split back into 32-bit loads; the volatile form is made system-coherent,
// which bypasses the L2 that neighbouring strips share.)
struct PackedIO {
  const float* st;  // packed state
  const float* sx;  // packed static fields
  unsigned off12;   // this lane's column, in bytes, in a packed row
  size_t row3;      // floats per packed row (3 nx)
};

__device__ __forceinline__ PackedIO packed_io(const Grid& G, const float* S, int ic) {
  PackedIO io;
  io.st = S;
  io.sx = G.pk;
  io.off12 = (unsigned)ic * 12u;
  io.row3 = 3 * (size_t)G.nx;
  return io;
}

__device__ __forceinline__ const float* packed_at(const float* base, size_t row3, int jj,
                                                  unsigned off12) {
  return (const float*)((const char*)(base + (size_t)jj * row3) + off12);
}

// Row j of the lane's column: clamped row index (numerics.py:29-55); the
// reflective-edge sign flips are applied in side_of.
__device__ __forceinline__ RawCell load_raw(const Grid& G, const PackedIO& io, int j) {
  const int jj = min(max(j, 0), G.ny - 1);
  const float* st = packed_at(io.st, io.row3, jj, io.off12);
  const float* sx = packed_at(io.sx, io.row3, jj, io.off12);
  RawCell r;
  r.h = st[0];
  r.hu = st[1];
  r.hv = st[2];
  r.b = sx[0];
  r.n = sx[1];
  r.m = sx[2];
  r.a = signbit(r.n) ? 0u : 1u;  // sign bit clear: active
  r.neg_hv = (j < 0 && G.bt[2] == 0) || (j >= G.ny && G.bt[3] == 0);
  return r;
}

__device__ __forceinline__ RawCell load_raw(const Grid& G, const StripIO& io, int j) {
  const int jj = min(max(j, 0), G.ny - 1);
  RawCell r;
  r.h = io.ld(io.h, jj);
  r.hu = io.ld(io.hu, jj);
  r.hv = io.ld(io.hv, jj);
  r.b = io.ld(io.bed, jj);
  r.a = io.ld_mask(jj);
  r.neg_hv = (j < 0 && G.bt[2] == 0) || (j >= G.ny && G.bt[3] == 0);
  return r;
}

struct RowCell {
  SideF s;
  float hu, hv;
};

// Cell quantities of a loaded row, with the normal momentum negated beyond a
// reflective edge (numerics.py:29-55).
__device__ __forceinline__ RowCell side_of(const RawCell& r, bool neg_hu, float dry) {
  RowCell c;
  const float hu = neg_hu ? -r.hu : r.hu;
  const float hv = r.neg_hv ? -r.hv : r.hv;
  c.s = make_side(r.h, hu, hv, r.b, r.a != 0, dry);
  c.hu = hu;
  c.hv = hv;
  return c;
}

// One wavefront's strip: columns sx*62-1 .. sx*62+62, rows sy*rows ..
// sy*rows+rows-1.  sx, sy and rows must be wave-uniform.  The raw inputs of
// row j+2 are in flight while row j is computed.  PACKED: H and Ho are packed
// states (3 floats per cell, HU/HV unused) and the static fields come from
// G.pk; the arithmetic is the same, so both layouts give identical bits.
template <bool LOOP, bool PACKED = false>
__device__ __forceinline__ void march_strip_fast(const Grid& G, const float* __restrict__ H,
                                                 const float* __restrict__ HU,
                                                 const float* __restrict__ HV,
                                                 float* __restrict__ Ho, float* __restrict__ HUo,
                                                 float* __restrict__ HVo, int sx, int sy, int rows,
                                                 float dt, float dtrate, float dtg, double dxd,
                                                 double dyd, int rec_lo, int rec_hi,
                                                 float* __restrict__ cube_h,
                                                 float* __restrict__ cube_hu,
                                                 float* __restrict__ cube_hv, StripAcc& acc) {
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  const int lane = threadIdx.x & (kWave - 1);
  const int i = sx * kStripCols - 1 + lane;
  const bool own = lane >= 1 && lane <= kStripCols && i < G.nx;
  const int ic = min(max(i, 0), G.nx - 1);
  const bool neg_hu = (i < 0 && G.bt[0] == 0) || (i >= G.nx && G.bt[1] == 0);
  // Only the first and last strip of a row band touch the west/east edge faces.
  const bool edge_x = LOOP && (sx == 0 || (sx + 1) * kStripCols >= G.nx);
  const int j0 = sy * rows;
  const int j1 = min(j0 + rows, G.ny);
  const float rdx = G.rdx, rdy = G.rdy, dry = G.dry;
  using IO = std::conditional_t<PACKED, PackedIO, StripIO>;
  IO io;
  if constexpr (PACKED)
    io = packed_io(G, H, ic);
  else
    io = strip_io(G, H, HU, HV, max(j0 - 1, 0), ic);
  const size_t out0 = (size_t)j0 * (size_t)G.nx, out_cells = (size_t)(j1 - j0) * (size_t)G.nx;
  Rsrc oh, ohu, ohv;
  if constexpr (!PACKED) {
    oh = make_rsrc(Ho + out0, out_cells * 4);
    ohu = make_rsrc(HUo + out0, out_cells * 4);
    ohv = make_rsrc(HVo + out0, out_cells * 4);
  }

  // Prologue: the south face of row j0, rows j0 (current) and j0+1 (in flight).
  RowCell cur = side_of(load_raw(G, io, j0 - 1), neg_hu, dry);
  RawCell next = load_raw(G, io, j0);
  float yt0, yt1, yt2;
  float nn, mult;
  {
    const RowCell c0 = side_of(next, neg_hu, dry);
    if constexpr (PACKED) {
      nn = next.n;
      mult = next.m;
    }
    next = load_raw(G, io, j0 + 1);
    const FaceFlux fy = face_fast_any<false>(G, cur.s, c0.s, __all(cur.s.a && c0.s.a));
    if (LOOP && j0 == 0 && own) acc.bnd += (double)fy.f0 * dxd;
    yt0 = fy.f0;
    yt1 = fy.ft;
    yt2 = fy.fn + fy.cB;
    cur = c0;
  }
  if constexpr (!PACKED) {
    nn = io.ld(io.n, j0);
    mult = io.ld(io.mult, j0);
  }
  float nf_probe = 0.0f;
  float clip32 = 0.0f;  // clip depth is small; float32 per lane over the strip

  // Row j: `cur` is row j, `north` row j+1 (loaded one row ago).  Issues the
  // loads of row j+2 into `ahead` and of row j+1's n/multiplier into
  // nn_next/mult_next, and leaves row j+1's side in `up`.
  auto row = [&](int j, const RowCell& cur, RowCell& up, const RawCell& north, RawCell& ahead,
                 float nn_j, float mult_j, float& nn_next, float& mult_next) {
    // Unconditional, so the wait counts stay exact; past the strip the
    // clamped rows were just loaded and come from the cache.
    ahead = load_raw(G, io, min(j + 2, j1));
    if constexpr (PACKED) {  // row j+1's fields arrived with its state
      nn_next = north.n;
      mult_next = north.m;
    } else {
      nn_next = io.ld(io.n, min(j + 1, j1 - 1));
      mult_next = io.ld(io.mult, min(j + 1, j1 - 1));
    }

    SideF w;
    w.h = from_west(cur.s.h);
    w.eta = from_west(cur.s.eta);
    w.b = from_west(cur.s.b);
    w.u = from_west(cur.s.u);
    w.v = from_west(cur.s.v);
    w.h2 = from_west(cur.s.h2);
    w.a = from_west(cur.s.a);
    const FaceFlux fx = face_fast_any<true>(G, w, cur.s, __all(lane == 0 || (w.a && cur.s.a)));
    const float xr0 = fx.f0, xr1 = fx.fn + fx.cB, xr2 = fx.ft;
    const float xl0 = from_east(fx.f0);
    const float xl1 = from_east(fx.fn + fx.cA);
    const float xl2 = from_east(fx.ft);
    if (edge_x) {
      if (i == 0) acc.bnd += (double)fx.f0 * dyd;
      if (i == G.nx) acc.bnd -= (double)fx.f0 * dyd;
    }
    up = side_of(north, neg_hu, dry);
    const FaceFlux gy = face_fast_any<false>(G, cur.s, up.s, __all(cur.s.a && up.s.a));
    const float yb0 = gy.f0, yb1 = gy.ft, yb2 = gy.fn + gy.cA;

    const float d0 = fmaf(yb0 - yt0, rdy, (xl0 - xr0) * rdx);
    const float d1 = fmaf(yb1 - yt1, rdy, (xl1 - xr1) * rdx);
    const float d2 = fmaf(yb2 - yt2, rdy, (xl2 - xr2) * rdx);
    const float hs = fmaf(dtrate, mult_j, fmaf(-dt, d0, cur.s.h));
    const float hus = fmaf(-dt, d1, cur.hu);
    const float hvs = fmaf(-dt, d2, cur.hv);
    const float hn = fmaxf(hs, 0.0f);
    const float hd = fmaxf(hn, dry);
    const float rhd = __builtin_amdgcn_rcpf(hd);
    const float q = __builtin_amdgcn_sqrtf(fmaf(hus, hus, hvs * hvs));
    // K = dtg n^2 q / hd^(7/3), the power as 2^(-7/3 log2 hd) on the hardware log/exp
    const float ip73 = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(hd) * -kPow73);
    const float K = ((dtg * (nn_j * nn_j)) * q) * ip73;
    const float rD = __builtin_amdgcn_rcpf(1.0f + K);
    const bool wet = cur.s.a && (hn > dry);
    const float h_o = cur.s.a ? hn : 0.0f;
    const float hu_o = wet ? hus * rD : 0.0f;
    const float hv_o = wet ? hvs * rD : 0.0f;
    if (own) {
      if constexpr (PACKED) {
        float* o = const_cast<float*>(packed_at(Ho, io.row3, j, io.off12));
        o[0] = h_o;
        o[1] = hu_o;
        o[2] = hv_o;
      } else {
        const unsigned so = (unsigned)(j - j0) * io.pitch;
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, h_o), oh, io.off4, so,
                                              0);
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, hu_o), ohu, io.off4,
                                              so, 0);
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, hv_o), ohv, io.off4,
                                              so, 0);
      }
      if (LOOP) {
        // Output or checkpoint slots (rare steps; rec_lo/rec_hi are
        // wave-uniform, so this is a scalar branch).
        if (__builtin_expect(rec_lo < rec_hi, 0)) {
          const size_t idx = (size_t)j * (size_t)G.nx + (size_t)ic;
          for (int k = rec_lo; k < rec_hi; ++k) {
            cube_h[(size_t)k * N + idx] = h_o;
            cube_hu[(size_t)k * N + idx] = hu_o;
            cube_hv[(size_t)k * N + idx] = hv_o;
          }
        }
        // CFL of the new state (numerics.py:299-303): u = hu/max(h,dry) where
        // wet, and hu is already 0 where not wet.
        const float ws = __builtin_amdgcn_sqrtf(G.g * h_o);
        const float inv = fmaf(fabsf(hv_o * rhd) + ws, rdy, (fabsf(hu_o * rhd) + ws) * rdx);
        acc.inv = fmaxf(acc.inv, cur.s.a ? inv : 0.0f);
        acc.vol += (double)h_o;
        clip32 += cur.s.a ? fmaxf(-hs, 0.0f) : 0.0f;
        // (x + y + z) * 0 stays 0 unless one of them is infinite or NaN
        nf_probe = fmaf(h_o + hu_o + hv_o, 0.0f, nf_probe);
      }
    }
    yt0 = gy.f0;
    yt1 = gy.ft;
    yt2 = gy.fn + gy.cB;
  };

  // Two rows per trip with the carried values alternating between two sets
  // of registers, so nothing is copied at the loop edge.
  RowCell alt;
  RawCell next2;
  float nn2 = 0.0f, mult2 = 0.0f;
  int j = j0;
  for (; j + 1 < j1; j += 2) {
    row(j, cur, alt, next, next2, nn, mult, nn2, mult2);
    row(j + 1, alt, cur, next2, next, nn2, mult2, nn, mult);
  }
  if (j < j1) row(j, cur, alt, next, next2, nn, mult, nn2, mult2);
  // yt0 now holds the mass flux through the north face of row j1-1.
  if (LOOP && j1 == G.ny && own) acc.bnd -= (double)yt0 * dxd;
  if (LOOP) {
    acc.nf |= !isfinite(nf_probe);
    acc.clip += (double)clip32;
  }
}

}  // namespace swe
