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

// max(a, b) as one v_med3_f32 against +inf.  fmaxf costs one or two extra
// v_max_f32 x, x, x on operands the compiler cannot prove canonical (loads,
// cross-lane moves, selects, loop-carried values): it quiets signalling NaNs
// first, which no value of this solver can be.  The +inf goes through an
// empty asm statement (no instruction) so that the compiler cannot fold the
// med3 back into a max.
__device__ __forceinline__ float max_nc(float a, float b) {
  float inf = __builtin_inff();
  asm("" : "+s"(inf));
  return __builtin_amdgcn_fmed3f(a, b, inf);
}

// True when p holds on every active lane of the wavefront.  The ballot of a
// comparison is the comparison's own lane mask, so this costs scalar
// instructions only.
__device__ __forceinline__ bool wave_all(bool p) {
  return __builtin_amdgcn_ballot_w64(p) == __builtin_amdgcn_read_exec();
}

struct SideF {
  float h, eta, b, u, v;
  float p;  // hydrostatic pressure term g/2 h^2, as (g/2 h) h
  bool a;   // active (inside the domain mask); kept as a lane mask
};

__device__ __forceinline__ SideF make_side(float h, float hu, float hv, float b, bool a,
                                           float dry, float half_g) {
  SideF s;
  // 1/h serves only where h > dry (elsewhere the velocity is selected to 0),
  // so it needs no max(h, dry) guard.
  const float r = __builtin_amdgcn_rcpf(h);
  const bool wet = h > dry;
  s.h = h;
  s.b = b;
  s.eta = h + b;
  s.p = (half_g * h) * h;
  s.u = wet ? hu * r : 0.0f;
  s.v = wet ? hv * r : 0.0f;
  s.a = a;
  return s;
}

// un/ut of each side are (u, v) for x faces and (v, u) for y faces.
template <bool XDIR, bool MASKED>
__device__ __forceinline__ FaceFlux face_fast(const Grid& G, float hA, float etaA, float bA,
                                              float unA, float utA, float pA, bool aA, float hB,
                                              float etaB, float bB, float unB, float utB,
                                              float pB, bool aB) {
  const bool wallA = MASKED && aA && !aB;
  const bool wallB = MASKED && !aA && aB;
  if (MASKED) {
    hB = wallA ? hA : hB;
    etaB = wallA ? etaA : etaB;
    bB = wallA ? bA : bB;
    unB = wallA ? -unA : unB;
    utB = wallA ? utA : utB;
    pB = wallA ? pA : pB;
    hA = wallB ? hB : hA;
    etaA = wallB ? etaB : etaA;
    bA = wallB ? bB : bA;
    unA = wallB ? -unB : unA;
    utA = wallB ? utB : utA;
    pA = wallB ? pB : pA;
  }
  const float dry = G.dry;
  const float bf = max_nc(bA, bB);
  const float hrA = fmaxf(etaA - bf, 0.0f);
  const float hrB = fmaxf(etaB - bf, 0.0f);
  const bool wA = hrA > dry, wB = hrB > dry;
  const float vnA = wA ? unA : 0.0f, vnB = wB ? unB : 0.0f;
  const float QnA = hrA * unA, QnB = hrB * unB;
  const float cA = __builtin_amdgcn_sqrtf(G.g * hrA);
  const float cB = __builtin_amdgcn_sqrtf(G.g * hrB);
  const float sL = fminf(fminf(vnA - cA, vnB - cB), 0.0f);
  const float sR = fmaxf(fmaxf(vnA + cA, vnB + cB), 0.0f);
  // g/2 hr^2 in the same rounding order as the cell's g/2 h^2, so that the
  // source correction below is exactly zero where the reconstruction keeps h
  const float prA = (G.half_g * hrA) * hrA, prB = (G.half_g * hrB) * hrB;
  const float FnA = fmaf(QnA, vnA, prA);
  const float FnB = fmaf(QnB, vnB, prB);
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
  // The mass flux of a wall face (one side mirrored: hrB = hrA, QnB = -QnA,
  // ka = kb) must be exactly zero: two separately rounded products cancel
  // there, a fused one would leave its rounding error.  Every other face
  // takes the fused form, in the masked and the all-active code alike, so a
  // face gives the same bits whichever path its wavefront takes.
  const float mass = fmaf(ka, QnA, kb * QnB);
  F.f0 = fmaf(kc, hrB - hrA, MASKED && (wallA || wallB) ? ka * QnA + kb * QnB : mass);
  F.fn = fmaf(kc, QnB - QnA, fmaf(ka, FnA, kb * FnB));
  F.ft = fmaf(kc, dQt, fmaf(ka, FtA, kb * FtB));
  F.cA = pA - prA;
  F.cB = pB - prB;
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
    return face_fast<XDIR, false>(G, A.h, A.eta, A.b, unA, utA, A.p, A.a, B.h, B.eta, B.b, unB,
                                  utB, B.p, B.a);
  return face_fast<XDIR, true>(G, A.h, A.eta, A.b, unA, utA, A.p, A.a, B.h, B.eta, B.b, unB, utB,
                               B.p, B.a);
}

// Raw inputs of one lane's cell in one row, as loaded (nothing is applied to
// them until they are used, so the loads stay in flight).
struct RawCell {
  float h, hu, hv, b;
  bool a;
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
// n' = n^2 on an active cell and -1 on an inactive one -- the friction law
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
// With a rain multiplier that has one value on every active cell (UNI: the
// reference's default multiplier, the domain mask, and any scalar multiple of
// it) the static record drops it, 8 bytes per cell (bed, n'), and the step
// takes the value as a constant: the rain source is the same fused multiply-add
// with the same operands, so the bits do not change.
struct PackedIO {
  const float* st;  // packed state
  const float* sx;  // packed static fields
  float* out;       // packed output state
  unsigned off12;   // this lane's column, in bytes, in a packed state row
  unsigned offsx;   // this lane's column, in bytes, in a packed static row
  size_t row3;      // floats per packed state row (3 nx)
  size_t rowsx;     // floats per packed static row (3 nx, or 2 nx with UNI)
};

// A wave-uniform pointer, read from lane 0 so that it sits in scalar
// registers even where the compiler holds it in vector ones (a pointer
// loaded from device memory that the kernel also writes).
template <class T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (T*)(((unsigned long long)hi << 32) | lo);
}

template <bool UNI>
__device__ __forceinline__ PackedIO packed_io(const Grid& G, const float* S, float* O, int ic) {
  PackedIO io;
  io.st = uniform_ptr(S);
  io.sx = uniform_ptr(G.pk);
  io.out = uniform_ptr(O);
  io.off12 = (unsigned)ic * 12u;
  io.offsx = (unsigned)ic * (UNI ? 8u : 12u);
  io.row3 = 3 * (size_t)G.nx;
  io.rowsx = (UNI ? 2 : 3) * (size_t)G.nx;
  return io;
}

// base, row3 and jj must be wave-uniform.  The row's address is pinned in
// scalar registers (an empty asm statement, no instruction) so that the
// compiler keeps the scalar-base + 32-bit lane-offset form of the global
// load/store instead of folding the loop-invariant lane offset into a 64-bit
// per-lane pointer that costs a vector add per row.
using GlobalF = __attribute__((address_space(1))) float;

__device__ __forceinline__ const GlobalF* packed_at(const float* base, size_t row3, int jj,
                                                    unsigned off12) {
  const __attribute__((address_space(1))) char* row =
      (const __attribute__((address_space(1))) char*)(base + (size_t)jj * row3);
  asm("" : "+s"(row));
  asm("" : "+v"(off12));  // keeps the 32-bit offset visible at the access
  return (const GlobalF*)(row + off12);
}

// Row j of the lane's column: clamped row index (numerics.py:29-55); the
// reflective-edge sign flips are applied in side_of.  mc: the uniform
// multiplier (UNI only).
template <bool UNI>
__device__ __forceinline__ RawCell load_raw(const Grid& G, const PackedIO& io, int j, float mc) {
  const int jj = min(max(j, 0), G.ny - 1);
  const GlobalF* st = packed_at(io.st, io.row3, jj, io.off12);
  const GlobalF* sx = packed_at(io.sx, io.rowsx, jj, io.offsx);
  RawCell r;
  r.h = st[0];
  r.hu = st[1];
  r.hv = st[2];
  r.b = sx[0];
  r.n = sx[1];
  r.m = UNI ? mc : sx[2];
  r.a = !signbit(r.n);  // sign bit clear: active
  r.neg_hv = (j < 0 && G.bt[2] == 0) || (j >= G.ny && G.bt[3] == 0);
  return r;
}

template <bool UNI>
__device__ __forceinline__ RawCell load_raw(const Grid& G, const StripIO& io, int j, float) {
  const int jj = min(max(j, 0), G.ny - 1);
  RawCell r;
  r.h = io.ld(io.h, jj);
  r.hu = io.ld(io.hu, jj);
  r.hv = io.ld(io.hv, jj);
  r.b = io.ld(io.bed, jj);
  r.a = io.ld_mask(jj) != 0;
  r.neg_hv = (j < 0 && G.bt[2] == 0) || (j >= G.ny && G.bt[3] == 0);
  return r;
}

struct RowCell {
  SideF s;
  float hu, hv;
};

// Cell quantities of a loaded row, with the normal momentum negated beyond a
// reflective edge (numerics.py:29-55).
__device__ __forceinline__ RowCell side_of(const RawCell& r, bool neg_hu, float dry,
                                           float half_g) {
  RowCell c;
  const float hu = neg_hu ? -r.hu : r.hu;
  const float hv = r.neg_hv ? -r.hv : r.hv;
  c.s = make_side(r.h, hu, hv, r.b, r.a, dry, half_g);
  c.hu = hu;
  c.hv = hv;
  return c;
}

// One wavefront's strip: columns sx*62-1 .. sx*62+62, rows sy*rows ..
// sy*rows+rows-1.  sx, sy and rows must be wave-uniform.  The raw inputs of
// row j+2 are in flight while row j is computed.  PACKED: H and Ho are packed
// states (3 floats per cell, HU/HV unused) and the static fields come from
// G.pk; the arithmetic is the same, so both layouts give identical bits.
template <bool LOOP, bool PACKED, bool SOUTH, bool UNI>
__device__ __forceinline__ void march_strip_dir(const Grid& G, const float* __restrict__ H,
                                                 const float* __restrict__ HU,
                                                 const float* __restrict__ HV,
                                                 float* __restrict__ Ho, float* __restrict__ HUo,
                                                 float* __restrict__ HVo, int sx, int sy, int rows,
                                                 float dtx, float dtrate, float dtg, double dxd,
                                                 double dyd, int rec_lo, int rec_hi,
                                                 float* __restrict__ cube_h,
                                                 float* __restrict__ cube_hu,
                                                 float* __restrict__ cube_hv, StripAcc& acc,
                                                 float mc) {
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
  const float dry = G.dry;
  using IO = std::conditional_t<PACKED, PackedIO, StripIO>;
  IO io;
  if constexpr (PACKED)
    io = packed_io<UNI>(G, H, Ho, ic);
  else
    io = strip_io(G, H, HU, HV, max(j0 - 1, 0), ic);
  const size_t out0 = (size_t)j0 * (size_t)G.nx, out_cells = (size_t)(j1 - j0) * (size_t)G.nx;
  Rsrc oh, ohu, ohv;
  if constexpr (!PACKED) {
    oh = make_rsrc(Ho + out0, out_cells * 4);
    ohu = make_rsrc(HUo + out0, out_cells * 4);
    ohv = make_rsrc(HVo + out0, out_cells * 4);
  }

  // The march runs north (rows j0 .. j1-1) or, SOUTH, the other way
  // (j1-1 .. j0); dj is the step.  "Ahead" below means the march direction.
  constexpr int dj = SOUTH ? -1 : 1;
  const int jf = SOUTH ? j1 - 1 : j0;  // first row of the march
  // Prologue: the face behind row jf, rows jf (current), jf+dj and jf+2dj
  // (in flight).
  RowCell cur = side_of(load_raw<UNI>(G, io, jf - dj, mc), neg_hu, dry, G.half_g);
  RawCell next = load_raw<UNI>(G, io, jf, mc);
  RawCell r2;
  float yc0, yc1, yc2;  // carried y face: between the current row and the one behind it
  float nn, mult;
  {
    const RowCell c0 = side_of(next, neg_hu, dry, G.half_g);
    if constexpr (PACKED) {
      nn = next.n;
      mult = next.m;
    }
    next = load_raw<UNI>(G, io, jf + dj, mc);
    r2 = load_raw<UNI>(G, io, SOUTH ? max(jf - 2, j0 - 1) : min(jf + 2, j1), mc);
    if constexpr (!SOUTH) {  // south face of row j0
      const FaceFlux fy = face_fast_any<false>(G, cur.s, c0.s, wave_all(cur.s.a && c0.s.a));
      if (LOOP && j0 == 0 && own) acc.bnd += (double)fy.f0 * dxd;
      yc0 = fy.f0;
      yc1 = fy.ft;
      yc2 = fy.fn + fy.cB;
    } else {  // north face of row j1-1
      const FaceFlux fy = face_fast_any<false>(G, c0.s, cur.s, wave_all(cur.s.a && c0.s.a));
      if (LOOP && j1 == G.ny && own) acc.bnd -= (double)fy.f0 * dxd;
      yc0 = fy.f0;
      yc1 = fy.ft;
      yc2 = fy.fn + fy.cA;
    }
    cur = c0;
  }
  if constexpr (!PACKED) {
    nn = io.ld(io.n, jf);
    mult = io.ld(io.mult, jf);
  }
  float nf_probe = 0.0f;
  float clip32 = 0.0f;  // clip depth is small; float32 per lane over the strip

  // Row j: `cur` is row j, `north` row j+dj (the next row of the march,
  // loaded two rows ago; row j+2dj is in flight too).  Issues the loads of
  // row j+3dj into `ahead`, takes row j+dj's n/multiplier into
  // nn_next/mult_next, and leaves row j+dj's side in `up`.
  auto row = [&](int j, const RowCell& cur, RowCell& up, const RawCell& north, RawCell& ahead,
                 float nn_j, float mult_j, float& nn_next, float& mult_next) {
    // Unconditional, so the wait counts stay exact; past the strip the
    // clamped rows were just loaded and come from the cache.
    ahead = load_raw<UNI>(G, io, SOUTH ? max(j - 3, j0 - 1) : min(j + 3, j1), mc);
    if constexpr (PACKED) {  // row j+dj's fields arrived with its state
      nn_next = north.n;
      mult_next = north.m;
    } else {
      const int jn = SOUTH ? max(j - 1, j0) : min(j + 1, j1 - 1);
      nn_next = io.ld(io.n, jn);
      mult_next = io.ld(io.mult, jn);
    }

    // Active flag of row j.  In the packed layout it is re-derived from the
    // sign of row j's n' (one compare, straight into a lane mask) rather than
    // carried over from the previous row, which the compiler does through a
    // vector register and two conversions.
    const bool ca = PACKED ? !signbit(nn_j) : cur.s.a;

    // x face between lane-1 (A) and this lane (B).  Every face of lanes
    // 1..63 joins two active cells when the active mask, shifted by one
    // lane, still covers those lanes: scalar work on the ballot.
    const float weta = from_west(cur.s.eta), wb = from_west(cur.s.b), wu = from_west(cur.s.u),
                wv = from_west(cur.s.v), wp = from_west(cur.s.p);
    const unsigned long long am = __builtin_amdgcn_ballot_w64(ca);
    const unsigned long long need = __builtin_amdgcn_read_exec() & ~1ull;
    FaceFlux fx;
    if (((am << 1) & am & need) == need) {
      fx = face_fast<true, false>(G, 0.0f, weta, wb, wu, wv, wp, true, cur.s.h, cur.s.eta,
                                  cur.s.b, cur.s.u, cur.s.v, cur.s.p, true);
    } else {
      const bool wa = from_west((int)ca) != 0;
      fx = face_fast<true, true>(G, from_west(cur.s.h), weta, wb, wu, wv, wp, wa, cur.s.h,
                                 cur.s.eta, cur.s.b, cur.s.u, cur.s.v, cur.s.p, ca);
    }
    const float xr0 = fx.f0, xr1 = fx.fn + fx.cB, xr2 = fx.ft;
    const float xl0 = from_east(fx.f0);
    const float xl1 = from_east(fx.fn + fx.cA);
    const float xl2 = from_east(fx.ft);
    if (edge_x) {
      if (i == 0) acc.bnd += (double)fx.f0 * dyd;
      if (i == G.nx) acc.bnd -= (double)fx.f0 * dyd;
    }
    // The y face ahead of row j; a y face's A side is always its south cell.
    up = side_of(north, neg_hu, dry, G.half_g);
    const SideF& ys = SOUTH ? up.s : cur.s;  // south cell of the face
    const SideF& yn = SOUTH ? cur.s : up.s;  // north cell
    const bool as = SOUTH ? up.s.a : ca, an = SOUTH ? ca : up.s.a;
    FaceFlux gy;
    if (wave_all(ca && up.s.a))
      gy = face_fast<false, false>(G, ys.h, ys.eta, ys.b, ys.v, ys.u, ys.p, true, yn.h, yn.eta,
                                   yn.b, yn.v, yn.u, yn.p, true);
    else
      gy = face_fast<false, true>(G, ys.h, ys.eta, ys.b, ys.v, ys.u, ys.p, as, yn.h, yn.eta, yn.b,
                                  yn.v, yn.u, yn.p, an);
    // Row j's share of the new face (its own source correction) and the
    // next row's share, carried to it.
    const float yf0 = gy.f0, yf1 = gy.ft, yf2 = gy.fn + (SOUTH ? gy.cB : gy.cA);
    const float yn0 = SOUTH ? yc0 : yf0, yn1 = SOUTH ? yc1 : yf1, yn2 = SOUTH ? yc2 : yf2;  // north
    const float ys0 = SOUTH ? yf0 : yc0, ys1 = SOUTH ? yf1 : yc1, ys2 = SOUTH ? yf2 : yc2;  // south

    // dt ((xl - xr)/dx + (yn - ys)/dy) as (dt/dx) ((xl - xr) + (yn - ys) dx/dy)
    const float d0 = fmaf(yn0 - ys0, G.rhox, xl0 - xr0);
    const float d1 = fmaf(yn1 - ys1, G.rhox, xl1 - xr1);
    const float d2 = fmaf(yn2 - ys2, G.rhox, xl2 - xr2);
    const float hs = fmaf(dtrate, mult_j, fmaf(-dtx, d0, cur.s.h));
    const float hus = fmaf(-dtx, d1, cur.hu);
    const float hvs = fmaf(-dtx, d2, cur.hv);
    const float hn = fmaxf(hs, 0.0f);
    const float hd = fmaxf(hs, dry);  // = max(hn, dry): dry > 0
    const float rhd = __builtin_amdgcn_rcpf(hd);
    const float q = __builtin_amdgcn_sqrtf(fmaf(hus, hus, hvs * hvs));
    // K = dtg n^2 q / hd^(7/3), the power as 2^(-7/3 log2 hd) on the hardware log/exp
    const float ip73 = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(hd) * -kPow73);
    const float n2 = PACKED ? nn_j : nn_j * nn_j;  // the packed field holds n^2
    const float K = ((dtg * n2) * q) * ip73;
    const float rD = __builtin_amdgcn_rcpf(1.0f + K);
    const bool wet = ca && (hn > dry);
    const float h_o = ca ? hn : 0.0f;
    const float hu_o = wet ? hus * rD : 0.0f;
    const float hv_o = wet ? hvs * rD : 0.0f;
    if (own) {
      if constexpr (PACKED) {
        GlobalF* o = const_cast<GlobalF*>(packed_at(io.out, io.row3, j, io.off12));
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
          // j through an empty asm statement: keeps the compiler from
          // strength-reducing these addresses into per-lane pointers that
          // every row would advance.
          int jr = j;
          asm("" : "+s"(jr));
          const size_t idx = (size_t)jr * (size_t)G.nx + (size_t)ic;
          for (int k = rec_lo; k < rec_hi; ++k) {
            cube_h[(size_t)k * N + idx] = h_o;
            cube_hu[(size_t)k * N + idx] = hu_o;
            cube_hv[(size_t)k * N + idx] = hv_o;
          }
        }
        // CFL of the new state (numerics.py:299-303), u = hu/max(h,dry) where
        // wet (hu is already 0 where not wet), in units of 1/dy:
        // (|u| + c) dy/dx + (|v| + c); the lane's maximum is scaled by 1/dy
        // once, after the march (cfl_inverse<FastMath> takes the same form).
        const float ws = __builtin_amdgcn_sqrtf(G.g * h_o);
        const float su = fmaf(fabsf(hu_o), rhd, ws);
        const float sv = fmaf(fabsf(hv_o), rhd, ws);
        acc.inv = max_nc(acc.inv, ca ? fmaf(su, G.rho, sv) : 0.0f);
        acc.vol += (double)h_o;
        clip32 += ca ? hn - hs : 0.0f;  // max(-hs, 0), exactly
        // (x + y + z) * 0 stays 0 unless one of them is infinite or NaN
        nf_probe = fmaf(h_o + hu_o + hv_o, 0.0f, nf_probe);
      }
    }
    yc0 = gy.f0;
    yc1 = gy.ft;
    yc2 = gy.fn + (SOUTH ? gy.cA : gy.cB);
  };

  // Four rows per trip with the raw rows rotating through four sets of
  // registers: a row's raw inputs stay live over four rows (loaded three
  // rows ahead, read as the next row, then as the current row), so fewer
  // sets would make the compiler copy one at every loop edge, and wait there
  // for the loads in flight.  The derived row values alternate between two.
  // The raw sets, by the row (relative to the trip's first row j) they hold
  // when row j starts: next = j+1, r2 = j+2 (both from the prologue or the
  // previous trip), r3 and r1 are filled with rows j+3 and j+4 on the way.
  RowCell c1;
  RawCell r1, r3;
  float nn1 = 0.0f, mult1 = 0.0f;
  const int n = j1 - j0;
  int k = 0;
  for (; k + 3 < n; k += 4) {
    const int j = jf + k * dj;
    row(j, cur, c1, next, r3, nn, mult, nn1, mult1);
    row(j + dj, c1, cur, r2, r1, nn1, mult1, nn, mult);
    row(j + 2 * dj, cur, c1, r3, next, nn, mult, nn1, mult1);
    row(j + 3 * dj, c1, cur, r1, r2, nn1, mult1, nn, mult);
  }
  if (k < n) row(jf + k * dj, cur, c1, next, r3, nn, mult, nn1, mult1);
  if (k + 1 < n) row(jf + (k + 1) * dj, c1, cur, r2, r1, nn1, mult1, nn, mult);
  if (k + 2 < n) row(jf + (k + 2) * dj, cur, c1, r3, next, nn, mult, nn1, mult1);
  // yc0 now holds the mass flux through the last face of the march: the north
  // face of row j1-1, or (SOUTH) the south face of row j0.
  if constexpr (!SOUTH) {
    if (LOOP && j1 == G.ny && own) acc.bnd -= (double)yc0 * dxd;
  } else {
    if (LOOP && j0 == 0 && own) acc.bnd += (double)yc0 * dxd;
  }
  if (LOOP) {
    acc.nf |= !isfinite(nf_probe);
    acc.clip += (double)clip32;
    acc.inv *= G.rdy;  // max(x) / dy = max(x / dy): rounding is monotonic
  }
}

// One wavefront's strip (see march_strip_dir).  Strips of odd row bands march
// south and the others north, so two strips that share halo rows (the last
// two rows of one band and the first two of the next) read them near the
// same time -- both at the start or both at the end of their marches -- and
// the second read is an L2 hit.  The direction does not change any cell's
// arithmetic: every face takes the same operands, and a cell combines its
// faces in the same order (north minus south, east minus west).
template <bool LOOP, bool PACKED = false, bool UNI = false>
__device__ __forceinline__ void march_strip_fast(const Grid& G, const float* __restrict__ H,
                                                 const float* __restrict__ HU,
                                                 const float* __restrict__ HV,
                                                 float* __restrict__ Ho, float* __restrict__ HUo,
                                                 float* __restrict__ HVo, int sx, int sy, int rows,
                                                 float dtx, float dtrate, float dtg, double dxd,
                                                 double dyd, int rec_lo, int rec_hi,
                                                 float* __restrict__ cube_h,
                                                 float* __restrict__ cube_hu,
                                                 float* __restrict__ cube_hv, StripAcc& acc,
                                                 float mc = 0.0f) {
  static_assert(PACKED || !UNI, "the uniform-multiplier record is a packed layout");
  if (sy & 1)
    march_strip_dir<LOOP, PACKED, true, UNI>(G, H, HU, HV, Ho, HUo, HVo, sx, sy, rows, dtx, dtrate,
                                             dtg, dxd, dyd, rec_lo, rec_hi, cube_h, cube_hu,
                                             cube_hv, acc, mc);
  else
    march_strip_dir<LOOP, PACKED, false, UNI>(G, H, HU, HV, Ho, HUo, HVo, sx, sy, rows, dtx, dtrate,
                                              dtg, dxd, dyd, rec_lo, rec_hi, cube_h, cube_hu,
                                              cube_hv, acc, mc);
}

}  // namespace swe
