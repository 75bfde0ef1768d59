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

// Lane mask of the cells whose packed n' has its sign bit clear (active),
// straight from the comparison: a ballot of a bool that lives across loop
// iterations goes through a vector register (two extra vector instructions).
__device__ __forceinline__ unsigned long long active_lanes(float np) {
  constexpr int kSignedGreater = 38;
  return __builtin_amdgcn_sicmp(__float_as_int(np), -1, kSignedGreater);
}

// a[lane + 1] - b (lane 63: 0 - b), as one subtraction with a DPP operand.
// Written out because the compiler keeps some of these as a move and a
// subtraction.  A DPP operand must not be read within two wait states of the
// vector instruction that wrote it (the s_nop).
__device__ __forceinline__ float east_minus(float a, float b) {
  float d;
  asm("s_nop 1\n\tv_sub_f32_dpp %0, %1, %2 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1"
      : "=v"(d)
      : "v"(a), "v"(b));
  return d;
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
  // Tangential flux hr un vt of each side as (hr ut) vn: the same three
  // factors (vn and vt are zero together, where the side is not wet), one
  // select fewer than with vt, and the same form for x and y faces.
  const float QtA = hrA * utA, QtB = hrB * utB;
  const float FtA = QtA * vnA;
  const float FtB = QtB * vnB;
  const float dQt = QtB - QtA;
  // sR >= 0 >= sL.  Where both wave speeds vanish (spread <= 1e-12: both
  // sides dry to within (1e-12)^2 / g) the reference takes the quotient as 1;
  // clamping the spread at 1e-12 leaves such a face's flux as negligible
  // (every term carries a factor hr <= 1e-25) with one instruction fewer.
  const float spread = sR - sL;
  const float r = __builtin_amdgcn_rcpf(fmaxf(spread, 1.0e-12f));
  const float ka = sR * r;          // weight of the A-side flux
  const float kb = -sL * r;         // weight of the B-side flux
  const float kc = ka * sL;         // weight of the state jump, sL sR / (sR - sL)
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

// The x face west of this lane's cell when both of its cells are active:
// side A is lane - 1's cell, side B this lane's.  The same operations as
// face_fast<true, false> with A = from_west of each value, so the same bits;
// here each of A's values is fetched at its use, so that the compiler folds
// the cross-lane move into the instruction that uses it (a DPP operand of a
// subtraction or product) instead of issuing it as a move of its own.
__device__ __forceinline__ FaceFlux face_x_west(const Grid& G, const SideF& B) {
  const float dry = G.dry;
  const float bf = max_nc(from_west(B.b), B.b);
  const float hrA = fmaxf(from_west(B.eta) - bf, 0.0f);
  const float hrB = fmaxf(B.eta - bf, 0.0f);
  const bool wA = hrA > dry, wB = hrB > dry;
  const float unA = from_west(B.u);
  const float vnA = wA ? unA : 0.0f, vnB = wB ? B.u : 0.0f;
  const float QnA = hrA * unA, QnB = hrB * B.u;
  const float cA = __builtin_amdgcn_sqrtf(G.g * hrA);
  const float cB = __builtin_amdgcn_sqrtf(G.g * hrB);
  const float sL = fminf(fminf(vnA - cA, vnB - cB), 0.0f);
  const float sR = fmaxf(fmaxf(vnA + cA, vnB + cB), 0.0f);
  const float prA = (G.half_g * hrA) * hrA, prB = (G.half_g * hrB) * hrB;
  const float FnA = fmaf(QnA, vnA, prA);
  const float FnB = fmaf(QnB, vnB, prB);
  const float QtA = from_west(B.v) * hrA, QtB = hrB * B.v;  // hr ut (x face: ut = v)
  const float FtA = QtA * vnA;
  const float FtB = QtB * vnB;
  const float dQt = QtB - QtA;
  const float spread = sR - sL;
  const float r = __builtin_amdgcn_rcpf(fmaxf(spread, 1.0e-12f));
  const float ka = sR * r;
  const float kb = -sL * r;
  const float kc = ka * sL;
  FaceFlux F;
  const float mass = fmaf(ka, QnA, kb * QnB);
  F.f0 = fmaf(kc, hrB - hrA, mass);
  F.fn = fmaf(kc, QnB - QnA, fmaf(ka, FnA, kb * FnB));
  F.ft = fmaf(kc, dQt, fmaf(ka, FtA, kb * FtB));
  F.cA = from_west(B.p) - prA;
  F.cB = B.p - prB;
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
  unsigned c;   // kStaticCls: the cell's class, until settled
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
// Static layouts of the packed loop (LAY):
//   kStatic12  (bed, n', multiplier), 12 bytes per cell;
//   kStatic8   (bed, n'), 8 bytes, when the rain multiplier has one value on
//              every active cell (the reference's default multiplier, the
//              domain mask, and any scalar multiple of it), taken as a
//              constant: the rain source is the same fused multiply-add with
//              the same operands, so the bits do not change;
//   kStaticCls bed (4 bytes) plus a one-byte class per cell, when moreover the
//              Manning field takes at most 255 values on the active cells (a
//              land-use roughness map): class c < 255 stands for n' = n^2 of
//              that value, read from a table in LDS, and class 255 for an
//              inactive cell (n' = -1).  The table holds the same float32
//              products as the 8-byte record and the active flag is again
//              the sign of n', so again the bits do not change.  Stored in
//              blocks of 64 columns, 64 beds then 64 classes (320 bytes), so
//              the static fields stay one stream; a row is
//              ceil(nx / 64) blocks.
enum { kStatic12 = 0, kStatic8 = 1, kStaticCls = 2 };
constexpr unsigned kClsInactive = 255u;
constexpr unsigned kClsBlockBytes = 320u;  // 64 beds + 64 classes

__host__ __device__ __forceinline__ size_t cls_row_bytes(int nx) {
  return (size_t)((nx + 63) / 64) * kClsBlockBytes;
}

// The class table of kStaticCls of the single-domain step, one per
// workgroup: dtg n^2 of each class (filled by the step kernel once the step's
// dt is known, before the march).  The march reads whatever table it is given
// (PackedIO::tab), so a kernel whose wavefronts step different domains gives
// each wavefront its own.
__device__ __forceinline__ float* class_table() {
  __shared__ float t[256];
  return t;
}

// Per-strip addressing of the packed layout through raw buffer resources, as
// StripIO: each array gets a resource based at the strip's lowest row, a row
// is a wave-uniform byte offset (scalar register) and each lane adds its
// fixed column offset, so no vector instruction runs per row for addressing.
struct PackedIO {
  Rsrc st, sx, out;  // packed state in, packed static fields, packed state out
  unsigned off12;    // this lane's column, in bytes, in a packed state row
  unsigned offsx;    // this lane's column, in bytes, in a packed static row
  unsigned offcls;   // kStaticCls: this lane's class byte in a packed static row
  unsigned row_st;   // bytes per packed state row (12 nx)
  unsigned row_sx;   // bytes per packed static row (12 nx, 8 nx, or 320 ceil(nx / 64))
  int jlo;           // row the resources start at
  const float* tab;  // kStaticCls: the class table in LDS (dtg n^2, or -1 for class 255)
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

// Cache policy of the state's loads and stores: non-temporal (it streams
// through once per step; measured 1-2 % faster at 4096^2 than plain ones).
// The static fields keep plain loads.
constexpr int kStreamNT = 2;

// The multi-dword buffer loads return vectors; their elements are read as
// .x/.y/.z of these types (a bit_cast of an indexed element of the builtin's
// own return type has compiled to element 0 for every index).
typedef unsigned U32x2 __attribute__((ext_vector_type(2)));
typedef unsigned U32x3 __attribute__((ext_vector_type(3)));

template <int LAY>
__device__ __forceinline__ PackedIO packed_io(const Grid& G, const float* S, float* O, int ic,
                                              int jlo, const float* tab) {
  PackedIO io;
  io.jlo = jlo;
  io.tab = tab;
  io.row_st = 12u * (unsigned)G.nx;
  const size_t rows = (size_t)(G.ny - jlo);
  const size_t st0 = (size_t)jlo * io.row_st;
  io.st = make_rsrc(reinterpret_cast<const char*>(uniform_ptr(S)) + st0, rows * io.row_st);
  io.out = make_rsrc(reinterpret_cast<char*>(uniform_ptr(O)) + st0, rows * io.row_st);
  io.off12 = (unsigned)ic * 12u;
  if constexpr (LAY == kStaticCls) {
    const unsigned blk = (unsigned)ic / 64u * kClsBlockBytes, i = (unsigned)ic % 64u;
    io.offsx = blk + 4u * i;
    io.offcls = blk + 256u + i;
    io.row_sx = (unsigned)cls_row_bytes(G.nx);
  } else {
    constexpr unsigned per = LAY == kStatic12 ? 12u : 8u;
    io.offsx = (unsigned)ic * per;
    io.offcls = 0;
    io.row_sx = per * (unsigned)G.nx;
  }
  io.sx = make_rsrc(reinterpret_cast<const char*>(uniform_ptr(G.pk)) + (size_t)jlo * io.row_sx,
                    rows * io.row_sx);
  return io;
}

// Row j of the lane's column: clamped row index (numerics.py:29-55); the
// reflective-edge sign flips are applied in side_of.  mc: the uniform
// multiplier (kStatic8 and kStaticCls).
template <int LAY, bool EDGE = true>
__device__ __forceinline__ RawCell load_raw(const Grid& G, const PackedIO& io, int j, float mc) {
  const int jj = min(max(j, 0), G.ny - 1);
  const unsigned rs = (unsigned)(jj - io.jlo);
  RawCell r;
  const U32x3 st = __builtin_amdgcn_raw_buffer_load_b96(io.st, io.off12, rs * io.row_st, kStreamNT);
  r.h = __uint_as_float(st.x);
  r.hu = __uint_as_float(st.y);
  r.hv = __uint_as_float(st.z);
  const unsigned so = rs * io.row_sx;
  if constexpr (LAY == kStatic12) {
    const U32x3 sx = __builtin_amdgcn_raw_buffer_load_b96(io.sx, io.offsx, so, 0);
    r.b = __uint_as_float(sx.x);
    r.n = __uint_as_float(sx.y);
    r.m = __uint_as_float(sx.z);
  } else if constexpr (LAY == kStatic8) {
    const U32x2 sx = __builtin_amdgcn_raw_buffer_load_b64(io.sx, io.offsx, so, 0);
    r.b = __uint_as_float(sx.x);
    r.n = __uint_as_float(sx.y);
    r.m = mc;
  } else {
    r.b = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(io.sx, io.offsx, so, 0));
    // the class is looked up when the row is used (settled), not here
    r.c = __builtin_amdgcn_raw_buffer_load_b8(io.sx, io.offcls, so, 0);
    r.m = mc;
  }
  // r.a is set where the row is first used (settled): taken here, at the
  // load, its comparison made the compiler wait for this newest load at once.
  r.neg_hv = EDGE && ((j < 0 && G.bt[2] == 0) || (j >= G.ny && G.bt[3] == 0));
  return r;
}

// The row's class looked up: n' and the active flag as the 8-byte record
// gives them.  Called where the row is first used, so the loads of the rows
// behind it stay in flight.
template <int LAY, class IO>
__device__ __forceinline__ RawCell settled(RawCell r, const IO& io) {
  if constexpr (LAY == kStaticCls) r.n = io.tab[r.c];
  if constexpr (std::is_same_v<IO, PackedIO>) r.a = !signbit(r.n);  // sign bit clear: active
  return r;
}

template <int LAY, bool EDGE = true>
__device__ __forceinline__ RawCell load_raw(const Grid& G, const StripIO& io, int j, float) {
  const int jj = min(max(j, 0), G.ny - 1);
  RawCell r;
  r.h = io.ld(io.h, jj);
  r.hu = io.ld(io.hu, jj);
  r.hv = io.ld(io.hv, jj);
  r.b = io.ld(io.bed, jj);
  r.a = io.ld_mask(jj) != 0;
  r.neg_hv = EDGE && ((j < 0 && G.bt[2] == 0) || (j >= G.ny && G.bt[3] == 0));
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

// Rows whose loads are in flight ahead of the current row of a march.  At
// 4096^2: 2 rows (78 registers, 6 wavefronts per SIMD) ran 1-2 % slower than
// 3 (85 registers, 5 per SIMD); 4 and 5 no faster than 3.
constexpr int kRowsAhead = 3;

// f(integral_constant<int, U>) for U = B .. E-1, in order: indices derived
// from U are compile-time constants in the source, so arrays indexed by them
// stay in registers.
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// One wavefront's strip: columns sx*62-1 .. sx*62+62, rows sy*rows ..
// sy*rows+rows-1.  sx, sy and rows must be wave-uniform.  EDGE: the strip may
// touch the grid's edges (ghost columns or rows, open-boundary faces); a
// strip that does not can take the march without them.  The raw inputs of
// row j+2 are in flight while row j is computed.  PACKED: H and Ho are packed
// states (3 floats per cell, HU/HV unused) and the static fields come from
// G.pk; the arithmetic is the same, so both layouts give identical bits.
template <bool LOOP, bool PACKED, bool SOUTH, int LAY, bool EDGE, bool VOL, bool SR>
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
                                                 float mc, const float* tab, int sr_step) {
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  const int lane = threadIdx.x & (kWave - 1);
  const int i = sx * kStripCols - 1 + lane;
  const bool own = lane >= 1 && lane <= kStripCols && i < G.nx;
  const int ic = min(max(i, 0), G.nx - 1);
  const bool neg_hu = EDGE && ((i < 0 && G.bt[0] == 0) || (i >= G.nx && G.bt[1] == 0));
  // Only the first and last strip of a row band touch the west/east edge faces.
  const int edge_x = EDGE && LOOP && (sx == 0 || (sx + 1) * kStripCols >= G.nx);
  const int j0 = sy * rows;
  const int j1 = min(j0 + rows, G.ny);
  const float dry = G.dry;
  // dry in a vector register too: max(hs, dry) below is a med3 whose other
  // bound (+inf) is a scalar, and a med3 takes one scalar operand at most.
  float dry_v = dry;
  asm("" : "+v"(dry_v));
  // With the uniform multiplier (a scalar) the rain term's two factors would
  // both be scalars, which one vector instruction cannot read: dt * rate is
  // kept in a vector register for the march, not copied into one per row.
  float dtrate_v = dtrate;
  if constexpr (LAY != kStatic12) asm("" : "+v"(dtrate_v));
  // SR: the rain is added with stochastic rounding (sr_add), drawn for
  // (cell, step sr_step).
  const unsigned sr_seed_s = SR ? __builtin_amdgcn_readfirstlane(sr_seed(sr_step)) : 0u;
  using IO = std::conditional_t<PACKED, PackedIO, StripIO>;
  IO io;
  if constexpr (PACKED)
    io = packed_io<LAY>(G, H, Ho, ic, max(j0 - 1, 0), tab);
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
  // Raw rows rotate through kRowsAhead + 1 register sets (raw[]): a row is
  // loaded kRowsAhead rows before it becomes the current row.  Prologue: the
  // face behind row jf, row jf (current) and rows jf+dj .. jf+(kRowsAhead-1)dj
  // (in flight, raw[0 ..]).
  constexpr int kSets = kRowsAhead + 1;
  RawCell raw[kSets];
  RowCell cur = side_of(settled<LAY>(load_raw<LAY, EDGE>(G, io, jf - dj, mc), io), neg_hu, dry, G.half_g);
  float yc0, yc1, yc2;  // carried y face: between the current row and the one behind it
  float nn, mult;
  {
    const RawCell first = settled<LAY>(load_raw<LAY, EDGE>(G, io, jf, mc), io);
    const RowCell c0 = side_of(first, neg_hu, dry, G.half_g);
    if constexpr (PACKED) {
      nn = first.n;
      mult = first.m;
    }
    raw[0] = load_raw<LAY, EDGE>(G, io, jf + dj, mc);
    static_for<2, kRowsAhead>([&](auto mc_) __attribute__((always_inline)) {
      constexpr int m = decltype(mc_)::value;
      raw[m - 1] = load_raw<LAY, EDGE>(G, io, SOUTH ? max(jf - m, j0 - 1) : min(jf + m, j1), mc);
    });
    if constexpr (!SOUTH) {  // south face of row j0
      const FaceFlux fy = face_fast_any<false>(G, cur.s, c0.s, wave_all(cur.s.a && c0.s.a));
      if (EDGE && LOOP && j0 == 0 && own) acc.bnd += (double)fy.f0 * dxd;
      yc0 = fy.f0;
      yc1 = fy.ft;
      yc2 = fy.fn + fy.cB;
    } else {  // north face of row j1-1
      const FaceFlux fy = face_fast_any<false>(G, c0.s, cur.s, wave_all(cur.s.a && c0.s.a));
      if (EDGE && LOOP && j1 == G.ny && own) acc.bnd -= (double)fy.f0 * dxd;
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
  // Lane maximum of the CFL terms below, taken on their bit patterns: the
  // terms are +0 or positive, which order as their patterns do, and a
  // non-finite new h, hu or hv makes its term infinite or NaN, whose pattern
  // (either sign) is at least +inf's, so the same maximum flags it.
  unsigned inv_bits = 0u;
  float clip32 = 0.0f;  // clip depth is small; float32 per lane over the strip

  // Row j: `cur` is row j, `north` row j+dj (the next row of the march,
  // loaded kRowsAhead - 1 rows ago; the rows up to j+(kRowsAhead-1)dj are in
  // flight too).  Issues the loads of row j+kRowsAhead dj into `ahead`, takes row j+dj's n/multiplier into
  // nn_next/mult_next, and leaves row j+dj's side in `up`.
  auto row = [&](int j, const RowCell& cur, RowCell& up, const RawCell& north, RawCell& ahead,
                 float nn_j, float mult_j, float& nn_next, float& mult_next) {
    // Unconditional, so the wait counts stay exact; past the strip the
    // clamped rows were just loaded and come from the cache.
    ahead = load_raw<LAY, EDGE>(G, io, SOUTH ? max(j - kRowsAhead, j0 - 1) : min(j + kRowsAhead, j1), mc);
    const RawCell nrow = settled<LAY>(north, io);
    if constexpr (PACKED) {  // row j+dj's fields arrived with its state
      nn_next = nrow.n;
      mult_next = nrow.m;
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
    const unsigned long long am = __builtin_amdgcn_ballot_w64(ca);
    const unsigned long long need = __builtin_amdgcn_read_exec() & ~1ull;
    FaceFlux fx;
    if (((am << 1) & am & need) == need) {
      fx = face_x_west(G, cur.s);
    } else {
      const bool wa = from_west((int)ca) != 0;
      fx = face_fast<true, true>(G, from_west(cur.s.h), from_west(cur.s.eta), from_west(cur.s.b),
                                 from_west(cur.s.u), from_west(cur.s.v), from_west(cur.s.p), wa,
                                 cur.s.h, cur.s.eta, cur.s.b, cur.s.u, cur.s.v, cur.s.p, ca);
    }
    const float xr0 = fx.f0, xr1 = fx.fn + fx.cB, xr2 = fx.ft;
    const float xe1 = fx.fn + fx.cA;  // this lane's share of the face west of lane+1
    // edge_x and the record range are uniform integers tested in every row
    // through an empty asm statement (scalar register, no instruction): a
    // loop-invariant test would be hoisted out of the march as a 64-bit lane
    // mask, and with scalar registers short, spilled and re-expanded per row.
    int ex = __builtin_amdgcn_readfirstlane(edge_x);
    asm volatile("" : "+s"(ex));
    if (ex) {
      if (i == 0) acc.bnd += (double)fx.f0 * dyd;
      if (i == G.nx) acc.bnd -= (double)fx.f0 * dyd;
    }
    // The y face ahead of row j; a y face's A side is always its south cell.
    up = side_of(nrow, neg_hu, dry, G.half_g);
    const SideF& ys = SOUTH ? up.s : cur.s;  // south cell of the face
    const SideF& yn = SOUTH ? cur.s : up.s;  // north cell
    const bool as = SOUTH ? up.s.a : ca, an = SOUTH ? ca : up.s.a;
    FaceFlux gy;
    bool all_y;  // both cells of every lane's y face are active
    if constexpr (PACKED)
      all_y = (active_lanes(nn_j) & active_lanes(nrow.n)) == __builtin_amdgcn_read_exec();
    else
      all_y = wave_all(ca && up.s.a);
    if (all_y)
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

    // The east faces from lane+1, each fetched at its one use: a
    // subtraction that takes the cross-lane move as its (DPP) operand.
    // dt ((xl - xr)/dx + (yn - ys)/dy) as (dt/dx) ((xl - xr) + (yn - ys) dx/dy)
    const float d0 = fmaf(yn0 - ys0, G.rhox, east_minus(fx.f0, xr0));
    const float d1 = fmaf(yn1 - ys1, G.rhox, east_minus(xe1, xr1));
    const float d2 = fmaf(yn2 - ys2, G.rhox, east_minus(fx.ft, xr2));
    const float hpre = fmaf(-dtx, d0, cur.s.h);
    float hs;
    if constexpr (SR) {
      const unsigned cell = (unsigned)j * (unsigned)G.nx + (unsigned)ic;
      hs = sr_add(hpre, dtrate_v * mult_j, sr_hash(cell ^ sr_seed_s));
    } else {
      hs = fmaf(dtrate_v, mult_j, hpre);
    }
    const float hus = fmaf(-dtx, d1, cur.hu);
    const float hvs = fmaf(-dtx, d2, cur.hv);
    // (sr_add's result is a selection of bit patterns, which the compiler
    // would quiet before a plain max: max_nc takes it as it is)
    const float hn = SR ? max_nc(hs, 0.0f) : fmaxf(hs, 0.0f);
    const float hd = max_nc(hs, dry_v);  // = max(hn, dry): dry > 0
    const float rhd = __builtin_amdgcn_rcpf(hd);
    const float q = __builtin_amdgcn_sqrtf(fmaf(hus, hus, hvs * hvs));
    // K = dtg n^2 q / hd^(7/3), the power as 2^(-7/3 log2 hd) on the hardware log/exp
    const float ip73 = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(hd) * -kPow73);
    const float n2 = PACKED ? nn_j : nn_j * nn_j;  // the packed field holds n^2
    // the class table holds dtg n^2 already
    const float dn2 = LAY == kStaticCls ? n2 : dtg * n2;
    const float K = (dn2 * q) * ip73;
    const float rD = __builtin_amdgcn_rcpf(1.0f + K);
    const bool wet = ca && (hn > dry);
    const float h_o = ca ? hn : 0.0f;
    const float hu_o = wet ? hus * rD : 0.0f;
    const float hv_o = wet ? hvs * rD : 0.0f;
    if (own) {
      if constexpr (PACKED) {
        const U32x3 v = {__float_as_uint(h_o), __float_as_uint(hu_o), __float_as_uint(hv_o)};
        __builtin_amdgcn_raw_buffer_store_b96(v, io.out, io.off12,
                                              (unsigned)(j - io.jlo) * io.row_st, kStreamNT);
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
        int lo = __builtin_amdgcn_readfirstlane(rec_lo);
        int hi = __builtin_amdgcn_readfirstlane(rec_hi);
        asm volatile("" : "+s"(lo), "+s"(hi));
        if (__builtin_expect(lo < hi, 0)) {
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
        // An inactive cell gives exactly +0 here (h_o, hu_o and hv_o are 0
        // and rhd is finite: hd >= dry > 0), as the lane's maximum starts from.
        inv_bits = max(inv_bits, __float_as_uint(fmaf(su, G.rho, sv)));
        if constexpr (VOL) acc.vol += (double)h_o;
        clip32 += ca ? hn - hs : 0.0f;  // max(-hs, 0), exactly
      }
    }
    yc0 = gy.f0;
    yc1 = gy.ft;
    yc2 = gy.fn + (SOUTH ? gy.cA : gy.cB);
  };

  // kSets rows per trip, so every raw set keeps its register slot: a row's
  // raw inputs stay live over kSets rows (loaded, read as the next row, then
  // as the current row), and fewer sets would make the compiler copy one at
  // every loop edge, and wait there for the loads in flight.  Row j+u of a
  // trip reads row j+u+1 from raw[u % kSets] and loads row j+u+kRowsAhead into
  // raw[(u + kRowsAhead - 1) % kSets].  The derived row values alternate
  // between two sets and are handed back after an odd-length trip; they are
  // computed values, so that copy waits for nothing.  All indices are
  // compile-time constants after unrolling, so raw[] stays in registers.
  RowCell rc[2] = {cur, cur};
  float nnv[2] = {nn, 0.0f}, multv[2] = {mult, 0.0f};
  const int n = j1 - j0;
  int k = 0;
  for (; k + kSets - 1 < n; k += kSets) {
    const int j = jf + k * dj;
    static_for<0, kSets>([&](auto uc) __attribute__((always_inline)) {
      constexpr int u = decltype(uc)::value;
      row(j + u * dj, rc[u & 1], rc[(u + 1) & 1], raw[u % kSets], raw[(u + kRowsAhead - 1) % kSets],
          nnv[u & 1], multv[u & 1], nnv[(u + 1) & 1], multv[(u + 1) & 1]);
    });
    if constexpr (kSets & 1) {
      rc[0] = rc[1];
      nnv[0] = nnv[1];
      multv[0] = multv[1];
    }
  }
  static_for<0, kSets - 1>([&](auto uc) __attribute__((always_inline)) {
    constexpr int u = decltype(uc)::value;
    if (k + u < n)
      row(jf + (k + u) * dj, rc[u & 1], rc[(u + 1) & 1], raw[u % kSets],
          raw[(u + kRowsAhead - 1) % kSets], nnv[u & 1], multv[u & 1], nnv[(u + 1) & 1],
          multv[(u + 1) & 1]);
  });
  // yc0 now holds the mass flux through the last face of the march: the north
  // face of row j1-1, or (SOUTH) the south face of row j0.
  if constexpr (!SOUTH) {
    if (EDGE && LOOP && j1 == G.ny && own) acc.bnd -= (double)yc0 * dxd;
  } else {
    if (EDGE && LOOP && j0 == 0 && own) acc.bnd += (double)yc0 * dxd;
  }
  if (LOOP) {
    const bool nonfinite = inv_bits >= 0x7F800000u;  // +inf, NaN or a negative pattern
    acc.nf |= nonfinite;
    acc.clip += (double)clip32;
    // max(x) / dy = max(x / dy): rounding is monotonic
    acc.inv = nonfinite ? 0.0f : __uint_as_float(inv_bits) * G.rdy;
  }
}

// One wavefront's strip (see march_strip_dir).  Strips of odd row bands march
// south and the others north, so two strips that share halo rows (the last
// two rows of one band and the first two of the next) read them near the
// same time -- both at the start or both at the end of their marches -- and
// the second read is an L2 hit.  The direction does not change any cell's
// arithmetic: every face takes the same operands, and a cell combines its
// faces in the same order (north minus south, east minus west).
template <bool LOOP, bool PACKED = false, int LAY = kStatic12, bool VOL = true>
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
                                                 float mc = 0.0f, const float* tab = nullptr,
                                                 int sr_step = -1) {
  static_assert(PACKED || LAY == kStatic12, "the compact static records are packed layouts");
  using S = std::true_type;
  using N = std::false_type;
  // A device loop's step with stochastic rounding of its rain (sr_step >= 0)
  // takes its own instantiation of the march, so the other steps' code does
  // not change.
  auto march = [&](auto south, auto edge) __attribute__((always_inline)) {
    if constexpr (LOOP) {
      if (sr_step >= 0)
        march_strip_dir<LOOP, PACKED, decltype(south)::value, LAY, decltype(edge)::value, VOL,
                        true>(G, H, HU, HV, Ho, HUo, HVo, sx, sy, rows, dtx, dtrate, dtg, dxd,
                              dyd, rec_lo, rec_hi, cube_h, cube_hu, cube_hv, acc, mc, tab,
                              sr_step);
      else
        march_strip_dir<LOOP, PACKED, decltype(south)::value, LAY, decltype(edge)::value, VOL,
                        false>(G, H, HU, HV, Ho, HUo, HVo, sx, sy, rows, dtx, dtrate, dtg, dxd,
                               dyd, rec_lo, rec_hi, cube_h, cube_hu, cube_hv, acc, mc, tab, -1);
    } else {
      march_strip_dir<LOOP, PACKED, decltype(south)::value, LAY, decltype(edge)::value, VOL,
                      false>(G, H, HU, HV, Ho, HUo, HVo, sx, sy, rows, dtx, dtrate, dtg, dxd, dyd,
                             rec_lo, rec_hi, cube_h, cube_hu, cube_hv, acc, mc, tab, -1);
    }
  };
  if constexpr (PACKED) {
    // A strip whose 64 columns and two halo rows all lie inside the grid
    // (every strip but the first and last of a row band, in every band but
    // the southmost and northmost) has no reflective-edge sign flips and no
    // open-boundary faces: it takes the march without that work.
    const int j0 = sy * rows, j1 = min(j0 + rows, G.ny);
    const bool interior = sx >= 1 && (sx + 1) * kStripCols < G.nx && j0 >= 1 && j1 < G.ny;
    if (sy & 1) {
      if (interior) march(S{}, N{});
      else march(S{}, S{});
    } else {
      if (interior) march(N{}, N{});
      else march(N{}, S{});
    }
  } else {
    if (sy & 1) march(S{}, S{});
    else march(N{}, S{});
  }
}

}  // namespace swe
