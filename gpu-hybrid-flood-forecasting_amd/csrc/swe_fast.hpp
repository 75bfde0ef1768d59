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
//   - one reciprocal of max(h, dry) shared by the friction law and the CFL
//     epilogue; x^(7/3) as x^2 * 2^(log2(x)/3 * ...) on the hardware log/exp;
//   - depth and clip volume summed per lane in float32 over the strip's rows,
//     then in float64 across lanes and workgroups.
// Results stay within the stated tolerance of the FP64 oracle; the lake at
// rest, wall mass flux and open-boundary accounting keep their exact forms.
#pragma once

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

struct RowCell {
  SideF s;
  float hu, hv;
};

template <class M>
__device__ __forceinline__ RowCell load_fast(const ColumnReader<M>& rd, const Grid& G, int j) {
  const int jj = min(max(j, 0), rd.ny - 1);
  const size_t idx = (size_t)jj * (size_t)rd.nx + (size_t)rd.ic;
  float h = rd.H[idx], hu = rd.HU[idx], hv = rd.HV[idx];
  const float b = G.bed[idx];
  const int a = G.mask[idx] != 0;
  if (rd.neg_hu) hu = -hu;
  if ((j < 0 && rd.refl_south) || (j >= rd.ny && rd.refl_north)) hv = -hv;
  RowCell c;
  c.s = make_side(h, hu, hv, b, a, G.dry);
  c.hu = hu;
  c.hv = hv;
  return c;
}

template <bool LOOP>
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
  ColumnReader<FastMath> rd;
  rd.H = H;
  rd.HU = HU;
  rd.HV = HV;
  rd.ic = min(max(i, 0), G.nx - 1);
  rd.neg_hu = (i < 0 && G.bt[0] == 0) || (i >= G.nx && G.bt[1] == 0);
  rd.nx = G.nx;
  rd.ny = G.ny;
  rd.refl_south = G.bt[2] == 0;
  rd.refl_north = G.bt[3] == 0;
  const int j0 = sy * rows;
  const int j1 = min(j0 + rows, G.ny);
  const float rdx = G.rdx, rdy = G.rdy, dry = G.dry;

  const RowCell below = load_fast(rd, G, j0 - 1);
  RowCell cur = load_fast(rd, G, j0);
  FaceFlux fy = face_fast_any<false>(G, below.s, cur.s, __all(below.s.a && cur.s.a));
  float yt0 = fy.f0, yt1 = fy.ft, yt2 = fy.fn + fy.cB;
  if (LOOP && j0 == 0 && own) acc.bnd += (double)fy.f0 * dxd;
  float vol32 = 0.0f, clip32 = 0.0f;

  for (int j = j0; j < j1; ++j) {
    const RowCell up = load_fast(rd, G, j + 1);
    const size_t idx = (size_t)j * (size_t)G.nx + (size_t)rd.ic;
    const float nn = G.n[idx];
    const float mult = G.mult[idx];

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
    if (LOOP) {
      if (i == 0) acc.bnd += (double)fx.f0 * dyd;
      if (i == G.nx) acc.bnd -= (double)fx.f0 * dyd;
    }
    const FaceFlux gy = face_fast_any<false>(G, cur.s, up.s, __all(cur.s.a && up.s.a));
    const float yb0 = gy.f0, yb1 = gy.ft, yb2 = gy.fn + gy.cA;
    if (LOOP && j + 1 == G.ny && own) acc.bnd -= (double)gy.f0 * dxd;

    const float d0 = fmaf(yb0 - yt0, rdy, (xl0 - xr0) * rdx);
    const float d1 = fmaf(yb1 - yt1, rdy, (xl1 - xr1) * rdx);
    const float d2 = fmaf(yb2 - yt2, rdy, (xl2 - xr2) * rdx);
    const float hs = fmaf(dtrate, mult, fmaf(-dt, d0, cur.s.h));
    const float hus = fmaf(-dt, d1, cur.hu);
    const float hvs = fmaf(-dt, d2, cur.hv);
    const float hn = fmaxf(hs, 0.0f);
    const float hd = fmaxf(hn, dry);
    const float rhd = __builtin_amdgcn_rcpf(hd);
    const float q = __builtin_amdgcn_sqrtf(fmaf(hus, hus, hvs * hvs));
    // 1 / hd^(7/3) = (1/hd)^2 * 2^(-log2(hd) * (7/3 - 2)); K = dtg n^2 q / hd^(7/3)
    const float ip13 = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(hd) * -kPow73MinusTwo);
    const float K = ((dtg * (nn * nn)) * q) * ((rhd * rhd) * ip13);
    const float rD = __builtin_amdgcn_rcpf(1.0f + K);
    const bool wet = cur.s.a && (hn > dry);
    const float h_o = cur.s.a ? hn : 0.0f;
    const float hu_o = wet ? hus * rD : 0.0f;
    const float hv_o = wet ? hvs * rD : 0.0f;
    if (own) {
      Ho[idx] = h_o;
      HUo[idx] = hu_o;
      HVo[idx] = hv_o;
      if (LOOP) {
        for (int k = rec_lo; k < rec_hi; ++k) {
          cube_h[(size_t)k * N + idx] = h_o;
          cube_hu[(size_t)k * N + idx] = hu_o;
          cube_hv[(size_t)k * N + idx] = hv_o;
        }
        // CFL of the new state (numerics.py:299-303): u = hu/max(h,dry) where
        // wet, and hu is already 0 where not wet.
        const float ws = __builtin_amdgcn_sqrtf(G.g * h_o);
        const float inv = fmaf(fabsf(hv_o * rhd) + ws, rdy, (fabsf(hu_o * rhd) + ws) * rdx);
        acc.inv = fmaxf(acc.inv, cur.s.a ? inv : 0.0f);
        vol32 += h_o;
        clip32 += cur.s.a ? fmaxf(-hs, 0.0f) : 0.0f;
        acc.nf |= !(isfinite(h_o) && isfinite(hu_o) && isfinite(hv_o));
      }
    }
    yt0 = gy.f0;
    yt1 = gy.ft;
    yt2 = gy.fn + gy.cB;
    cur = up;
  }
  if (LOOP) {
    acc.vol += (double)vol32;
    acc.clip += (double)clip32;
  }
}

}  // namespace swe
