// The per-wavefront strip march of the fused step, shared by the single-domain
// step kernel and the ensemble kernel.
//
// A wavefront owns columns [62*sx - 1, 62*sx + 63) of one grid (lanes 1..62
// update their column, lanes 0 and 63 are the halo) and rows
// [rows*sy, rows*sy + rows).  It marches north one row at a time: the x face
// on each lane's west side is evaluated once (the west cell arrives by a DPP
// wave shift, the face's west-side output leaves by the opposite shift), the
// y face on the north side is evaluated once and carried to the next row.
#pragma once

#include "swe_device.hpp"

namespace swe {

// Per-lane accumulators of the loop epilogue.
struct StripAcc {
  float inv = 0.0f;   // max inverse CFL time of the new state
  double vol = 0.0;   // sum of new depths
  double clip = 0.0;  // depth created by max(h, 0)
  double bnd = 0.0;   // signed inflow (flux * face length) through the outer faces
  int nf = 0;         // non-finite seen
};

// Row loader for one lane's column with the ghost rules of
// boundary_conditions.py:36-51 (rows/columns outside the grid copy the edge
// and negate the normal momentum on reflective sides).
template <class M>
struct ColumnReader {
  const float* __restrict__ H;
  const float* __restrict__ HU;
  const float* __restrict__ HV;
  int ic;       // clamped column
  bool neg_hu;  // this lane is a reflective ghost column
  int nx, ny;
  bool refl_south, refl_north;

  __device__ __forceinline__ Cell load(const Grid& G, int j) const {
    const int jj = min(max(j, 0), ny - 1);
    const size_t idx = (size_t)jj * (size_t)nx + (size_t)ic;
    Cell c;
    c.h = H[idx];
    c.hu = HU[idx];
    c.hv = HV[idx];
    c.b = G.bed[idx];
    c.a = G.mask[idx] != 0;
    if (neg_hu) c.hu = -c.hu;
    if ((j < 0 && refl_south) || (j >= ny && refl_north)) c.hv = -c.hv;
    M::velocities(c.h, c.hu, c.hv, G.dry, c.u, c.v);
    return c;
  }
};

template <class M, bool LOOP>
__device__ __forceinline__ void march_strip(const Grid& G, const float* __restrict__ H,
                                            const float* __restrict__ HU,
                                            const float* __restrict__ HV, float* __restrict__ Ho,
                                            float* __restrict__ HUo, float* __restrict__ HVo,
                                            int sx, int sy, int rows, float dt, float dtrate,
                                            float dtg, double dxd, double dyd, int rec_lo,
                                            int rec_hi, float* __restrict__ cube_h,
                                            float* __restrict__ cube_hu,
                                            float* __restrict__ cube_hv, StripAcc& acc,
                                            int sr_step = -1) {
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  const int lane = threadIdx.x & (kWave - 1);
  // sr_step >= 0: the rain is added with stochastic rounding, drawn for
  // (cell, step) -- the same draw as the fast march and the oracle.
  const bool sr = sr_step >= 0;
  const unsigned seed = sr ? sr_seed(sr_step) : 0u;
  const int i = sx * kStripCols - 1 + lane;
  const bool own = lane >= 1 && lane <= kStripCols && i < G.nx;
  ColumnReader<M> rd;
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

  const Cell below = rd.load(G, j0 - 1);
  Cell cur = rd.load(G, j0);
  FaceFlux fy;
  if (__all(below.a && cur.a))
    fy = face_flux<M, false, false>(G, below.h, below.v, below.u, below.b, below.a, cur.h, cur.v,
                                    cur.u, cur.b, cur.a);
  else
    fy = face_flux<M, false, true>(G, below.h, below.v, below.u, below.b, below.a, cur.h, cur.v,
                                   cur.u, cur.b, cur.a);
  // South face of the first row, seen from its north (top) side.
  float yt0 = fy.f0, yt1 = fy.ft, yt2 = fy.fn + fy.cB;
  if (LOOP && j0 == 0 && own) acc.bnd += (double)fy.f0 * dxd;

  for (int j = j0; j < j1; ++j) {
    const Cell up = rd.load(G, j + 1);
    const size_t idx = (size_t)j * (size_t)G.nx + (size_t)rd.ic;
    const float nn = G.n[idx];
    const float mult = G.mult[idx];

    // West face of this lane's column; the lane to the west supplies its cell.
    const float wh = from_west(cur.h), wu = from_west(cur.u), wv = from_west(cur.v);
    const float wb = from_west(cur.b);
    const int wa = from_west(cur.a);
    FaceFlux fx;
    if (__all(lane == 0 || (wa && cur.a)))
      fx = face_flux<M, true, false>(G, wh, wu, wv, wb, wa, cur.h, cur.u, cur.v, cur.b, cur.a);
    else
      fx = face_flux<M, true, true>(G, wh, wu, wv, wb, wa, cur.h, cur.u, cur.v, cur.b, cur.a);
    // This cell is the east (right) side of its west face ...
    const float xr0 = fx.f0, xr1 = fx.fn + fx.cB, xr2 = fx.ft;
    // ... and the west (left) side of its east face, evaluated by lane+1.
    const float xl0 = from_east(fx.f0);
    const float xl1 = from_east(fx.fn + fx.cA);
    const float xl2 = from_east(fx.ft);
    if (LOOP) {
      if (i == 0) acc.bnd += (double)fx.f0 * dyd;
      if (i == G.nx) acc.bnd -= (double)fx.f0 * dyd;
    }

    // North face: this cell is its south (bottom) side.
    FaceFlux gy;
    if (__all(cur.a && up.a))
      gy = face_flux<M, false, false>(G, cur.h, cur.v, cur.u, cur.b, cur.a, up.h, up.v, up.u,
                                      up.b, up.a);
    else
      gy = face_flux<M, false, true>(G, cur.h, cur.v, cur.u, cur.b, cur.a, up.h, up.v, up.u,
                                     up.b, up.a);
    const float yb0 = gy.f0, yb1 = gy.ft, yb2 = gy.fn + gy.cA;
    if (LOOP && j + 1 == G.ny && own) acc.bnd -= (double)gy.f0 * dxd;

    const float d0 = M::per_dx(xl0 - xr0, G.dx, G.rdx) + M::per_dx(yb0 - yt0, G.dy, G.rdy);
    const float d1 = M::per_dx(xl1 - xr1, G.dx, G.rdx) + M::per_dx(yb1 - yt1, G.dy, G.rdy);
    const float d2 = M::per_dx(xl2 - xr2, G.dx, G.rdx) + M::per_dx(yb2 - yt2, G.dy, G.rdy);
    const CellUpdate r = cell_update<M>(G, cur.h, cur.hu, cur.hv, d0, d1, d2, nn, mult, cur.a,
                                        dt, dtrate, dtg, sr, sr_hash((unsigned)idx ^ seed));
    if (own) {
      Ho[idx] = r.h;
      HUo[idx] = r.hu;
      HVo[idx] = r.hv;
      if (LOOP) {
        for (int k = rec_lo; k < rec_hi; ++k) {
          cube_h[(size_t)k * N + idx] = r.h;
          cube_hu[(size_t)k * N + idx] = r.hu;
          cube_hv[(size_t)k * N + idx] = r.hv;
        }
        acc.inv = fmaxf(acc.inv, cfl_inverse<M>(G, r.h, r.hu, r.hv, cur.a));
        acc.vol += (double)r.h;
        if (cur.a) acc.clip += (double)fmaxf(-r.h_star, 0.0f);
        acc.nf |= !(isfinite(r.h) && isfinite(r.hu) && isfinite(r.hv));
      }
    }
    yt0 = gy.f0;
    yt1 = gy.ft;
    yt2 = gy.fn + gy.cB;
    cur = up;
  }
}

}  // namespace swe
