// Adjoint (reverse) of one fused step, by hand: given the state U_n entering
// step n and the cotangent lambda_{n+1} of the state leaving it, produce
// lambda_n = (dU_{n+1}/dU_n)^T lambda_{n+1} and accumulate the cell-wise
// sensitivities to the Manning field and to the rain multiplier.  The step
// size is data (the recorded schedule of the forward run).
//
// One wavefront owns 60 columns (lanes 2..61; two halo lanes each side,
// because a cell's lambda needs both its faces' reverse, and a face's
// reverse needs the pre-source cotangent lambda* of both its cells, which in
// turn needs that cell's two x faces).  It marches north: at row r it
// recomputes the forward of row r, reverses the pointwise part into
// lambda*_r, reverses its west x face (the west cell's share leaves by a DPP
// wave shift) and the y face r-1/2 below it, which completes row r-1.
// Cotangents that land on ghost cells are folded back into the edge cell
// they copy (negated normal momentum on reflective sides).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <utility>

#include "../../include/swe_hip.h"
#include "swe_adjoint.hpp"
#include "swe_device.hpp"
#include "swe_host.hpp"
#include "swe_strip.hpp"

namespace swe {

constexpr int kAdjCols = kWave - 4;  // columns completed per wavefront

// M: arithmetic of the forward recompute (the forward kernel's own, so every
// branch is the one the forward took).  T: arithmetic and storage of the
// cotangents and sensitivities -- float, or double, which gfx950 runs at the
// float rate for add/mul/FMA and which keeps the reverse sweep accurate where
// the cotangent is a small difference of large terms (transcritical cells).
template <class M, class T>
__global__ __launch_bounds__(256) void k_adjoint_step(
    Grid G, const float* __restrict__ H, const float* __restrict__ HU,
    const float* __restrict__ HV, const T* __restrict__ LH, const T* __restrict__ LHU,
    const T* __restrict__ LHV, T* __restrict__ OH, T* __restrict__ OHU, T* __restrict__ OHV,
    T* __restrict__ grad_n, T* __restrict__ grad_mult, float dt, float dtrate, float dtg,
    int rows, int segs_x, int n_waves) {
  // Hardware reciprocal/sqrt/log2/exp2 in the reverse pieces of the fast
  // variant's float sweep (adj::quot and friends).
  constexpr bool kFastAdj = M::kFast && sizeof(T) == sizeof(float);
  const int gw = blockIdx.x * (blockDim.x / kWave) + wave_in_block();
  if (gw >= n_waves) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int sx = gw % segs_x, sy = gw / segs_x;
  const int i = sx * kAdjCols - 2 + lane;
  const bool real = i >= 0 && i < G.nx;
  const bool own = lane >= 2 && lane <= kAdjCols + 1 && i < G.nx;
  const bool refl_w = G.bt[0] == 0, refl_e = G.bt[1] == 0;
  const bool refl_s = G.bt[2] == 0, refl_n = G.bt[3] == 0;
  ColumnReader<M> rd;
  rd.H = H;
  rd.HU = HU;
  rd.HV = HV;
  rd.ic = min(max(i, 0), G.nx - 1);
  rd.neg_hu = (i < 0 && refl_w) || (i >= G.nx && refl_e);
  rd.nx = G.nx;
  rd.ny = G.ny;
  rd.refl_south = refl_s;
  rd.refl_north = refl_n;
  const int j0 = sy * rows;
  const int j1 = min(j0 + rows, G.ny);
  const T cx = T(dt) * T(G.rdx), cy = T(dt) * T(G.rdy);
  const T g = G.g, hg = G.half_g, dry = G.dry;

  Cell Rprev = rd.load(G, j0 - 2);
  Cell Rcur = rd.load(G, j0 - 1);
  FaceFlux Gs = face_flux<M, false>(G, Rprev.h, Rprev.v, Rprev.u, Rprev.b, Rprev.a, Rcur.h, Rcur.v,
                                    Rcur.u, Rcur.b, Rcur.a);
  T lsp0 = T(0), lsp1 = T(0), lsp2 = T(0);  // lambda* of row r-1
  T lup0 = T(0), lup1 = T(0), lup2 = T(0);  // lambda of row r-1 (partial)

  for (int r = j0 - 1; r <= j1; ++r) {
    const bool valid_row = r >= 0 && r < G.ny;
    const Cell Rnext = rd.load(G, r + 1);
    T ls0 = T(0), ls1 = T(0), ls2 = T(0);
    T lu0 = T(0), lu1 = T(0), lu2 = T(0);
    // Forward flux of the face r+1/2; needed as row r+1's south face even when
    // row r itself is the south ghost.
    const FaceFlux Gn = face_flux<M, false>(G, Rcur.h, Rcur.v, Rcur.u, Rcur.b, Rcur.a, Rnext.h,
                                            Rnext.v, Rnext.u, Rnext.b, Rnext.a);
    if (valid_row) {
      const size_t idx = (size_t)r * (size_t)G.nx + (size_t)rd.ic;
      // ---- forward of row r (as the step kernel computes it) ----
      const float wh = from_west(Rcur.h), wu = from_west(Rcur.u), wv = from_west(Rcur.v);
      const float wb = from_west(Rcur.b);
      const float whu = from_west(Rcur.hu), whv = from_west(Rcur.hv);
      const int wa = from_west(Rcur.a);
      const FaceFlux Fx = face_flux<M, true>(G, wh, wu, wv, wb, wa, Rcur.h, Rcur.u, Rcur.v,
                                             Rcur.b, Rcur.a);
      const float xr0 = Fx.f0, xr1 = Fx.fn + Fx.cB, xr2 = Fx.ft;
      const float xl0 = from_east(Fx.f0), xl1 = from_east(Fx.fn + Fx.cA), xl2 = from_east(Fx.ft);
      const float yb0 = Gn.f0, yb1 = Gn.ft, yb2 = Gn.fn + Gn.cA;
      const float yt0 = Gs.f0, yt1 = Gs.ft, yt2 = Gs.fn + Gs.cB;
      const float d0 = M::per_dx(xl0 - xr0, G.dx, G.rdx) + M::per_dx(yb0 - yt0, G.dy, G.rdy);
      const float d1 = M::per_dx(xl1 - xr1, G.dx, G.rdx) + M::per_dx(yb1 - yt1, G.dy, G.rdy);
      const float d2 = M::per_dx(xl2 - xr2, G.dx, G.rdx) + M::per_dx(yb2 - yt2, G.dy, G.rdy);
      float hs = Rcur.h - dt * d0;
      const float hus = Rcur.hu - dt * d1;
      const float hvs = Rcur.hv - dt * d2;
      const float mult = G.mult[idx];
      hs = hs + dtrate * mult;
      // ---- reverse of the pointwise part ----
      const adj::PointVjp<T> pv =
          adj::point_vjp<T, kFastAdj>(dry, T(kPow73), T(hs), T(hus), T(hvs), T(G.n[idx]), Rcur.a,
                            T(dtrate), T(dtg), LH[idx], LHU[idx], LHV[idx]);
      if (real) {
        ls0 = pv.lh;
        ls1 = pv.lhu;
        ls2 = pv.lhv;
      }
      if (own && r >= j0 && r < j1) {
        if (grad_n) grad_n[idx] += pv.gn;
        if (grad_mult) grad_mult[idx] += pv.gmult;
      }
      lu0 = ls0;
      lu1 = ls1;
      lu2 = ls2;
      // ---- reverse of this lane's west x face (A = west cell, B = this cell) ----
      // Only for the strip's own rows: on its two halo rows the x faces feed
      // only that row's cotangent, which another strip owns and stores (the
      // halo rows' lambda* above is what the boundary y faces need).  r is
      // wave-uniform, so the branch is uniform and the DPP moves stay valid.
      if (r >= j0 && r < j1) {
      const T lw0 = from_west(ls0), lw1 = from_west(ls1), lw2 = from_west(ls2);
      adj::SideGrad<T> dA, dB;
      adj::face_vjp<true, T, kFastAdj>(g, hg, dry, T(wh), T(whu), T(whv), T(wb), wa, T(Rcur.h),
                             T(Rcur.hu), T(Rcur.hv), T(Rcur.b), Rcur.a, cx * (ls0 - lw0),
                             cx * (ls1 - lw1), cx * (ls2 - lw2), -cx * lw1, cx * ls1, dA, dB);
      // x: (h, q_n, q_t) = (h, hu, hv)
      lu0 += dB.h;
      lu1 += dB.qn;
      lu2 += dB.qt;
      T sw0 = dA.h, sw1 = dA.qn, sw2 = dA.qt;
      if (i == 0) {  // west cell is the left ghost: it copies this cell
        lu0 += dA.h;
        lu1 += refl_w ? -dA.qn : dA.qn;
        lu2 += dA.qt;
        sw0 = sw1 = sw2 = T(0);
      }
      if (i == G.nx) {  // this lane is the right ghost: its share belongs to cell nx-1
        sw0 += dB.h;
        sw1 += refl_e ? -dB.qn : dB.qn;
        sw2 += dB.qt;
      }
      lu0 += from_east(sw0);
      lu1 += from_east(sw1);
      lu2 += from_east(sw2);
      }
    }
    // ---- reverse of the y face r-1/2 (A = row r-1, B = row r) ----
    if (r >= j0) {
      adj::SideGrad<T> dA, dB;
      // y: normal momentum hv, tangential hu
      adj::face_vjp<false, T, kFastAdj>(g, hg, dry, T(Rprev.h), T(Rprev.hv), T(Rprev.hu), T(Rprev.b),
                              Rprev.a, T(Rcur.h), T(Rcur.hv), T(Rcur.hu), T(Rcur.b), Rcur.a,
                              cy * (ls0 - lsp0), cy * (ls2 - lsp2), cy * (ls1 - lsp1),
                              -cy * lsp2, cy * ls2, dA, dB);
      if (r - 1 < 0) {  // row r-1 is the south ghost: it copies row 0 (= row r)
        lu0 += dA.h;
        lu1 += dA.qt;
        lu2 += refl_s ? -dA.qn : dA.qn;
      } else {
        lup0 += dA.h;
        lup1 += dA.qt;
        lup2 += dA.qn;
      }
      if (r >= G.ny) {  // row r is the north ghost: it copies row ny-1 (= row r-1)
        lup0 += dB.h;
        lup1 += dB.qt;
        lup2 += refl_n ? -dB.qn : dB.qn;
      } else {
        lu0 += dB.h;
        lu1 += dB.qt;
        lu2 += dB.qn;
      }
      if (own && r - 1 >= j0 && r - 1 < j1) {
        const size_t idp = (size_t)(r - 1) * (size_t)G.nx + (size_t)rd.ic;
        OH[idp] = lup0;
        OHU[idp] = lup1;
        OHV[idp] = lup2;
      }
    }
    Gs = Gn;
    Rprev = Rcur;
    Rcur = Rnext;
    lsp0 = ls0;
    lsp1 = ls1;
    lsp2 = ls2;
    lup0 = lu0;
    lup1 = lu1;
    lup2 = lu2;
  }
}

}  // namespace swe

using namespace swe;

namespace {
template <class T>
int launch_adjoint(SweContext* c, const float* h, const float* hu, const float* hv, const T* lh,
                   const T* lhu, const T* lhv, T* oh, T* ohu, T* ohv, double dt, double time_s,
                   T* grad_manning, T* grad_mult, hipStream_t s);
}

namespace {
template <class T, class S>
int adjoint_step_any(SweContext* c, SweState state, S lam_next, S lam_out, double dt,
                     double time_s, T* grad_manning, T* grad_mult, void* stream) {
  if (!c || !state.h || !state.hu || !state.hv || !lam_next.h || !lam_next.hu || !lam_next.hv ||
      !lam_out.h || !lam_out.hu || !lam_out.hv)
    return set_error(SWE_EINVAL, "swe_adjoint_step: null argument");
  if (lam_out.h == lam_next.h || lam_out.hu == lam_next.hu || lam_out.hv == lam_next.hv)
    return set_error(SWE_EINVAL, "swe_adjoint_step: lam_out must not alias lam_next");
  return launch_adjoint<T>(c, state.h, state.hu, state.hv, lam_next.h, lam_next.hu, lam_next.hv,
                           lam_out.h, lam_out.hu, lam_out.hv, dt, time_s, grad_manning, grad_mult,
                           (hipStream_t)stream);
}

template <class T, class S>
int adjoint_sweep_any(SweContext* c, const float* states_h, const float* states_hu,
                      const float* states_hv, int32_t n, const double* dts, const double* ts,
                      S lam, S lam_tmp, T* grad_manning, T* grad_mult, void* stream) {
  if (!c || n < 0 || (n > 0 && (!states_h || !states_hu || !states_hv || !dts || !ts)) || !lam.h ||
      !lam.hu || !lam.hv || !lam_tmp.h || !lam_tmp.hu || !lam_tmp.hv || lam.h == lam_tmp.h)
    return set_error(SWE_EINVAL, "swe_adjoint_sweep: bad arguments");
  const Grid& G = *context_grid(c);
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  hipStream_t s = (hipStream_t)stream;
  // lambda ping-pongs between lam and lam_tmp; an odd count ends in lam_tmp,
  // so copy back once at the end.
  S a = lam, b = lam_tmp;
  for (int k = n - 1; k >= 0; --k) {
    int rc = launch_adjoint<T>(c, states_h + (size_t)k * N, states_hu + (size_t)k * N,
                               states_hv + (size_t)k * N, a.h, a.hu, a.hv, b.h, b.hu, b.hv, dts[k],
                               ts[k], grad_manning, grad_mult, s);
    if (rc) return rc;
    std::swap(a, b);
  }
  if (a.h != lam.h) {
    const size_t bytes = sizeof(T) * N;
    if (hipMemcpyAsync(lam.h, a.h, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpyAsync(lam.hu, a.hu, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpyAsync(lam.hv, a.hv, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
      return set_error(SWE_EHIP, "swe_adjoint_sweep: copy back");
  }
  return SWE_OK;
}
}  // namespace

extern "C" int swe_adjoint_step(SweContext* c, SweState state, SweState lam_next, SweState lam_out,
                                double dt, double time_s, float* grad_manning, float* grad_mult,
                                void* stream) {
  return adjoint_step_any<float>(c, state, lam_next, lam_out, dt, time_s, grad_manning, grad_mult,
                                 stream);
}

extern "C" int swe_adjoint_step_f64(SweContext* c, SweState state, SweStateF64 lam_next,
                                    SweStateF64 lam_out, double dt, double time_s,
                                    double* grad_manning, double* grad_mult, void* stream) {
  return adjoint_step_any<double>(c, state, lam_next, lam_out, dt, time_s, grad_manning,
                                  grad_mult, stream);
}

extern "C" int swe_adjoint_sweep(SweContext* c, const float* states_h, const float* states_hu,
                                 const float* states_hv, int32_t n, const double* dts,
                                 const double* ts, SweState lam, SweState lam_tmp,
                                 float* grad_manning, float* grad_mult, void* stream) {
  return adjoint_sweep_any<float>(c, states_h, states_hu, states_hv, n, dts, ts, lam, lam_tmp,
                                  grad_manning, grad_mult, stream);
}

extern "C" int swe_adjoint_sweep_f64(SweContext* c, const float* states_h, const float* states_hu,
                                     const float* states_hv, int32_t n, const double* dts,
                                     const double* ts, SweStateF64 lam, SweStateF64 lam_tmp,
                                     double* grad_manning, double* grad_mult, void* stream) {
  return adjoint_sweep_any<double>(c, states_h, states_hu, states_hv, n, dts, ts, lam, lam_tmp,
                                   grad_manning, grad_mult, stream);
}

namespace {
template <class T>
int launch_adjoint(SweContext* c, const float* h, const float* hu, const float* hv, const T* lh,
                   const T* lhu, const T* lhv, T* oh, T* ohu, T* ohv, double dt, double time_s,
                   T* grad_manning, T* grad_mult, hipStream_t s) {
  const Grid& G = *context_grid(c);
  const double rate = context_rate(c, time_s);
  const float dt_f = (float)dt, dtrate_f = (float)(dt * rate);
  const float dtg_f = (float)(dt * context_gravity(c));
  const int rows = context_adjoint_rows(c);
  const int segs_x = (G.nx + kAdjCols - 1) / kAdjCols;
  const int n_waves = segs_x * ((G.ny + rows - 1) / rows);
  const int blocks = (n_waves + 3) / 4;
  if (context_variant(c) == SWE_VARIANT_REFERENCE)
    hipLaunchKernelGGL((k_adjoint_step<RefMath, T>), dim3(blocks), dim3(256), 0, s, G, h, hu, hv,
                       lh, lhu, lhv, oh, ohu, ohv, grad_manning, grad_mult, dt_f, dtrate_f, dtg_f,
                       rows, segs_x, n_waves);
  else
    hipLaunchKernelGGL((k_adjoint_step<FastMath, T>), dim3(blocks), dim3(256), 0, s, G, h, hu, hv,
                       lh, lhu, lhv, oh, ohu, ohv, grad_manning, grad_mult, dt_f, dtrate_f, dtg_f,
                       rows, segs_x, n_waves);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(SWE_EHIP, std::string("swe_adjoint_step: ") + hipGetErrorString(e));
  return SWE_OK;
}
}  // namespace
