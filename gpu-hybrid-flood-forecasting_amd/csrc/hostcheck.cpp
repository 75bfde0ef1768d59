// Host build of the hand-written reverse-mode pieces (swe_adjoint.hpp) in
// double precision, exported for the CPU tests that check them against
// finite differences of the same forward.  Not part of the GPU path.
#include <cstdint>

#include "swe_adjoint.hpp"

using namespace swe::adj;

extern "C" {

// in: g, hg, dry, hA, qnA, qtA, bA, aA, hB, qnB, qtB, bB, aB ; out: f0, fn, ft, cA, cB
void swe_hc_face_fwd(int32_t xdir, const double* in, double* out) {
  Face5<double> o = xdir ? face_fwd<true>(in[0], in[1], in[2], in[3], in[4], in[5], in[6], (int)in[7],
                                          in[8], in[9], in[10], in[11], (int)in[12])
                         : face_fwd<false>(in[0], in[1], in[2], in[3], in[4], in[5], in[6], (int)in[7],
                                           in[8], in[9], in[10], in[11], (int)in[12]);
  out[0] = o.f0;
  out[1] = o.fn;
  out[2] = o.ft;
  out[3] = o.cA;
  out[4] = o.cB;
}

// gout: cotangents of (f0, fn, ft, cA, cB); gin: d/d(hA, qnA, qtA, hB, qnB, qtB)
void swe_hc_face_vjp(int32_t xdir, const double* in, const double* gout, double* gin) {
  SideGrad<double> a, b;
  if (xdir)
    face_vjp<true>(in[0], in[1], in[2], in[3], in[4], in[5], in[6], (int)in[7], in[8], in[9],
                   in[10], in[11], (int)in[12], gout[0], gout[1], gout[2], gout[3], gout[4], a, b);
  else
    face_vjp<false>(in[0], in[1], in[2], in[3], in[4], in[5], in[6], (int)in[7], in[8], in[9],
                    in[10], in[11], (int)in[12], gout[0], gout[1], gout[2], gout[3], gout[4], a, b);
  gin[0] = a.h;
  gin[1] = a.qn;
  gin[2] = a.qt;
  gin[3] = b.h;
  gin[4] = b.qn;
  gin[5] = b.qt;
}

// in: dry, p, hs, hus, hvs, n, active, dtg ; out: h, hu, hv
void swe_hc_point_fwd(const double* in, double* out) {
  PointOut<double> o = point_fwd<double>(in[0], in[1], in[2], in[3], in[4], in[5], (int)in[6], in[7]);
  out[0] = o.h;
  out[1] = o.hu;
  out[2] = o.hv;
}

// in: dry, p, hs, hus, hvs, n, active, dtrate, dtg, lam_h, lam_hu, lam_hv
// out: lh, lhu, lhv, gn, gmult
void swe_hc_point_vjp(const double* in, double* out) {
  PointVjp<double> r = point_vjp<double>(in[0], in[1], in[2], in[3], in[4], in[5], (int)in[6],
                                         in[7], in[8], in[9], in[10], in[11]);
  out[0] = r.lh;
  out[1] = r.lhu;
  out[2] = r.lhv;
  out[3] = r.gn;
  out[4] = r.gmult;
}

}  // extern "C"
