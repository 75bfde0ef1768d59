// Reverse-mode pieces of the step, written by hand.
//
// face_fwd     one interface in exact arithmetic form (the HLL velocity is
//              the cell velocity, (hr*u)/hr = u), the function face_vjp
//              differentiates (jax_solver/numerics.py:62-283).
// face_vjp     cotangents of one interface's outputs (flux F and the two
//              pressure corrections) -> cotangents of the conserved state of
//              its two cells, through HLL, hydrostatic reconstruction and
//              wall mirroring.
// point_vjp    cotangent of a cell's new state -> cotangent of its
//              pre-source state U* plus the local Manning-n and rain-
//              multiplier sensitivities, through masking, semi-implicit
//              friction, the positivity clip and rain (numerics.py:350-369).
//
// The step size is data (a recorded schedule), not differentiated.  Kinks
// take the branch the forward took: max(x, 0) passes the cotangent where
// x > 0, min/max of the wave speeds pass it to the selected argument, the
// HLL region (left/right upwind or star) follows the forward's speeds.  One
// deliberate departure from the exact derivative: the wave speed of a wet
// cell's side reconstructed to 0 < hr <= dry is not differentiated (see
// face_vjp).
//
// Templated on the scalar type: float on the GPU, double in the host build
// that checks them against finite differences.  No HIP header is needed.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define SWE_HD __host__ __device__
#else
#define SWE_HD
#endif

// The reverse pieces may fuse a*b+c into FMAs on the GPU: they carry no
// bit-parity contract (the library is otherwise built with
// -ffp-contract=off for the forward step's); the host build that finite
// differences check stays uncontracted.
#if defined(__HIP_DEVICE_COMPILE__)
#define SWE_VJP_CONTRACT _Pragma("clang fp contract(fast)")
#else
#define SWE_VJP_CONTRACT
#endif

namespace swe {
namespace adj {

template <class T>
struct Face5 {
  T f0, fn, ft, cA, cB;
};

template <class T>
struct SideGrad {
  T h, qn, qt;  // d/d(h, normal momentum, tangential momentum) of one cell
};

template <class T>
SWE_HD inline T tmax(T a, T b) { return a > b ? a : b; }

// Arithmetic of the reverse pieces.  FAST (float on the GPU, for the fast
// kernel variant) uses the hardware reciprocal, square root and log2/exp2,
// ~1 ulp each, as the fast forward step does; otherwise (the host build that
// finite differences check, double cotangents, the reference variant) the
// operations are correctly rounded.  Quotients whose divisor can be
// denormal stay exact in both.
template <bool FAST, class T>
SWE_HD inline T quot(T a, T b) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (FAST && sizeof(T) == 4) return a * __builtin_amdgcn_rcpf(b);
#endif
  return a / b;
}

template <bool FAST, class T>
SWE_HD inline T tsqrt(T x) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (FAST && sizeof(T) == 4) return __builtin_amdgcn_sqrtf(x);
#endif
  return std::sqrt(x);
}

template <bool FAST, class T>
SWE_HD inline T tpow(T x, T p) {  // x > 0
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (FAST && sizeof(T) == 4) return __builtin_amdgcn_exp2f(p * __builtin_amdgcn_logf(x));
#endif
  return std::pow(x, p);
}

// Face in face-normal coordinates: side A is west/south, q_n the normal and
// q_t the tangential momentum of each cell (before mirroring).  XDIR: the x
// face's tangential flux is q_n v_t, the y face's q_t v_n.
template <bool XDIR, class T>
SWE_HD inline Face5<T> face_fwd(T g, T hg, T dry, T hA, T qnA, T qtA, T bA, int aA, T hB, T qnB,
                                T qtB, T bB, int aB) {
  Face5<T> o{T(0), T(0), T(0), T(0), T(0)};
  if (!(aA || aB)) return o;
  if (aA && !aB) { hB = hA; qnB = -qnA; qtB = qtA; bB = bA; }
  if (!aA && aB) { hA = hB; qnA = -qnB; qtA = qtB; bA = bB; }
  const T uA = hA > dry ? qnA / hA : T(0), vA = hA > dry ? qtA / hA : T(0);
  const T uB = hB > dry ? qnB / hB : T(0), vB = hB > dry ? qtB / hB : T(0);
  const T bf = tmax(bA, bB);
  const T eA = (hA + bA) - bf, eB = (hB + bB) - bf;
  const T hrA = eA > T(0) ? eA : T(0), hrB = eB > T(0) ? eB : T(0);
  const T QnA = hrA * uA, QtA = hrA * vA, QnB = hrB * uB, QtB = hrB * vB;
  const T vnA = hrA > dry ? uA : T(0), vtA = hrA > dry ? vA : T(0);
  const T vnB = hrB > dry ? uB : T(0), vtB = hrB > dry ? vB : T(0);
  const T cA = std::sqrt(g * hrA), cB = std::sqrt(g * hrB);
  const T s1A = vnA - cA, s1B = vnB - cB;
  const T sA = s1A <= s1B ? s1A : s1B;
  const T s2A = vnA + cA, s2B = vnB + cB;
  const T sB = s2A >= s2B ? s2A : s2B;
  const T FA[3] = {QnA, QnA * vnA + hg * (hrA * hrA), XDIR ? QnA * vtA : QtA * vnA};
  const T FB[3] = {QnB, QnB * vnB + hg * (hrB * hrB), XDIR ? QnB * vtB : QtB * vnB};
  const T dU[3] = {hrB - hrA, QnB - QnA, QtB - QtA};
  T F[3];
  const T spread = sB - sA;
  const T den = spread > T(1.0e-12) ? spread : T(1);
  for (int k = 0; k < 3; ++k) {
    const T mid = ((sB * FA[k] - sA * FB[k]) + (sA * sB) * dU[k]) / den;
    F[k] = sA >= T(0) ? FA[k] : (sB <= T(0) ? FB[k] : mid);
  }
  o.f0 = F[0];
  o.fn = F[1];
  o.ft = F[2];
  o.cA = hg * (hA * hA - hrA * hrA);
  o.cB = hg * (hB * hB - hrB * hrB);
  return o;
}

template <bool XDIR, class T, bool FAST = false>
SWE_HD inline void face_vjp(T g, T hg, T dry, T hA, T qnA, T qtA, T bA, int aA, T hB, T qnB,
                            T qtB, T bB, int aB, T gF0, T gFn, T gFt, T gcA, T gcB,
                            SideGrad<T>& dA, SideGrad<T>& dB) {
  SWE_VJP_CONTRACT
  dA = {T(0), T(0), T(0)};
  dB = {T(0), T(0), T(0)};
  if (!(aA || aB)) return;  // irrelevant face: zero flux, no dependence
  const bool wallA = aA && !aB;
  const bool wallB = !aA && aB;
  if (wallA) { hB = hA; qnB = -qnA; qtB = qtA; bB = bA; }
  if (wallB) { hA = hB; qnA = -qnB; qtA = qtB; bA = bB; }
  const bool dwA = hA > dry, dwB = hB > dry;
  const T uA = dwA ? quot<FAST>(qnA, hA) : T(0), vA = dwA ? quot<FAST>(qtA, hA) : T(0);
  const T uB = dwB ? quot<FAST>(qnB, hB) : T(0), vB = dwB ? quot<FAST>(qtB, hB) : T(0);
  const T bf = tmax(bA, bB);
  const T eA = (hA + bA) - bf, eB = (hB + bB) - bf;
  const T hrA = eA > T(0) ? eA : T(0), hrB = eB > T(0) ? eB : T(0);
  const T QnA = hrA * uA, QtA = hrA * vA, QnB = hrB * uB, QtB = hrB * vB;
  const bool wA = hrA > dry, wB = hrB > dry;
  const T vnA = wA ? uA : T(0), vtA = wA ? vA : T(0);
  const T vnB = wB ? uB : T(0), vtB = wB ? vB : T(0);
  const T cA = tsqrt<FAST>(g * hrA), cB = tsqrt<FAST>(g * hrB);
  const T s1A = vnA - cA, s1B = vnB - cB;
  const bool minA = s1A <= s1B;
  const T sA = minA ? s1A : s1B;
  const T s2A = vnA + cA, s2B = vnB + cB;
  const bool maxA = s2A >= s2B;
  const T sB = maxA ? s2A : s2B;
  const T FA[3] = {QnA, QnA * vnA + hg * (hrA * hrA), XDIR ? QnA * vtA : QtA * vnA};
  const T FB[3] = {QnB, QnB * vnB + hg * (hrB * hrB), XDIR ? QnB * vtB : QtB * vnB};
  const T dU[3] = {hrB - hrA, QnB - QnA, QtB - QtA};
  const T gF[3] = {gF0, gFn, gFt};

  T gFA[3] = {T(0), T(0), T(0)}, gFB[3] = {T(0), T(0), T(0)};
  T gUA[3] = {T(0), T(0), T(0)}, gUB[3] = {T(0), T(0), T(0)};
  T gsA = T(0), gsB = T(0);
  if (sA >= T(0)) {
    for (int k = 0; k < 3; ++k) gFA[k] = gF[k];
  } else if (sB <= T(0)) {
    for (int k = 0; k < 3; ++k) gFB[k] = gF[k];
  } else {
    const T spread = sB - sA;
    const bool clamped = !(spread > T(1.0e-12));
    const T r = clamped ? T(1) : quot<FAST>(T(1), spread);
    for (int k = 0; k < 3; ++k) {
      const T mid = ((sB * FA[k] - sA * FB[k]) + (sA * sB) * dU[k]) * r;
      gFA[k] = gF[k] * sB * r;
      gFB[k] = -gF[k] * sA * r;
      const T gd = gF[k] * sA * sB * r;
      gUB[k] += gd;
      gUA[k] -= gd;
      const T m = clamped ? T(0) : mid;
      gsB += gF[k] * (FA[k] + sA * dU[k] - m) * r;
      gsA += gF[k] * (-FB[k] + sB * dU[k] + m) * r;
    }
  }
  T gvnA = T(0), gvnB = T(0), gcsA = T(0), gcsB = T(0);
  if (minA) { gvnA += gsA; gcsA -= gsA; } else { gvnB += gsA; gcsB -= gsA; }
  if (maxA) { gvnA += gsB; gcsA += gsB; } else { gvnB += gsB; gcsB += gsB; }

  // Fluxes (Qn, Qn vn + hg hr^2, Ft) and state (hr, Qn, Qt) of each side.
  T ghrA = gUA[0], gQnA = gUA[1] + gFA[0], gQtA = gUA[2], gvtA = T(0);
  gQnA += gFA[1] * vnA;
  gvnA += gFA[1] * QnA;
  ghrA += gFA[1] * T(2) * hg * hrA;
  T ghrB = gUB[0], gQnB = gUB[1] + gFB[0], gQtB = gUB[2], gvtB = T(0);
  gQnB += gFB[1] * vnB;
  gvnB += gFB[1] * QnB;
  ghrB += gFB[1] * T(2) * hg * hrB;
  if (XDIR) {
    gQnA += gFA[2] * vtA;
    gvtA += gFA[2] * QnA;
    gQnB += gFB[2] * vtB;
    gvtB += gFB[2] * QnB;
  } else {
    gQtA += gFA[2] * vnA;
    gvnA += gFA[2] * QtA;
    gQtB += gFB[2] * vnB;
    gvnB += gFB[2] * QtB;
  }
  // d sqrt(g hr)/d hr = g / (2c) is unbounded as hr -> 0+.  When a wet cell
  // (h > dry) is reconstructed to a dry side (hr <= dry) -- its water level
  // within the dry tolerance of the neighbour's higher bed, i.e. of the kink
  // of hr = max(eta - bf, 0) -- the wave speed is held fixed: the exact
  // derivative there holds only for perturbations smaller than hr and,
  // accumulated over a rollout, swamps the gradient that finite differences
  // of any practical size see.  Shallow cells (h <= dry) keep it: their thin
  // films change smoothly with the forcing.
  if (cA > T(0) && (wA || !dwA)) ghrA += quot<FAST>(gcsA * g, T(2) * cA);
  if (cB > T(0) && (wB || !dwB)) ghrB += quot<FAST>(gcsB * g, T(2) * cB);
  T guA = wA ? gvnA : T(0), gvA = wA ? gvtA : T(0);
  T guB = wB ? gvnB : T(0), gvB = wB ? gvtB : T(0);
  ghrA += gQnA * uA + gQtA * vA;
  guA += gQnA * hrA;
  gvA += gQtA * hrA;
  ghrB += gQnB * uB + gQtB * vB;
  guB += gQnB * hrB;
  gvB += gQtB * hrB;
  // pressure corrections hg (h^2 - hr^2)
  T ghA = gcA * T(2) * hg * hA;
  ghrA -= gcA * T(2) * hg * hrA;
  T ghB = gcB * T(2) * hg * hB;
  ghrB -= gcB * T(2) * hg * hrB;
  if (eA > T(0)) ghA += ghrA;
  if (eB > T(0)) ghB += ghrB;
  T gqnA = T(0), gqtA = T(0), gqnB = T(0), gqtB = T(0);
  if (dwA) {
    gqnA = quot<FAST>(guA, hA);
    gqtA = quot<FAST>(gvA, hA);
    ghA -= quot<FAST>(guA * uA + gvA * vA, hA);
  }
  if (dwB) {
    gqnB = quot<FAST>(guB, hB);
    gqtB = quot<FAST>(gvB, hB);
    ghB -= quot<FAST>(guB * uB + gvB * vB, hB);
  }
  // Undo the mirroring: the mirrored side is a function of the active side.
  if (wallA) {
    dA = {ghA + ghB, gqnA - gqnB, gqtA + gqtB};
  } else if (wallB) {
    dB = {ghB + ghA, gqnB - gqnA, gqtB + gqtA};
  } else {
    dA = {ghA, gqnA, gqtA};
    dB = {ghB, gqnB, gqtB};
  }
}

template <class T>
struct PointOut {
  T h, hu, hv;
};

template <class T>
struct PointVjp {
  T lh, lhu, lhv;  // cotangent of U* (before the sources)
  T gn;            // d/d(manning n) of this cell's step
  T gmult;         // d/d(rain multiplier) of this cell's step
};

// Forward of the pointwise part: h* = hs (rain included), hn = max(h*, 0),
// q = |(hu*, hv*)|, D = 1 + dtg n^2 q / max(hn, dry)^p, (hu, hv) = (hu*, hv*)/D,
// then masking.  p is the friction exponent (float32(7/3) on the GPU).
template <class T>
SWE_HD inline PointOut<T> point_fwd(T dry, T p, T hs, T hus, T hvs, T n, int active, T dtg) {
  if (!active) return {T(0), T(0), T(0)};
  const T hn = hs > T(0) ? hs : T(0);
  const T q = std::sqrt(hus * hus + hvs * hvs);
  const T P = std::pow(hn > dry ? hn : dry, p);
  const T D = T(1) + ((dtg * (n * n)) * q) / P;
  const bool wet = hn > dry;
  return {hn, wet ? hus / D : T(0), wet ? hvs / D : T(0)};
}

template <class T, bool FAST = false>
SWE_HD inline PointVjp<T> point_vjp(T dry, T p, T hs, T hus, T hvs, T n, int active, T dtrate,
                                    T dtg, T lam_h, T lam_hu, T lam_hv) {
  SWE_VJP_CONTRACT
  PointVjp<T> r{T(0), T(0), T(0), T(0), T(0)};
  if (!active) return r;
  const T hn = hs > T(0) ? hs : T(0);
  const T q = tsqrt<FAST>(hus * hus + hvs * hvs);
  const T hd = hn > dry ? hn : dry;
  const T P = tpow<FAST>(hd, p);
  const T K = quot<FAST>((dtg * (n * n)) * q, P);
  const T D = T(1) + K;
  const bool wet = hn > dry;
  const T lhu_n = wet ? lam_hu : T(0);
  const T lhv_n = wet ? lam_hv : T(0);
  const T s = quot<FAST>(lhu_n * hus + lhv_n * hvs, D * D);  // = -dL/dD
  T lhn = lam_h;
  T lhus = quot<FAST>(lhu_n, D), lhvs = quot<FAST>(lhv_n, D);
  if (q > T(0)) {
    const T lq = -s * K / q;  // dD/dq = K/q; q may be denormal: exact quotients
    lhus += lq * hus / q;
    lhvs += lq * hvs / q;
  }
  if (hn > dry) lhn += quot<FAST>(s * p * K, hn);  // dD/dh = -p K / h
  r.gn = (n != T(0)) ? -s * T(2) * K / n : T(0);
  const T lhs = hs > T(0) ? lhn : T(0);
  r.lh = lhs;
  r.lhu = lhus;
  r.lhv = lhvs;
  r.gmult = lhs * dtrate;
  return r;
}

}  // namespace adj
}  // namespace swe
