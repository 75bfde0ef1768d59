/* CPU restatement of the explicit 2-D shallow-water step in C, float32, run
 * on all host cores with OpenMP.
 *
 * TEST INFRASTRUCTURE ONLY.  Like oracle/swe_oracle.py, whose arithmetic it
 * repeats operation for operation, it is a checker and the CPU baseline that
 * bench.py times beside the GPU; the product path never calls it.
 *
 * Follows src/hybrid_flood/jax_solver/numerics.py:
 *   safe velocity 58-59, HLL 62-159, face fluxes with wall mirroring and
 *   hydrostatic reconstruction 162-283, CFL 297-309, the step 312-369;
 * ghost cells from jax_solver/boundary_conditions.py:29-52.  The order of the
 * operations is oracle/swe_oracle.py's (face_fluxes, hll, finite_volume_step,
 * cfl_inverse_dt), so with -ffp-contract=off every quantity rounds as NumPy
 * rounds it in float32; the one library function, powf of the friction term,
 * is glibc's here and may differ by an ulp from NumPy's own SIMD power.
 *
 * Layout: row-major (ny, nx) float32 fields, uint8 mask, boundary types
 * (left, right, bottom, top; 0 reflective, 1 open).  Each thread owns a block
 * of rows; within it the y face between two rows is evaluated once and the x
 * faces of a row once, so the work per cell is that of the vectorised NumPy
 * step.  Results do not depend on the thread count.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

typedef struct {
  float f0, fn, ft; /* flux: mass, normal and tangential momentum */
  float cA, cB;     /* pressure corrections of the A (west/south) and B side */
} Face;

typedef struct {
  float g, half_g, dry;
} Consts;

/* h > dry ? q / max(h, dry) : 0 */
static inline float safe_velocity(float q, float h, float dry) {
  return h > dry ? q / (h > dry ? h : dry) : 0.0f;
}

static inline float max0(float x) { return x > 0.0f ? x : 0.0f; }
/* np.maximum / np.minimum of two finite floats */
static inline float fmx(float a, float b) { return a >= b ? a : b; }
static inline float fmn(float a, float b) { return a <= b ? a : b; }

/* One face in face-normal coordinates; Y selects the y-face tangential flux
 * q_t * v_n' (numerics.py:135) instead of q_n * v_t' (numerics.py:85). */
static inline Face face(int Y, float hA, float qnA, float qtA, float bA, int aA, float hB,
                        float qnB, float qtB, float bB, int aB, Consts k) {
  Face F = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  const int wallA = aA && !aB, wallB = !aA && aB;
  if (wallA) {
    hB = hA;
    qnB = -qnA;
    qtB = qtA;
    bB = bA;
  }
  if (wallB) {
    hA = hB;
    qnA = -qnB;
    qtA = qtB;
    bA = bB;
  }
  const float bf = fmx(bA, bB);
  const float hrA = max0((hA + bA) - bf), hrB = max0((hB + bB) - bf);
  const float unA = safe_velocity(qnA, hA, k.dry), utA = safe_velocity(qtA, hA, k.dry);
  const float unB = safe_velocity(qnB, hB, k.dry), utB = safe_velocity(qtB, hB, k.dry);
  /* HLL on the reconstructed states (h, q_n, q_t) = (hr, hr u_n, hr u_t) */
  const float QnA = hrA * unA, QtA = hrA * utA, QnB = hrB * unB, QtB = hrB * utB;
  const float vnA = safe_velocity(QnA, hrA, k.dry), vnB = safe_velocity(QnB, hrB, k.dry);
  const float cA = sqrtf(k.g * max0(hrA)), cB = sqrtf(k.g * max0(hrB));
  const float sA = fmn(vnA - cA, vnB - cB), sB = fmx(vnA + cA, vnB + cB);
  float ftA, ftB;
  if (Y) {
    ftA = QtA * vnA;
    ftB = QtB * vnB;
  } else {
    ftA = QnA * safe_velocity(QtA, hrA, k.dry);
    ftB = QnB * safe_velocity(QtB, hrB, k.dry);
  }
  const float fA[3] = {QnA, QnA * vnA + k.half_g * (hrA * hrA), ftA};
  const float fB[3] = {QnB, QnB * vnB + k.half_g * (hrB * hrB), ftB};
  const float uA[3] = {hrA, QnA, QtA}, uB[3] = {hrB, QnB, QtB};
  const float spread = sB - sA;
  const float den = spread > 1.0e-12f ? spread : 1.0f;
  const float sAsB = sA * sB;
  float out[3];
  for (int c = 0; c < 3; ++c) {
    const float mid = ((sB * fA[c] - sA * fB[c]) + sAsB * (uB[c] - uA[c])) / den;
    out[c] = sA >= 0.0f ? fA[c] : (sB <= 0.0f ? fB[c] : mid);
  }
  if (aA || aB) {
    F.f0 = out[0];
    F.fn = out[1];
    F.ft = out[2];
    F.cA = k.half_g * (hA * hA - hrA * hrA);
    F.cB = k.half_g * (hB * hB - hrB * hrB);
  }
  return F;
}

typedef struct {
  int ny, nx;
  const float *h, *hu, *hv, *bed, *n, *mult;
  const uint8_t* mask;
  int bt[4];
} Grid;

/* Cell (j, i) of the ghosted grid, j in [-1, ny], i in [-1, nx]: the edge cell,
 * with the normal momentum negated on reflective sides. */
static inline void ghosted(const Grid* G, int j, int i, float* h, float* hu, float* hv, float* b,
                           int* a) {
  const int jj = j < 0 ? 0 : (j >= G->ny ? G->ny - 1 : j);
  const int ii = i < 0 ? 0 : (i >= G->nx ? G->nx - 1 : i);
  const size_t idx = (size_t)jj * (size_t)G->nx + (size_t)ii;
  *h = G->h[idx];
  *hu = G->hu[idx];
  *hv = G->hv[idx];
  *b = G->bed[idx];
  *a = G->mask[idx] != 0;
  if ((i < 0 && G->bt[0] == 0) || (i >= G->nx && G->bt[1] == 0)) *hu = -*hu;
  if ((j < 0 && G->bt[2] == 0) || (j >= G->ny && G->bt[3] == 0)) *hv = -*hv;
}

/* y faces between rows j-1 and j (j in [0, ny]) for every column into F[0..nx) */
static void y_faces(const Grid* G, int j, Consts k, Face* F) {
  for (int i = 0; i < G->nx; ++i) {
    float hA, huA, hvA, bA, hB, huB, hvB, bB;
    int aA, aB;
    ghosted(G, j - 1, i, &hA, &huA, &hvA, &bA, &aA);
    ghosted(G, j, i, &hB, &huB, &hvB, &bB, &aB);
    F[i] = face(1, hA, hvA, huA, bA, aA, hB, hvB, huB, bB, aB, k);
  }
}

/* One step: out = step(in) with float32 scalars dt, dt*rate and dt*g as the
 * NumPy oracle rounds them.  Returns 0, or -1 on bad arguments. */
int swe_omp_step(int ny, int nx, const float* h, const float* hu, const float* hv,
                 const float* bed, const float* n, const uint8_t* mask, const float* mult,
                 const int32_t* bt, float g, float half_g, float dry, float dx, float dy, float dt,
                 float dtrate, float dtg, float p73, float* ho, float* huo, float* hvo) {
  if (ny < 1 || nx < 1 || !h || !hu || !hv || !bed || !n || !mask || !mult || !bt || !ho ||
      !huo || !hvo)
    return -1;
  Grid G = {ny, nx, h, hu, hv, bed, n, mult, mask, {bt[0], bt[1], bt[2], bt[3]}};
  const Consts k = {g, half_g, dry};
  int status = 0;
#pragma omp parallel
  {
    const int nt = omp_get_num_threads(), t = omp_get_thread_num();
    const int j0 = (int)((long long)ny * t / nt), j1 = (int)((long long)ny * (t + 1) / nt);
    Face* fx = (Face*)malloc(sizeof(Face) * (size_t)(nx + 1));
    Face* ys = (Face*)malloc(sizeof(Face) * (size_t)nx);
    Face* yn = (Face*)malloc(sizeof(Face) * (size_t)nx);
    if (!fx || !ys || !yn) {
#pragma omp atomic write
      status = -2;
    } else if (j0 < j1) {
      y_faces(&G, j0, k, ys);
      for (int j = j0; j < j1; ++j) {
        y_faces(&G, j + 1, k, yn);
        for (int f = 0; f <= nx; ++f) {  /* x face f between columns f-1 and f */
          float hA, huA, hvA, bA, hB, huB, hvB, bB;
          int aA, aB;
          ghosted(&G, j, f - 1, &hA, &huA, &hvA, &bA, &aA);
          ghosted(&G, j, f, &hB, &huB, &hvB, &bB, &aB);
          fx[f] = face(0, hA, huA, hvA, bA, aA, hB, huB, hvB, bB, aB, k);
        }
        for (int i = 0; i < nx; ++i) {
          const size_t idx = (size_t)j * (size_t)nx + (size_t)i;
          const Face W = fx[i], E = fx[i + 1], S = ys[i], N = yn[i];
          /* (x_left[i+1] - x_right[i]) / dx + (y_bottom[j+1] - y_top[j]) / dy,
           * the side's own correction on the normal component */
          const float dxh = ((E.f0 + 0.0f) - (W.f0 + 0.0f)) / dx;
          const float dxu = ((E.fn + E.cA) - (W.fn + W.cB)) / dx;
          const float dxv = ((E.ft + 0.0f) - (W.ft + 0.0f)) / dx;
          const float dyh = ((N.f0 + 0.0f) - (S.f0 + 0.0f)) / dy;
          const float dyu = ((N.ft + 0.0f) - (S.ft + 0.0f)) / dy;
          const float dyv = ((N.fn + N.cA) - (S.fn + S.cB)) / dy;
          const float hs = h[idx] - dt * (dxh + dyh);
          const float hus = hu[idx] - dt * (dxu + dyu);
          const float hvs = hv[idx] - dt * (dxv + dyv);
          const float h_star = hs + dtrate * mult[idx];
          const float h_new = max0(h_star);
          const float q = sqrtf(hus * hus + hvs * hvs);
          const float n2 = n[idx] * n[idx];
          const float denom = 1.0f + ((dtg * n2) * q) / powf(fmx(h_new, dry), p73);
          const float hu_new = hus / denom, hv_new = hvs / denom;
          const int a = mask[idx] != 0;
          const int wet = a && h_new > dry;
          ho[idx] = a ? h_new : 0.0f;
          huo[idx] = wet ? hu_new : 0.0f;
          hvo[idx] = wet ? hv_new : 0.0f;
        }
        Face* tmp = ys;
        ys = yn;
        yn = tmp;
      }
    }
    free(fx);
    free(ys);
    free(yn);
  }
  return status;
}

/* Maximum over active cells of (|u| + c)/dx + (|v| + c)/dy in float32
 * (numerics.py:297-303); 0 when no cell is active. */
float swe_omp_cfl_max(int ny, int nx, const float* h, const float* hu, const float* hv,
                      const uint8_t* mask, float g, float dry, float dx, float dy) {
  float best = 0.0f;
  const long long N = (long long)ny * nx;
#pragma omp parallel for reduction(max : best) schedule(static)
  for (long long c = 0; c < N; ++c) {
    const float u = safe_velocity(hu[c], h[c], dry), v = safe_velocity(hv[c], h[c], dry);
    const float wave = sqrtf(g * max0(h[c]));
    const float inv = (fabsf(u) + wave) / dx + (fabsf(v) + wave) / dy;
    const float x = mask[c] ? inv : 0.0f;
    best = x > best ? x : best;
  }
  return best;
}

int swe_omp_threads(void) { return omp_get_max_threads(); }

void swe_omp_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}
