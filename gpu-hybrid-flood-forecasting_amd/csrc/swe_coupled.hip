// Kernels of the hybrid coupled forecast around the step
// (hybrid/coupled_forecast.py:53-120):
//   - network_features: the six normalised CNN input channels of a state in
//     one pass (depth, the two safe velocities, the rain channel, terrain,
//     roughness), written in any channel/cell stride (NHWC or NCHW);
//   - apply_residual_correction, fused with the residual clip of
//     coupled_forecast.py:199-201: positivity clip, a global rescale that
//     restores the physics depth volume, and the velocity correction.  The
//     two volume sums are float64 (per-workgroup partials in a fixed order,
//     then one workgroup), so the rescale is deterministic and the corrected
//     volume matches the physics volume to float32 rounding of each cell.
// Also the batched staging of comparison/v2_staging.py (features of many
// primitive states at once, and the depth correction of a batch).
// Element-wise arithmetic follows the reference's float32 order (the build
// uses -ffp-contract=off and IEEE division).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <mutex>
#include <string>

#include "../../include/swe_hip.h"
#include "swe_device.hpp"
#include "swe_host.hpp"
#include "swe_loop.hpp"

namespace swe {
namespace {

// coupled_forecast.py:49-50: h > dry ? q / max(h, dry) : 0
__device__ __forceinline__ float safe_velocity(float q, float h, float dry) {
  return h > dry ? q / fmaxf(h, dry) : 0.0f;
}

// jnp.clip(x, -lim, lim), NaN passes through.
__device__ __forceinline__ float clip_to(float x, float lim) {
  return x < -lim ? -lim : (x > lim ? lim : x);
}

struct Norm {
  float mean[3], std[3];
};

// Write the six channels v of the cells base .. base+kBlock-1 (this thread's
// cell is base + threadIdx.x; cells >= N are skipped).  Channel c of cell k
// goes to o0[c * cs + k * ps].  For NHWC (cs 1, ps 6) at a 16-byte aligned
// o0 the block's cells are 6 * count consecutive floats: they are staged in
// LDS and written as 16-byte vectors, so each store instruction covers 1 KiB
// of consecutive bytes instead of six stores of 4 bytes at a 24-byte stride.
// Called by every thread of the block (it synchronises).
__device__ __forceinline__ void write_cells6(float* __restrict__ o0, size_t base, size_t N,
                                             const float (&v)[6], long long cs, long long ps) {
  __shared__ __align__(16) float tile[kBlock * 6];
  const size_t k = base + threadIdx.x;
  const bool vec = cs == 1 && ps == 6 && ((uintptr_t)o0 & 15) == 0;  // block-uniform
  if (!vec) {
    if (k < N) {
      float* o = o0 + (long long)k * ps;
#pragma unroll
      for (int c = 0; c < 6; ++c) o[c * cs] = v[c];
    }
    return;
  }
  if (k < N) {
#pragma unroll
    for (int c = 0; c < 6; ++c) tile[threadIdx.x * 6 + c] = v[c];
  }
  __syncthreads();
  const int nf = (int)(6 * min<size_t>(kBlock, N - base));
  float* dst = o0 + base * 6;  // base is a multiple of kBlock: 16-byte aligned
  for (int q = threadIdx.x; q < nf / 4; q += kBlock)
    reinterpret_cast<float4*>(dst)[q] = reinterpret_cast<const float4*>(tile)[q];
  for (int q = (nf / 4) * 4 + threadIdx.x; q < nf; q += kBlock) dst[q] = tile[q];
  __syncthreads();
}

__global__ __launch_bounds__(kBlock) void k_features(Grid G, const float* __restrict__ H,
                                                     const float* __restrict__ HU,
                                                     const float* __restrict__ HV,
                                                     const float* __restrict__ terrain,
                                                     const float* __restrict__ rough, Norm nz,
                                                     float rain_channel, float* __restrict__ out,
                                                     long long cs, long long ps) {
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  // Block-uniform loop over tiles of kBlock cells (write_cells6 synchronises).
  for (size_t base = (size_t)blockIdx.x * kBlock; base < N; base += (size_t)gridDim.x * kBlock) {
    const size_t k = base + threadIdx.x;
    float f[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (k < N) {
      const float h = H[k];
      const float u = safe_velocity(HU[k], h, G.dry);
      const float v = safe_velocity(HV[k], h, G.dry);
      const bool a = G.mask[k] != 0;
      f[0] = a ? (h - nz.mean[0]) / nz.std[0] : 0.0f;
      f[1] = a ? (u - nz.mean[1]) / nz.std[1] : 0.0f;
      f[2] = a ? (v - nz.mean[2]) / nz.std[2] : 0.0f;
      f[3] = rain_channel;
      f[4] = terrain[k];
      f[5] = rough[k];
    }
    write_cells6(out, base, N, f, cs, ps);
  }
}

// Four cells per thread for NHWC features (channel stride 1, cell stride 6)
// when the cell count is a multiple of 4 and every array is aligned: 16-byte
// loads, the mask as one 32-bit word, and the block's 24 * kBlock output
// floats staged in LDS and written as consecutive 16-byte vectors.  Each
// cell's arithmetic is that of k_features.
__global__ __launch_bounds__(kBlock) void k_features4(Grid G, const float* __restrict__ H,
                                                      const float* __restrict__ HU,
                                                      const float* __restrict__ HV,
                                                      const float* __restrict__ terrain,
                                                      const float* __restrict__ rough, Norm nz,
                                                      float rain_channel, float* __restrict__ out) {
  __shared__ __align__(16) float tile[kBlock * 24];
  const size_t N4 = (size_t)G.nx * (size_t)G.ny / 4;
  const uint32_t* __restrict__ M = reinterpret_cast<const uint32_t*>(G.mask);
  const float4* __restrict__ H4 = reinterpret_cast<const float4*>(H);
  const float4* __restrict__ HU4 = reinterpret_cast<const float4*>(HU);
  const float4* __restrict__ HV4 = reinterpret_cast<const float4*>(HV);
  const float4* __restrict__ T4 = reinterpret_cast<const float4*>(terrain);
  const float4* __restrict__ R4 = reinterpret_cast<const float4*>(rough);
  // Block-uniform loop over tiles of kBlock quads of cells.
  for (size_t base = (size_t)blockIdx.x * kBlock; base < N4; base += (size_t)gridDim.x * kBlock) {
    const size_t q = base + threadIdx.x;
    if (q < N4) {
      const float4 h4 = H4[q], hu4 = HU4[q], hv4 = HV4[q], t4 = T4[q], r4 = R4[q];
      const uint32_t m = M[q];
      const float hs[4] = {h4.x, h4.y, h4.z, h4.w};
      const float hus[4] = {hu4.x, hu4.y, hu4.z, hu4.w};
      const float hvs[4] = {hv4.x, hv4.y, hv4.z, hv4.w};
      const float ts[4] = {t4.x, t4.y, t4.z, t4.w};
      const float rs[4] = {r4.x, r4.y, r4.z, r4.w};
      float f[24];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float h = hs[i];
        const float u = safe_velocity(hus[i], h, G.dry);
        const float v = safe_velocity(hvs[i], h, G.dry);
        const bool a = ((m >> (8 * i)) & 0xffu) != 0;
        f[6 * i + 0] = a ? (h - nz.mean[0]) / nz.std[0] : 0.0f;
        f[6 * i + 1] = a ? (u - nz.mean[1]) / nz.std[1] : 0.0f;
        f[6 * i + 2] = a ? (v - nz.mean[2]) / nz.std[2] : 0.0f;
        f[6 * i + 3] = rain_channel;
        f[6 * i + 4] = ts[i];
        f[6 * i + 5] = rs[i];
      }
      float4* t = reinterpret_cast<float4*>(tile) + threadIdx.x * 6;
#pragma unroll
      for (int j = 0; j < 6; ++j) t[j] = make_float4(f[4 * j], f[4 * j + 1], f[4 * j + 2], f[4 * j + 3]);
    }
    __syncthreads();
    const int nf4 = 6 * (int)min<size_t>(kBlock, N4 - base);  // float4s of this tile
    float4* dst = reinterpret_cast<float4*>(out) + base * 6;
    for (int j = threadIdx.x; j < nf4; j += kBlock) dst[j] = reinterpret_cast<const float4*>(tile)[j];
    __syncthreads();
  }
}

struct Residual {
  const float* r;
  long long cs, ps;
  float relax;
  float lim[3];  // +inf: no clip
  __device__ __forceinline__ float at(size_t k, int c) const {
    return clip_to(r[(long long)k * ps + c * cs], lim[c]);
  }
};

// Pass 1: per-workgroup float64 sums of the masked physics depth and of the
// clipped provisional depth (coupled_forecast.py:97-105).
__global__ __launch_bounds__(kBlock) void k_corr_sums(Grid G, const float* __restrict__ H,
                                                      Residual R, double* __restrict__ part) {
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  double tgt = 0.0, prov = 0.0;
  for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < N;
       k += (size_t)gridDim.x * kBlock) {
    const float h = H[k];
    const bool a = G.mask[k] != 0;
    const float p = a ? fmaxf(h + R.relax * R.at(k, 0), 0.0f) : 0.0f;
    tgt += a ? (double)h : 0.0;
    prov += (double)p;
  }
  __shared__ double s_t[kWavesPerBlock], s_p[kWavesPerBlock];
  tgt = wave_sum(tgt);
  prov = wave_sum(prov);
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    s_t[w] = tgt;
    s_p[w] = prov;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int i = 0; i < kWavesPerBlock; ++i) {
      a += s_t[i];
      b += s_p[i];
    }
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// Pass 2 (one workgroup): totals and the rescale factor; out[0..3] =
// (target volume, provisional volume, scale, use_scaled) in cell units.
__global__ __launch_bounds__(kBlock) void k_corr_scale(const double* __restrict__ part, int nparts,
                                                       double* __restrict__ out) {
  double tgt = 0.0, prov = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kBlock) {
    tgt += part[2 * i];
    prov += part[2 * i + 1];
  }
  __shared__ double s_t[kWavesPerBlock], s_p[kWavesPerBlock];
  tgt = wave_sum(tgt);
  prov = wave_sum(prov);
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    s_t[w] = tgt;
    s_p[w] = prov;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int i = 0; i < kWavesPerBlock; ++i) {
      a += s_t[i];
      b += s_p[i];
    }
    const double eps = (double)FLT_EPSILON;  // jnp.finfo(float32).eps
    out[0] = a;
    out[1] = b;
    out[2] = b > eps ? a / b : 1.0;
    out[3] = a > eps ? 1.0 : 0.0;
  }
}

// Pass 3: corrected state (coupled_forecast.py:106-120).
// The corrected state may be the predicted one (each cell is read before it
// is written, by the same thread), so no pointer here is __restrict__.
__global__ __launch_bounds__(kBlock) void k_corr_apply(Grid G, const float* H, const float* HU,
                                                       const float* HV, Residual R,
                                                       const double* sc, float* Ho, float* HUo,
                                                       float* HVo) {
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  const double scale = sc[2];
  const bool use_scaled = sc[3] != 0.0;
  for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < N;
       k += (size_t)gridDim.x * kBlock) {
    const float h = H[k];
    const bool a = G.mask[k] != 0;
    const float u = safe_velocity(HU[k], h, G.dry);
    const float v = safe_velocity(HV[k], h, G.dry);
    const float p = a ? fmaxf(h + R.relax * R.at(k, 0), 0.0f) : 0.0f;
    const float ch = use_scaled ? (float)((double)p * scale) : h;
    const float cu = u + R.relax * R.at(k, 1);
    const float cv = v + R.relax * R.at(k, 2);
    const bool wet = a && ch > G.dry;
    Ho[k] = ch;
    HUo[k] = wet ? ch * cu : 0.0f;
    HVo[k] = wet ? ch * cv : 0.0f;
  }
}

// Four-cell form of pass 3: cells 4q..4q+3 per thread, 16-byte loads and
// stores, the mask as one 32-bit word (pass 1 gains nothing from it).  The host takes them only when the
// cell count is a multiple of 4, every array is 16-byte aligned and the
// residual is interleaved (cell stride 3, channel stride 1) or planar (cell
// stride 1, channel stride a multiple of 4).  Each cell's arithmetic is that
// of k_corr_apply.
template <bool INTERLEAVED>
__device__ __forceinline__ void residual4(const Residual& R, size_t q, float (&c)[3][4]) {
  const float4* r4 = reinterpret_cast<const float4*>(R.r);
  if (INTERLEAVED) {
    const float4 a = r4[3 * q], b = r4[3 * q + 1], d = r4[3 * q + 2];
    const float v[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, d.x, d.y, d.z, d.w};
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int ch = 0; ch < 3; ++ch) c[ch][i] = v[3 * i + ch];
  } else {
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
      const float4 a = r4[(size_t)ch * (size_t)(R.cs / 4) + q];
      c[ch][0] = a.x;
      c[ch][1] = a.y;
      c[ch][2] = a.z;
      c[ch][3] = a.w;
    }
  }
#pragma unroll
  for (int ch = 0; ch < 3; ++ch)
#pragma unroll
    for (int i = 0; i < 4; ++i) c[ch][i] = clip_to(c[ch][i], R.lim[ch]);
}

__device__ __forceinline__ float4 ld4(const float* p, size_t q) {
  return reinterpret_cast<const float4*>(p)[q];
}

// As k_corr_apply, the output may be the input: each thread reads its four
// cells before writing them, so no pointer is __restrict__.
template <bool INTERLEAVED>
__global__ __launch_bounds__(kBlock) void k_corr_apply4(Grid G, const float* H, const float* HU,
                                                        const float* HV, Residual R,
                                                        const double* sc, float* Ho, float* HUo,
                                                        float* HVo) {
  const size_t N4 = (size_t)G.nx * (size_t)G.ny / 4;
  const uint32_t* M = reinterpret_cast<const uint32_t*>(G.mask);
  const double scale = sc[2];
  const bool use_scaled = sc[3] != 0.0;
  for (size_t q = (size_t)blockIdx.x * kBlock + threadIdx.x; q < N4;
       q += (size_t)gridDim.x * kBlock) {
    const float4 h4 = ld4(H, q), hu4 = ld4(HU, q), hv4 = ld4(HV, q);
    const float hs[4] = {h4.x, h4.y, h4.z, h4.w};
    const float hus[4] = {hu4.x, hu4.y, hu4.z, hu4.w};
    const float hvs[4] = {hv4.x, hv4.y, hv4.z, hv4.w};
    const uint32_t m = M[q];
    float c[3][4];
    residual4<INTERLEAVED>(R, q, c);
    float oh[4], ou[4], ov[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float h = hs[i];
      const bool a = ((m >> (8 * i)) & 0xffu) != 0;
      const float u = safe_velocity(hus[i], h, G.dry);
      const float v = safe_velocity(hvs[i], h, G.dry);
      const float p = a ? fmaxf(h + R.relax * c[0][i], 0.0f) : 0.0f;
      const float ch = use_scaled ? (float)((double)p * scale) : h;
      const float cu = u + R.relax * c[1][i];
      const float cv = v + R.relax * c[2][i];
      const bool wet = a && ch > G.dry;
      oh[i] = ch;
      ou[i] = wet ? ch * cu : 0.0f;
      ov[i] = wet ? ch * cv : 0.0f;
    }
    reinterpret_cast<float4*>(Ho)[q] = make_float4(oh[0], oh[1], oh[2], oh[3]);
    reinterpret_cast<float4*>(HUo)[q] = make_float4(ou[0], ou[1], ou[2], ou[3]);
    reinterpret_cast<float4*>(HVo)[q] = make_float4(ov[0], ov[1], ov[2], ov[3]);
  }
}

// Feature staging of a batch of primitive states (comparison/v2_staging.py:
// 17-44): sample s, cell k, channel c of the (count, ny, nx, 3) NHWC state is
// state[(s*N + k)*3 + c]; the output channel c of (s, k) goes to
// out[s*ss + c*cs + k*ps].  blockIdx.y is the sample.
__global__ __launch_bounds__(kBlock) void k_stage(size_t N, const float* __restrict__ state,
                                                  const uint8_t* __restrict__ mask,
                                                  const float* __restrict__ terrain,
                                                  const float* __restrict__ rough, Norm nz,
                                                  const float* __restrict__ rain_ch,
                                                  float* __restrict__ out, long long ss,
                                                  long long cs, long long ps) {
  const size_t smp = blockIdx.y;
  const float rain = rain_ch[smp];
  const float* st = state + smp * N * 3;
  float* o0 = out + (long long)smp * ss;
  for (size_t base = (size_t)blockIdx.x * kBlock; base < N; base += (size_t)gridDim.x * kBlock) {
    const size_t k = base + threadIdx.x;
    float f[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (k < N) {
      // One 12-byte load per cell (a 96-bit load per lane, consecutive
      // across the wavefront) rather than three 4-byte loads at a 12-byte
      // stride.
      struct F3 {
        float d, u, v;
      };
      const F3 c3 = reinterpret_cast<const F3*>(st)[k];
      const bool a = mask[k] != 0;
      f[0] = a ? (c3.d - nz.mean[0]) / nz.std[0] : 0.0f;
      f[1] = a ? (c3.u - nz.mean[1]) / nz.std[1] : 0.0f;
      f[2] = a ? (c3.v - nz.mean[2]) / nz.std[2] : 0.0f;
      f[3] = rain;
      f[4] = terrain[k];
      f[5] = rough[k];
    }
    write_cells6(o0, base, N, f, cs, ps);
  }
}

// Four cells per thread of k_stage for NHWC output (channel stride 1, cell
// stride 6) when the cell count and the sample stride are multiples of 4 and
// every array is aligned: the (depth, u, v) of four cells as three 16-byte
// loads, the mask as one 32-bit word, and the block's 24 * kBlock output
// floats staged in LDS and written as 16-byte vectors.  Each cell's
// arithmetic is that of k_stage.  blockIdx.y is the sample.
__global__ __launch_bounds__(kBlock) void k_stage4(size_t N, const float* __restrict__ state,
                                                   const uint8_t* __restrict__ mask,
                                                   const float* __restrict__ terrain,
                                                   const float* __restrict__ rough, Norm nz,
                                                   const float* __restrict__ rain_ch,
                                                   float* __restrict__ out, long long ss) {
  __shared__ __align__(16) float tile[kBlock * 24];
  const size_t smp = blockIdx.y;
  const size_t N4 = N / 4;
  const float rain = rain_ch[smp];
  const float4* __restrict__ S4 = reinterpret_cast<const float4*>(state + smp * N * 3);
  const uint32_t* __restrict__ M = reinterpret_cast<const uint32_t*>(mask);
  const float4* __restrict__ T4 = reinterpret_cast<const float4*>(terrain);
  const float4* __restrict__ R4 = reinterpret_cast<const float4*>(rough);
  float4* __restrict__ o4 = reinterpret_cast<float4*>(out + (long long)smp * ss);
  for (size_t base = (size_t)blockIdx.x * kBlock; base < N4; base += (size_t)gridDim.x * kBlock) {
    const size_t q = base + threadIdx.x;
    if (q < N4) {
      const float4 a4 = S4[3 * q], b4 = S4[3 * q + 1], c4 = S4[3 * q + 2];
      const float sv[12] = {a4.x, a4.y, a4.z, a4.w, b4.x, b4.y, b4.z, b4.w, c4.x, c4.y, c4.z, c4.w};
      const float4 t4 = T4[q], r4 = R4[q];
      const float ts[4] = {t4.x, t4.y, t4.z, t4.w};
      const float rs[4] = {r4.x, r4.y, r4.z, r4.w};
      const uint32_t m = M[q];
      float f[24];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const bool a = ((m >> (8 * i)) & 0xffu) != 0;
        f[6 * i + 0] = a ? (sv[3 * i] - nz.mean[0]) / nz.std[0] : 0.0f;
        f[6 * i + 1] = a ? (sv[3 * i + 1] - nz.mean[1]) / nz.std[1] : 0.0f;
        f[6 * i + 2] = a ? (sv[3 * i + 2] - nz.mean[2]) / nz.std[2] : 0.0f;
        f[6 * i + 3] = rain;
        f[6 * i + 4] = ts[i];
        f[6 * i + 5] = rs[i];
      }
      float4* t = reinterpret_cast<float4*>(tile) + threadIdx.x * 6;
#pragma unroll
      for (int j = 0; j < 6; ++j) t[j] = make_float4(f[4 * j], f[4 * j + 1], f[4 * j + 2], f[4 * j + 3]);
    }
    __syncthreads();
    const int nf4 = 6 * (int)min<size_t>(kBlock, N4 - base);  // float4s of this tile
    float4* dst = o4 + base * 6;
    for (int j = threadIdx.x; j < nf4; j += kBlock) dst[j] = reinterpret_cast<const float4*>(tile)[j];
    __syncthreads();
  }
}

// max(depth + residual[..., 0], 0) for a batch (v2_staging.py:47-52); residual
// channel 0 of (s, k) is residual[s*ss + k*ps].
__global__ __launch_bounds__(kBlock) void k_correct_depth(size_t N, const float* __restrict__ depth,
                                                          const float* __restrict__ residual,
                                                          long long ss, long long ps,
                                                          float* __restrict__ out) {
  const size_t smp = blockIdx.y;
  for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < N;
       k += (size_t)gridDim.x * kBlock)
    out[smp * N + k] = fmaxf(depth[smp * N + k] + residual[(long long)smp * ss + (long long)k * ps],
                             0.0f);
}

int grid_blocks(size_t N) {
  return (int)std::max<size_t>(1, std::min<size_t>((N + kBlock - 1) / kBlock, 4096));
}

// Elementwise passes: one item per thread (their grid-stride loops then run
// once), so every load of the pass is in flight at once; a grid capped at a
// few thousand workgroups made each thread walk several items in turn.
int stream_blocks(size_t items) {
  return (int)std::max<size_t>(1, std::min<size_t>((items + kBlock - 1) / kBlock, 0x7FFFFFFF));
}

bool valid(const SweState& s) { return s.h && s.hu && s.hv; }

// apply_residual_correction on `scratch` (device: 4 result doubles, then 2 x nb
// partials).  Shared by the entry point that keeps its scratch in the
// context and the one that uses the file-scope array.
int correct_with(const Grid& G, SweState predicted, const float* residual, int64_t channel_stride,
                 int64_t cell_stride, float relaxation, const float* clip_limits,
                 SweState corrected, double* volumes, double* scratch, int nb, hipStream_t s) {
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  Residual R;
  R.r = residual;
  R.cs = channel_stride;
  R.ps = cell_stride;
  R.relax = relaxation;
  for (int i = 0; i < 3; ++i) R.lim[i] = clip_limits ? clip_limits[i] : HUGE_VALF;
  double* part = scratch + 4;
  hipLaunchKernelGGL(k_corr_sums, dim3(nb), dim3(kBlock), 0, s, G, predicted.h, R, part);
  hipLaunchKernelGGL(k_corr_scale, dim3(1), dim3(kBlock), 0, s, part, nb, scratch);
  // four-cell apply: N % 4 == 0, 16-byte arrays, a 4-byte mask, and an
  // interleaved (3, 1) or planar (1, 4k) residual
  auto al = [](const void* p, uintptr_t a) { return ((uintptr_t)p & (a - 1)) == 0; };
  const bool interleaved = channel_stride == 1 && cell_stride == 3;
  const bool planar = cell_stride == 1 && channel_stride % 4 == 0;
  const bool vec = N % 4 == 0 && (interleaved || planar) && al(residual, 16) && al(G.mask, 4) &&
                   al(predicted.h, 16) && al(predicted.hu, 16) && al(predicted.hv, 16) &&
                   al(corrected.h, 16) && al(corrected.hu, 16) && al(corrected.hv, 16);
  if (vec && interleaved)
    hipLaunchKernelGGL(k_corr_apply4<true>, dim3(stream_blocks(N / 4)), dim3(kBlock), 0, s, G,
                       predicted.h, predicted.hu, predicted.hv, R, scratch, corrected.h,
                       corrected.hu, corrected.hv);
  else if (vec)
    hipLaunchKernelGGL(k_corr_apply4<false>, dim3(stream_blocks(N / 4)), dim3(kBlock), 0, s, G,
                       predicted.h, predicted.hu, predicted.hv, R, scratch, corrected.h,
                       corrected.hu, corrected.hv);
  else
    hipLaunchKernelGGL(k_corr_apply, dim3(stream_blocks(N)), dim3(kBlock), 0, s, G, predicted.h,
                       predicted.hu, predicted.hv, R, scratch, corrected.h, corrected.hu,
                       corrected.hv);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(SWE_EHIP, std::string("residual correction: ") + hipGetErrorString(e));
  if (volumes) {
    e = hipMemcpyAsync(volumes, scratch, 4 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
      return set_error(SWE_EHIP, std::string("residual correction volumes: ") + hipGetErrorString(e));
  }
  return SWE_OK;
}

}  // namespace

// Scratch of swe_apply_residual_correction_noalloc: 4 result doubles and the
// 2 x kNoAllocParts per-workgroup volume partials, in device memory that the
// code object itself holds (each device that loads the library keeps its own
// copy, at its own address, until the process ends).  In host code the name
// `scratch` is only a stand-in for the symbol: kernels get the device address
// from hipGetSymbolAddress, kept per device in g_scratch_dev.
constexpr int kNoAllocParts = 2048;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[4 + 2 * kNoAllocParts];
namespace {
double* g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;
}  // namespace

int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    return set_error(SWE_ENODEV, "noalloc scratch: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[device]) return SWE_OK;
  double* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return set_error(SWE_EHIP, std::string("noalloc scratch: hipGetSymbolAddress: ") +
                                   hipGetErrorString(e));
  g_scratch_dev[device] = ptr;
  return SWE_OK;
}

namespace {
double* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_dev[device];
}
}  // namespace
}  // namespace swe

using namespace swe;

extern "C" {

int swe_network_features(SweContext* c, SweState state, const float* terrain_std,
                         const float* roughness_std, const float* state_mean,
                         const float* state_std, double rainfall_scale_mm_hr, double time_s,
                         float* features, int64_t channel_stride, int64_t cell_stride,
                         void* stream) {
  if (!c || !valid(state) || !terrain_std || !roughness_std || !state_mean || !state_std ||
      !features)
    return set_error(SWE_EINVAL, "swe_network_features: null argument");
  const Grid& G = *context_grid(c);
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  Norm nz;
  for (int i = 0; i < 3; ++i) {
    nz.mean[i] = state_mean[i];
    nz.std[i] = state_std[i];
  }
  // coupled_forecast.py:65-69 in float32: rate * 3.6e6, then / scale.
  const float mm_hr = (float)context_rate(c, time_s) * 3.6e6f;
  const float rain_channel = mm_hr / (float)rainfall_scale_mm_hr;
  auto al = [](const void* p, uintptr_t a) { return ((uintptr_t)p & (a - 1)) == 0; };
  if (channel_stride == 1 && cell_stride == 6 && N % 4 == 0 && al(features, 16) &&
      al(G.mask, 4) && al(state.h, 16) && al(state.hu, 16) && al(state.hv, 16) &&
      al(terrain_std, 16) && al(roughness_std, 16))
    hipLaunchKernelGGL(k_features4, dim3(stream_blocks(N / 4)), dim3(kBlock), 0,
                       (hipStream_t)stream, G, state.h, state.hu, state.hv, terrain_std,
                       roughness_std, nz, rain_channel, features);
  else
    hipLaunchKernelGGL(k_features, dim3(stream_blocks(N)), dim3(kBlock), 0, (hipStream_t)stream, G,
                       state.h, state.hu, state.hv, terrain_std, roughness_std, nz, rain_channel,
                       features, (long long)channel_stride, (long long)cell_stride);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(SWE_EHIP, std::string("k_features: ") + hipGetErrorString(e));
  return SWE_OK;
}

int swe_apply_residual_correction(SweContext* c, SweState predicted, const float* residual,
                                  int64_t channel_stride, int64_t cell_stride, float relaxation,
                                  const float* clip_limits, SweState corrected, double* volumes,
                                  void* stream) {
  if (!c || !valid(predicted) || !valid(corrected) || !residual)
    return set_error(SWE_EINVAL, "swe_apply_residual_correction: null argument");
  const Grid& G = *context_grid(c);
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  const int nb = std::min(grid_blocks(N), context_reduce_blocks(c));
  double* scratch = context_scratch(c, 2 * (size_t)nb + 4);
  if (!scratch) return set_error(SWE_EHIP, "swe_apply_residual_correction: scratch allocation failed");
  return correct_with(G, predicted, residual, channel_stride, cell_stride, relaxation, clip_limits,
                      corrected, volumes, scratch, nb, (hipStream_t)stream);
}

int swe_apply_residual_correction_noalloc(SweContext* c, SweState predicted,
                                          const float* residual, int64_t channel_stride,
                                          int64_t cell_stride, float relaxation,
                                          const float* clip_limits, SweState corrected,
                                          double* volumes, void* stream) {
  if (!c || !valid(predicted) || !valid(corrected) || !residual)
    return set_error(SWE_EINVAL, "swe_apply_residual_correction_noalloc: null argument");
  const Grid& G = *context_grid(c);
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  // The same workgroup count as swe_apply_residual_correction, so both sum
  // the partials in the same order and give the same bits.
  const int nb = std::min(grid_blocks(N), context_reduce_blocks(c));
  if (nb > kNoAllocParts)
    return set_error(SWE_EINVAL, "swe_apply_residual_correction_noalloc: the grid needs " +
                                     std::to_string(nb) + " volume partials, more than the " +
                                     std::to_string(kNoAllocParts) + " the static scratch holds");
  double* ptr = noalloc_scratch(context_device(c));
  if (!ptr)
    return set_error(SWE_EHIP, "swe_apply_residual_correction_noalloc: the static scratch of "
                               "this device was not resolved (swe_create does it)");
  return correct_with(G, predicted, residual, channel_stride, cell_stride, relaxation, clip_limits,
                      corrected, volumes, ptr, nb, (hipStream_t)stream);
}

int swe_stage_features(int32_t count, int32_t ny, int32_t nx, const float* state_nhwc,
                       const uint8_t* domain_mask, const float* terrain_std,
                       const float* roughness_std, const float* state_mean, const float* state_std,
                       const float* rain_channel, float* features, int64_t sample_stride,
                       int64_t channel_stride, int64_t cell_stride, void* stream) {
  if (count < 0 || count > 65535 || ny < 1 || nx < 1)
    return set_error(SWE_EINVAL, "swe_stage_features: bad sizes");
  if (count == 0) return SWE_OK;
  if (!state_nhwc || !domain_mask || !terrain_std || !roughness_std || !state_mean || !state_std ||
      !rain_channel || !features)
    return set_error(SWE_EINVAL, "swe_stage_features: null argument");
  const size_t N = (size_t)nx * (size_t)ny;
  Norm nz;
  for (int i = 0; i < 3; ++i) {
    nz.mean[i] = state_mean[i];
    nz.std[i] = state_std[i];
  }
  auto al = [](const void* p, uintptr_t a) { return ((uintptr_t)p & (a - 1)) == 0; };
  if (channel_stride == 1 && cell_stride == 6 && N % 4 == 0 && sample_stride % 4 == 0 &&
      al(features, 16) && al(state_nhwc, 16) && al(domain_mask, 4) && al(terrain_std, 16) &&
      al(roughness_std, 16)) {
    const int nb = std::min(stream_blocks(N / 4), 65535);
    hipLaunchKernelGGL(k_stage4, dim3(nb, count), dim3(kBlock), 0, (hipStream_t)stream, N,
                       state_nhwc, domain_mask, terrain_std, roughness_std, nz, rain_channel,
                       features, (long long)sample_stride);
  } else {
    const int nb = std::min(stream_blocks(N), 65535);
    hipLaunchKernelGGL(k_stage, dim3(nb, count), dim3(kBlock), 0, (hipStream_t)stream, N,
                       state_nhwc, domain_mask, terrain_std, roughness_std, nz, rain_channel,
                       features, (long long)sample_stride, (long long)channel_stride,
                       (long long)cell_stride);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(SWE_EHIP, std::string("k_stage: ") + hipGetErrorString(e));
  return SWE_OK;
}

int swe_correct_depth(int32_t count, int32_t ny, int32_t nx, const float* depth,
                      const float* residual, int64_t sample_stride, int64_t cell_stride,
                      float* out, void* stream) {
  if (count < 0 || count > 65535 || ny < 1 || nx < 1)
    return set_error(SWE_EINVAL, "swe_correct_depth: bad sizes");
  if (count == 0) return SWE_OK;
  if (!depth || !residual || !out) return set_error(SWE_EINVAL, "swe_correct_depth: null argument");
  const size_t N = (size_t)nx * (size_t)ny;
  // (a small pass, 12 B per cell: measured faster with each thread walking
  // 16 cells than with one cell per thread, 30.7 against 34.8 us at 4096^2)
  const int nb = std::min(grid_blocks(N), 1024);
  hipLaunchKernelGGL(k_correct_depth, dim3(nb, count), dim3(kBlock), 0, (hipStream_t)stream, N,
                     depth, residual, (long long)sample_stride, (long long)cell_stride, out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(SWE_EHIP, std::string("k_correct_depth: ") + hipGetErrorString(e));
  return SWE_OK;
}

}  // extern "C"
