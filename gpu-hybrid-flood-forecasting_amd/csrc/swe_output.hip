// Conserved -> primitive conversion of stacked output states into the
// reference's common forecast schema (jax_solver/shallow_water_2d.py:309-350,
// structured_io.py:160-211): depth, x/y velocity hu/h and hv/h where
// h > dry (0 elsewhere), speed hypot(u, v), and NaN for all four outside the
// domain mask; elevation is the bed with NaN outside the mask.  One pass over
// the (T, ny, nx) cubes, written to caller-owned device arrays.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "../../include/swe_hip.h"
#include "swe_device.hpp"
#include "swe_host.hpp"
#include "swe_loop.hpp"

namespace swe {
namespace {

// One cell of the schema: depth, velocities and speed, NaN outside the mask.
struct Prim {
  float d, u, v, s;
};

__device__ __forceinline__ Prim primitive_cell(float h, float hu, float hv, bool a, float dry) {
  const float nan = __builtin_nanf("");
  const bool wet = h > dry;
  // np.divide(q, h, where=wet) into zeros: a plain quotient, no max(h, dry)
  const float u = wet ? hu / h : 0.0f;
  const float v = wet ? hv / h : 0.0f;
  return {a ? h : nan, a ? u : nan, a ? v : nan, a ? hypotf(u, v) : nan};
}

// blockIdx.y is the output slot; the x grid strides over the cells.
__global__ __launch_bounds__(kBlock) void k_primitive(Grid G, const float* __restrict__ H,
                                                      const float* __restrict__ HU,
                                                      const float* __restrict__ HV,
                                                      float* __restrict__ depth,
                                                      float* __restrict__ vx,
                                                      float* __restrict__ vy,
                                                      float* __restrict__ speed) {
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  const size_t base = (size_t)blockIdx.y * N;
  for (size_t cell = (size_t)blockIdx.x * kBlock + threadIdx.x; cell < N;
       cell += (size_t)gridDim.x * kBlock) {
    const size_t k = base + cell;
    const Prim p = primitive_cell(H[k], HU[k], HV[k], G.mask[cell] != 0, G.dry);
    depth[k] = p.d;
    vx[k] = p.u;
    vy[k] = p.v;
    speed[k] = p.s;
  }
}

// Four cells per thread: 16-byte loads and stores, the mask as one 32-bit
// word.  The host takes this form only when the cell count is a multiple of 4
// and every array is aligned for it (the output slots then stay aligned too).
__global__ __launch_bounds__(kBlock) void k_primitive4(Grid G, const float4* __restrict__ H,
                                                       const float4* __restrict__ HU,
                                                       const float4* __restrict__ HV,
                                                       float4* __restrict__ depth,
                                                       float4* __restrict__ vx,
                                                       float4* __restrict__ vy,
                                                       float4* __restrict__ speed) {
  const size_t N4 = (size_t)G.nx * (size_t)G.ny / 4;
  const size_t base = (size_t)blockIdx.y * N4;
  const uint32_t* __restrict__ M = reinterpret_cast<const uint32_t*>(G.mask);
  for (size_t q = (size_t)blockIdx.x * kBlock + threadIdx.x; q < N4;
       q += (size_t)gridDim.x * kBlock) {
    const size_t k = base + q;
    const float4 h = H[k], hu = HU[k], hv = HV[k];
    const uint32_t m = M[q];
    const Prim p0 = primitive_cell(h.x, hu.x, hv.x, (m & 0xffu) != 0, G.dry);
    const Prim p1 = primitive_cell(h.y, hu.y, hv.y, (m & 0xff00u) != 0, G.dry);
    const Prim p2 = primitive_cell(h.z, hu.z, hv.z, (m & 0xff0000u) != 0, G.dry);
    const Prim p3 = primitive_cell(h.w, hu.w, hv.w, (m & 0xff000000u) != 0, G.dry);
    depth[k] = make_float4(p0.d, p1.d, p2.d, p3.d);
    vx[k] = make_float4(p0.u, p1.u, p2.u, p3.u);
    vy[k] = make_float4(p0.v, p1.v, p2.v, p3.v);
    speed[k] = make_float4(p0.s, p1.s, p2.s, p3.s);
  }
}

__global__ __launch_bounds__(kBlock) void k_elevation(Grid G, float* __restrict__ elev) {
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  const float nan = __builtin_nanf("");
  for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < N;
       k += (size_t)gridDim.x * kBlock)
    elev[k] = G.mask[k] != 0 ? G.bed[k] : nan;
}

__global__ __launch_bounds__(kBlock) void k_elevation4(Grid G, float4* __restrict__ elev) {
  const size_t N4 = (size_t)G.nx * (size_t)G.ny / 4;
  const float nan = __builtin_nanf("");
  const float4* __restrict__ B = reinterpret_cast<const float4*>(G.bed);
  const uint32_t* __restrict__ M = reinterpret_cast<const uint32_t*>(G.mask);
  for (size_t q = (size_t)blockIdx.x * kBlock + threadIdx.x; q < N4;
       q += (size_t)gridDim.x * kBlock) {
    const float4 b = B[q];
    const uint32_t m = M[q];
    elev[q] = make_float4((m & 0xffu) ? b.x : nan, (m & 0xff00u) ? b.y : nan,
                          (m & 0xff0000u) ? b.z : nan, (m & 0xff000000u) ? b.w : nan);
  }
}

bool aligned(const void* p, uintptr_t a) { return ((uintptr_t)p & (a - 1)) == 0; }

}  // namespace
}  // namespace swe

using namespace swe;

extern "C" int swe_conserved_to_primitive(SweContext* c, const float* h, const float* hu,
                                          const float* hv, int32_t n_times, float* depth,
                                          float* x_velocity, float* y_velocity, float* velocity,
                                          float* elevation, void* stream) {
  if (!c || n_times < 0 || (n_times > 0 && (!h || !hu || !hv || !depth || !x_velocity ||
                                            !y_velocity || !velocity)))
    return set_error(SWE_EINVAL, "swe_conserved_to_primitive: null argument");
  const Grid& G = *context_grid(c);
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  hipStream_t s = (hipStream_t)stream;
  if (n_times > 65535) return set_error(SWE_EINVAL, "swe_conserved_to_primitive: n_times > 65535");
  // the four-cell forms need N % 4 == 0 (so every output slot starts aligned)
  // and 16-byte arrays with a 4-byte mask
  const bool vec_grid = N % 4 == 0 && aligned(G.mask, 4);
  if (n_times > 0) {
    if (vec_grid && aligned(h, 16) && aligned(hu, 16) && aligned(hv, 16) && aligned(depth, 16) &&
        aligned(x_velocity, 16) && aligned(y_velocity, 16) && aligned(velocity, 16)) {
      const int nb = (int)std::max<size_t>(1, std::min<size_t>((N / 4 + kBlock - 1) / kBlock, 65535));
      hipLaunchKernelGGL(k_primitive4, dim3(nb, n_times), dim3(kBlock), 0, s, G,
                         (const float4*)h, (const float4*)hu, (const float4*)hv, (float4*)depth,
                         (float4*)x_velocity, (float4*)y_velocity, (float4*)velocity);
    } else {
      const int nb = (int)std::max<size_t>(1, std::min<size_t>((N + kBlock - 1) / kBlock, 65535));
      hipLaunchKernelGGL(k_primitive, dim3(nb, n_times), dim3(kBlock), 0, s, G, h, hu, hv, depth,
                         x_velocity, y_velocity, velocity);
    }
  }
  if (elevation) {
    if (vec_grid && aligned(G.bed, 16) && aligned(elevation, 16)) {
      const int nb = (int)std::max<size_t>(1, std::min<size_t>((N / 4 + kBlock - 1) / kBlock, 65535));
      hipLaunchKernelGGL(k_elevation4, dim3(nb), dim3(kBlock), 0, s, G, (float4*)elevation);
    } else {
      const int nb = (int)std::max<size_t>(1, std::min<size_t>((N + kBlock - 1) / kBlock, 65535));
      hipLaunchKernelGGL(k_elevation, dim3(nb), dim3(kBlock), 0, s, G, elevation);
    }
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(SWE_EHIP, std::string("swe_conserved_to_primitive: ") + hipGetErrorString(e));
  return SWE_OK;
}
