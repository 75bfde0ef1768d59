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
  const float nan = __builtin_nanf("");
  for (size_t cell = (size_t)blockIdx.x * kBlock + threadIdx.x; cell < N;
       cell += (size_t)gridDim.x * kBlock) {
    const size_t k = base + cell;
    const float h = H[k];
    const bool wet = h > G.dry;
    // np.divide(q, h, where=wet) into zeros: a plain quotient, no max(h, dry)
    const float u = wet ? HU[k] / h : 0.0f;
    const float v = wet ? HV[k] / h : 0.0f;
    const bool a = G.mask[cell] != 0;
    depth[k] = a ? h : nan;
    vx[k] = a ? u : nan;
    vy[k] = a ? v : nan;
    speed[k] = a ? hypotf(u, v) : nan;
  }
}

__global__ __launch_bounds__(kBlock) void k_elevation(Grid G, float* __restrict__ elev) {
  const size_t N = (size_t)G.nx * (size_t)G.ny;
  const float nan = __builtin_nanf("");
  for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < N;
       k += (size_t)gridDim.x * kBlock)
    elev[k] = G.mask[k] != 0 ? G.bed[k] : nan;
}

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
  if (n_times > 0) {
    const int nb = (int)std::max<size_t>(1, std::min<size_t>((N + kBlock - 1) / kBlock, 2048));
    hipLaunchKernelGGL(k_primitive, dim3(nb, n_times), dim3(kBlock), 0, s, G, h, hu, hv, depth,
                       x_velocity, y_velocity, velocity);
  }
  if (elevation) {
    const int nb = (int)std::max<size_t>(1, std::min<size_t>((N + kBlock - 1) / kBlock, 4096));
    hipLaunchKernelGGL(k_elevation, dim3(nb), dim3(kBlock), 0, s, G, elevation);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(SWE_EHIP, std::string("swe_conserved_to_primitive: ") + hipGetErrorString(e));
  return SWE_OK;
}
