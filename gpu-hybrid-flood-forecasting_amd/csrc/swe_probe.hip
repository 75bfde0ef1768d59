// Streaming probes: the achievable HBM bandwidth of this device for the
// access patterns of the step, measured by the benchmark beside the step
// itself (no counterpart in the reference, which never measured a device).
//   mode 0  copy: n floats read, n written (16-byte accesses)
//   mode 1  read: n floats read (16-byte accesses), one float per
//           workgroup written
//   mode 2  the packed device loop's streams without its arithmetic: per
//           cell, 12 bytes of packed state and 8 bytes of packed static
//           fields read and 12 bytes of packed state written, one cell per
//           lane, as the step's 96- and 64-bit accesses
//   mode 3  as mode 2 with the class-indexed static record: 4 bytes of bed
//           and 1 byte of class per cell, in blocks of 64 cells (64 beds,
//           then 64 classes: 320 bytes), src2 holding ceil(n / 64) blocks
//   mode 4  clock sample: one thread writes the shader clock counter and the
//           constant-rate real-time counter (two uint64) to dst slot n - 1;
//           two samples on the stream around other work give the average
//           shader clock over it
// One element per thread; the caller times the launches with events.

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/swe_hip.h"
#include "swe_host.hpp"

namespace swe {
namespace {

constexpr int kProbeBlock = 256;

__global__ __launch_bounds__(kProbeBlock) void k_probe_copy(const float4* __restrict__ src,
                                                            float4* __restrict__ dst, size_t n4) {
  const size_t k = (size_t)blockIdx.x * kProbeBlock + threadIdx.x;
  if (k < n4) dst[k] = src[k];
}

__global__ __launch_bounds__(kProbeBlock) void k_probe_read(const float4* __restrict__ src,
                                                            float* __restrict__ dst, size_t n4) {
  const size_t k = (size_t)blockIdx.x * kProbeBlock + threadIdx.x;
  float4 v = k < n4 ? src[k] : make_float4(0.f, 0.f, 0.f, 0.f);
  float x = (v.x + v.y) + (v.z + v.w);
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  __shared__ float s[kProbeBlock / 64];
  if ((threadIdx.x & 63) == 0) s[threadIdx.x / 64] = x;
  __syncthreads();
  if (threadIdx.x == 0) dst[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
}

struct Cell3 {
  float a, b, c;
};
struct Cell2 {
  float a, b;
};

__global__ __launch_bounds__(kProbeBlock) void k_probe_stepmix(const Cell3* __restrict__ st,
                                                               const Cell2* __restrict__ sx,
                                                               Cell3* __restrict__ out, size_t n) {
  const size_t k = (size_t)blockIdx.x * kProbeBlock + threadIdx.x;
  if (k < n) {
    const Cell3 u = st[k];
    const Cell2 f = sx[k];
    // A data dependence on every loaded word, so no load is dropped.
    out[k] = Cell3{u.a + f.a, u.b + f.b, u.c};
  }
}

__global__ __launch_bounds__(kProbeBlock) void k_probe_stepcls(const Cell3* __restrict__ st,
                                                               const unsigned char* __restrict__ sx,
                                                               Cell3* __restrict__ out, size_t n) {
  const size_t k = (size_t)blockIdx.x * kProbeBlock + threadIdx.x;
  if (k < n) {
    const Cell3 u = st[k];
    const unsigned char* blk = sx + k / 64 * 320;
    const float b = reinterpret_cast<const float*>(blk)[k % 64];
    const unsigned c = blk[256 + k % 64];
    out[k] = Cell3{u.a + b, u.b + (float)c, u.c};
  }
}

__global__ __launch_bounds__(64) void k_probe_clock(unsigned long long* __restrict__ out) {
  if (threadIdx.x == 0) {
    out[0] = (unsigned long long)clock64();
    out[1] = (unsigned long long)wall_clock64();
  }
}

}  // namespace
}  // namespace swe

using namespace swe;

extern "C" int swe_stream_probe(int32_t mode, const void* src, const void* src2, void* dst,
                                int64_t n, void* stream) {
  if (n <= 0 || !dst || mode < 0 || mode > 4 || (mode < 4 && !src) ||
      ((mode == 2 || mode == 3) && !src2))
    return set_error(SWE_EINVAL, "swe_stream_probe: bad arguments");
  if (mode == 4) {
    if ((uintptr_t)dst & 7) return set_error(SWE_EINVAL, "swe_stream_probe: misaligned buffers");
    hipLaunchKernelGGL(k_probe_clock, dim3(1), dim3(64), 0, (hipStream_t)stream,
                       static_cast<unsigned long long*>(dst) + 2 * (size_t)(n - 1));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(SWE_EHIP, std::string("swe_stream_probe: ") + hipGetErrorString(e));
    return SWE_OK;
  }
  if ((mode < 2 && (n % 4 != 0 || ((uintptr_t)src & 15) || ((uintptr_t)dst & 15))) ||
      (mode >= 2 && (((uintptr_t)src & 3) || ((uintptr_t)src2 & 7) || ((uintptr_t)dst & 3))))
    return set_error(SWE_EINVAL, "swe_stream_probe: misaligned buffers or n not a multiple of 4");
  hipStream_t s = (hipStream_t)stream;
  const size_t items = mode < 2 ? (size_t)n / 4 : (size_t)n;
  const size_t blocks = (items + kProbeBlock - 1) / kProbeBlock;
  if (blocks > 0x7FFFFFFFull) return set_error(SWE_EINVAL, "swe_stream_probe: n too large");
  const dim3 grid((unsigned)blocks), block(kProbeBlock);
  if (mode == 0)
    hipLaunchKernelGGL(k_probe_copy, grid, block, 0, s, (const float4*)src, (float4*)dst, items);
  else if (mode == 1)
    hipLaunchKernelGGL(k_probe_read, grid, block, 0, s, (const float4*)src, (float*)dst, items);
  else if (mode == 2)
    hipLaunchKernelGGL(k_probe_stepmix, grid, block, 0, s, (const Cell3*)src, (const Cell2*)src2,
                       (Cell3*)dst, items);
  else
    hipLaunchKernelGGL(k_probe_stepcls, grid, block, 0, s, (const Cell3*)src,
                       (const unsigned char*)src2, (Cell3*)dst, items);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(SWE_EHIP, std::string("swe_stream_probe: ") + hipGetErrorString(e));
  return SWE_OK;
}
