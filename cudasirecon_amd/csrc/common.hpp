// Shared host/device helpers for the MI355X SIM reconstruction engine.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "errors.hpp"

namespace simr {

#define SIMR_HIP(call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::simr::HipError(std::string(#call " failed: ") + hipGetErrorString(e_) +    \
                             " (" __FILE__ ":" + std::to_string(__LINE__) + ")");        \
  } while (0)

#define SIMR_LAUNCH_CHECK() SIMR_HIP(hipGetLastError())

#define SIMR_REQUIRE(cond, msg)                  \
  do {                                           \
    if (!(cond)) throw ::simr::ArgError(msg);    \
  } while (0)

__host__ __device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

// The dispatcher hands workgroups to the 8 XCDs (each with its own L2) in
// turn by linear id (x fastest).  xcd_tile renumbers a 2D grid so that each
// XCD takes one contiguous run of tiles (x fastest) in dispatch order: tiles
// next to each other in x, whose unaligned 128-byte row segments share
// cache lines, then meet in one L2 at about the same time.  A bijection on
// the grid for any size.
__device__ __forceinline__ void xcd_tile(unsigned& tx, unsigned& ty) {
  const unsigned total = gridDim.x * gridDim.y;
  const unsigned b = blockIdx.y * gridDim.x + blockIdx.x;
  const unsigned x = b % 8u, j = b / 8u, q = total / 8u, r = total % 8u;
  const unsigned t = x * q + (x < r ? x : r) + j;
  tx = t % gridDim.x;
  ty = t / gridDim.x;
}
// The same over a 3D grid (x fastest, then y, then z).
__device__ __forceinline__ void xcd_tile3(unsigned& tx, unsigned& ty, unsigned& tz) {
  const unsigned nxy = gridDim.x * gridDim.y;
  const unsigned total = nxy * gridDim.z;
  const unsigned b = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  const unsigned x = b % 8u, j = b / 8u, q = total / 8u, r = total % 8u;
  const unsigned t = x * q + (x < r ? x : r) + j;
  tx = t % gridDim.x;
  ty = (t % nxy) / gridDim.x;
  tz = t / nxy;
}

}  // namespace simr
