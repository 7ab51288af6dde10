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

}  // namespace simr
