// Device helpers shared by the stage kernels.
#pragma once
#include "common.hpp"
#include "stages.hpp"

namespace simr {

// Bilinear (kr, kz) lookup in a rotationally averaged OTF table, following
// dev_otfinterpolate (src/cudaSirecon/gpuFunctionsImpl.cu:1677-1715): kz wraps
// (negative kz index + nzotf; the last kz row interpolates towards row 0).
// Indices are clamped to the table, so a support wider than the table reads
// the table edge instead of memory past it.
__device__ __forceinline__ float2 otf_interp(const OtfTab& o, float kx, float ky, int kz, float kzscale) {
#pragma clang fp contract(off)
  const int nzotf = o.nzotf;
  float krindex = sqrtf(kx * kx + ky * ky) / o.dkrotf;
  float kzindex = (float)kz * kzscale;
  if (kzindex < 0) kzindex += (float)nzotf;
  int irindex = (int)floorf(krindex);
  int izindex = (int)floorf(kzindex);
  float ar = krindex - (float)irindex;
  float az = kzindex - (float)izindex;
  if (irindex > o.nxotf - 2) {
    irindex = o.nxotf - 2;
    ar = 1.f;
  }
  if (irindex < 0) {
    irindex = 0;
    ar = 0.f;
  }
  if (izindex > nzotf - 1) izindex = nzotf - 1;
  if (izindex < 0) izindex = 0;
  const int iz1 = (izindex == nzotf - 1) ? 0 : izindex + 1;
  const float2 a00 = o.t[irindex * nzotf + izindex];
  const float2 a01 = o.t[irindex * nzotf + iz1];
  const float2 a10 = o.t[(irindex + 1) * nzotf + izindex];
  const float2 a11 = o.t[(irindex + 1) * nzotf + iz1];
  float2 v;
  v.x = (1 - ar) * (a00.x * (1 - az) + a01.x * az) + ar * (a10.x * (1 - az) + a11.x * az);
  v.y = (1 - ar) * (a00.y * (1 - az) + a01.y * az) + ar * (a10.y * (1 - az) + a11.y * az);
  return v;
}

// Radial profile of one RA OTF order at a fixed kz: prof[i] is the kz-interpolated
// value at kr index i, computed with exactly the bracketed kz term of
// otf_interp, so that profile_lookup(prof, ...) == otf_interp(o, ..., kz, ...).
__device__ __forceinline__ float2 otf_profile_entry(const OtfTab& o, int i, int kz, float kzscale) {
#pragma clang fp contract(off)
  const int nzotf = o.nzotf;
  float kzindex = (float)kz * kzscale;
  if (kzindex < 0) kzindex += (float)nzotf;
  int izindex = (int)floorf(kzindex);
  const float az = kzindex - (float)izindex;
  if (izindex > nzotf - 1) izindex = nzotf - 1;
  if (izindex < 0) izindex = 0;
  const int iz1 = (izindex == nzotf - 1) ? 0 : izindex + 1;
  const float2 a0 = o.t[i * nzotf + izindex];
  const float2 a1 = o.t[i * nzotf + iz1];
  return make_float2(a0.x * (1 - az) + a1.x * az, a0.y * (1 - az) + a1.y * az);
}

__device__ __forceinline__ float2 profile_lookup(const float2* prof, int nxotf, float dkrotf, float kx, float ky) {
#pragma clang fp contract(off)
  const float krindex = sqrtf(kx * kx + ky * ky) / dkrotf;
  int irindex = (int)floorf(krindex);
  float ar = krindex - (float)irindex;
  if (irindex > nxotf - 2) {
    irindex = nxotf - 2;
    ar = 1.f;
  }
  if (irindex < 0) {
    irindex = 0;
    ar = 0.f;
  }
  const float2 t0 = prof[irindex], t1 = prof[irindex + 1];
  return make_float2((1 - ar) * t0.x + ar * t1.x, (1 - ar) * t0.y + ar * t1.y);
}

// Trilinear lookup in a full (not rotationally averaged) OTF table at
// lateral frequency (kx, ky) (1/um) and axial plane index kz, the
// extension of dev_otfinterpolate's (kr, kz) scheme to (kx, ky, kz): the
// table holds kx >= 0, so kx < 0 reads the Hermitian mirror conj(OTF(-k))
// of a real PSF's OTF; ky and kz wrap (FFT order, the last row
// interpolating towards row 0, as kz does in the RA table); kx clamps at
// the table's edge like kr.
__device__ __forceinline__ float2 otf3d_interp(const OtfTab& o, float kx, float ky, int kz, float kzscale) {
#pragma clang fp contract(off)
  const bool cj = kx < 0.f;
  if (cj) {
    kx = -kx;
    ky = -ky;
    kz = -kz;
  }
  const float fx = kx / o.dkrotf;
  float fy = ky / o.dkrotf;
  if (fy < 0) fy += (float)o.nyotf;
  float fz = (float)kz * kzscale;
  if (fz < 0) fz += (float)o.nzotf;
  int ix = (int)floorf(fx), iy = (int)floorf(fy), iz = (int)floorf(fz);
  float ax = fx - (float)ix, ay = fy - (float)iy, az = fz - (float)iz;
  if (ix > o.nxotf - 2) {
    ix = o.nxotf - 2;
    ax = 1.f;
  }
  if (ix < 0) {
    ix = 0;
    ax = 0.f;
  }
  if (iy > o.nyotf - 1) iy = o.nyotf - 1;
  if (iy < 0) iy = 0;
  if (iz > o.nzotf - 1) iz = o.nzotf - 1;
  if (iz < 0) iz = 0;
  const int iy1 = iy == o.nyotf - 1 ? 0 : iy + 1;
  const int iz1 = iz == o.nzotf - 1 ? 0 : iz + 1;
  const long long px = o.nxotf, pz = (long long)o.nxotf * o.nyotf;
  auto at = [&](int y, int z) {
    const float2* r = o.t + z * pz + y * px + ix;
    const float2 a = r[0], b = r[1];
    return make_float2(a.x * (1 - ax) + b.x * ax, a.y * (1 - ax) + b.y * ax);
  };
  const float2 v00 = at(iy, iz), v10 = at(iy1, iz), v01 = at(iy, iz1), v11 = at(iy1, iz1);
  const float2 w0 = make_float2(v00.x * (1 - ay) + v10.x * ay, v00.y * (1 - ay) + v10.y * ay);
  const float2 w1 = make_float2(v01.x * (1 - ay) + v11.x * ay, v01.y * (1 - ay) + v11.y * ay);
  float2 v = make_float2(w0.x * (1 - az) + w1.x * az, w0.y * (1 - az) + w1.y * az);
  if (cj) v.y = -v.y;
  return v;
}

__device__ __forceinline__ float2 cmul_nc(float2 a, float2 b) {
#pragma clang fp contract(off)
  float re = a.x * b.x - a.y * b.y;
  float im = a.x * b.y + a.y * b.x;
  return make_float2(re, im);
}

template <int BS>
__device__ __forceinline__ double block_sum_d(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

}  // namespace simr
