// Generalised Wiener filter with OTF lookup, applied in place to one
// direction/order band pair (reference: filterbands,
// src/cudaSirecon/gpuFunctionsImpl.cu:1222-1441; filterbands_kernel1
// 1443-1656; filterbands_kernel3 1660-1675; dev_suppress / dev_order0damping
// 1717-1739).
//
// The reference launches kernel1 twice per order: a first pass over
// x1 in [1, nx/2] and a second over x1 in [1-nx/2, 0]; every stored element
// with 1 <= x <= nx/2-1 is updated once by each pass (first as itself, then
// as the conjugate of the mirrored frequency), and kernel3 zeroes the planes
// outside +-zdistcutoff.  Here one thread owns one stored element and applies
// the same two updates in the same order in registers, so the band is read
// and written once and the result equals the two-pass one.
#include "device_math.hpp"
#include "stages.hpp"

namespace simr {

__device__ __forceinline__ float dev_suppress(float x) {
#pragma clang fp contract(off)
  float x6 = x * x * x;
  x6 *= x6;
  const float den = 1.0f + 20000.0f / (x6 + 20.0f);
  return (float)(1.0 / (double)den);
}

__device__ __forceinline__ float dev_order0damping(float radius, float zindex, float rlimit, int zlimit) {
#pragma clang fp contract(off)
  const float rf = radius / rlimit;
  const float zf = fabsf(zindex / (float)zlimit);
  return rf * rf + zf * zf * zf;
}

// One visit of filterbands_kernel1 at band coordinates (x1, y1, z0) whose
// rdist1 <= rdistcutoff; updates (re, im) of the stored element.
__device__ void filter_visit(const FilterParams& p, int x1, int y1, int z0, bool conj, float2& re, float2& im) {
#pragma clang fp contract(off)
  const int order = p.order;
  const float kx = (float)order * p.k0[p.dir].x;
  const float ky = (float)order * p.k0[p.dir].y;
  const float x1f = (float)x1 * p.dkx, y1f = (float)y1 * p.dky;
  const float rdist1 = sqrtf(x1f * x1f + y1f * y1f);
  const float xabs = x1f + kx, yabs = y1f + ky;
  const float rdistabs = sqrtf(xabs * xabs + yabs * yabs);
  const float2 otf1 = otf_interp(p.otf[order], x1f, y1f, z0, p.kzscale);
  float weight = otf1.x * otf1.x + otf1.y * otf1.y;
  if (order != 0) weight *= p.ampmag2_all[p.dir * p.norders + order];
  float dampfact = (float)(1. / (double)p.noiseVar[p.dir * p.norders + order]);
  if (p.bSuppress && order != 0 && rdist1 <= p.suppRadius)
    dampfact *= dev_suppress(rdist1 / p.min_dkr);
  else if (!p.bDampenOrder0 && p.bSuppress && order == 0 && rdist1 <= p.suppRadius)
    dampfact *= dev_suppress(rdist1 / p.min_dkr);
  else if (p.bDampenOrder0 && order == 0)
    dampfact *= dev_order0damping(rdist1, (float)z0, p.rdistcutoff, p.zd[0]);
  if (order == 0 && z0 == 0 && p.bNoKz0) dampfact = 0.f;
  weight *= dampfact;
  float sumweight = weight;
  for (int dir2 = 0; dir2 < p.ndirs; dir2++) {
    for (int order2 = -(p.norders - 1); order2 < p.norders; order2++) {
      if (dir2 == p.dir && order2 == order) continue;
      if (!p.bFilteroverlaps && !(order2 == 0 && order == 0)) continue;
      const int ao2 = order2 < 0 ? -order2 : order2;
      const float amp2mag2 = p.ampmag2_all[dir2 * p.norders + ao2];
      const float kx2 = (float)order2 * p.k0[dir2].x, ky2 = (float)order2 * p.k0[dir2].y;
      const float x2 = xabs - kx2, y2 = yabs - ky2;
      const float rdist2 = sqrtf(x2 * x2 + y2 * y2);
      if (rdist2 < p.rdistcutoff) {
        const float2 otf2 = otf_interp(p.otf[ao2], x2, y2, z0, p.kzscale);
        float w = (otf2.x * otf2.x + otf2.y * otf2.y) / p.noiseVar[dir2 * p.norders + ao2];
        if (order2 != 0) w *= amp2mag2;
        if (p.bSuppress && order2 != 0 && rdist2 <= p.suppRadius)
          w *= dev_suppress(rdist2 / p.min_dkr);
        else if (!p.bDampenOrder0 && p.bSuppress && order2 == 0 && rdist2 <= p.suppRadius)
          w *= dev_suppress(rdist2 / p.min_dkr);
        else if (p.bDampenOrder0 && order2 == 0)
          w *= dev_order0damping(rdist2, (float)z0, p.rdistcutoff, p.zd[0]);
        if (p.bNoKz0 && order2 == 0 && z0 == 0) w = 0.0f;
        sumweight += w;
      }
    }
  }
  sumweight += p.wiener;
  float2 scale = make_float2(dampfact * otf1.x / sumweight, dampfact * (-otf1.y) / sumweight);
  if (p.apodizeoutput) {
    float rho;
    const float zdistabs = (float)abs(z0);
    if (p.zapocutoff > 0)
      rho = sqrtf((rdistabs / p.apocutoff) * (rdistabs / p.apocutoff) +
                  (zdistabs / p.zapocutoff) * (zdistabs / p.zapocutoff));
    else
      rho = sqrtf((rdistabs / p.apocutoff) * (rdistabs / p.apocutoff));
    if (rho > 1.f) rho = 1.0f;
    float apofact = 0.f;
    if (p.apodizeoutput == 1)
      apofact = cosf((float)((3.14159265358979323846 * 0.5f) * (double)rho));
    else if (p.apodizeoutput == 2)
      apofact = 1.0f - rho;
    scale.x *= apofact;
    scale.y *= apofact;
  }
  if (order == 0) {
    re = cmul_nc(re, scale);
  } else {
    scale = cmul_nc(scale, p.conjamp);
    if (conj) {
      re.y = -re.y;
      im.y = -im.y;
    }
    float2 plus = make_float2(re.x - im.y, re.y + im.x);
    const float2 minus = make_float2(re.x + im.y, re.y - im.x);
    plus = cmul_nc(plus, scale);
    re.x = 0.5f * (plus.x + minus.x);
    re.y = 0.5f * (plus.y + minus.y);
    im.x = 0.5f * (plus.y - minus.y);
    im.y = 0.5f * (-plus.x + minus.x);
    if (conj) {
      re.y = -re.y;
      im.y = -im.y;
    }
  }
}

// One thread per stored element (jin, iin, iz) of the band pair; planes whose
// z lies outside +-zd are zeroed (kernel3).
__global__ __launch_bounds__(256) void k_filter(FilterParams p) {
#pragma clang fp contract(off)
  const long long nplane = (long long)p.nxh * p.ny;
  const long long total = nplane * p.nz;
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= total) return;
  const int iz = (int)(e / nplane);
  const long long rem = e - (long long)iz * nplane;
  const int iin = (int)(rem / p.nxh), jin = (int)(rem - (long long)iin * p.nxh);
  const int zd = p.zd[p.order];
  int z0n;
  bool visited = true;
  if (iz <= zd)
    z0n = iz;
  else if (iz >= p.nz - zd)
    z0n = iz - p.nz;
  else {
    z0n = 0;
    visited = false;
  }
  float2* bp = p.bp;
  float2* bp2 = p.bp2;
  const bool two = p.order != 0;
  if (!visited) {
    bp[e] = make_float2(0.f, 0.f);
    if (two) bp2[e] = make_float2(0.f, 0.f);
    return;
  }
  const int y1n = (iin < p.ny - p.ny / 2) ? iin : iin - p.ny;
  const float x1f = (float)jin * p.dkx, y1f = (float)y1n * p.dky;
  const float rdist1 = sqrtf(x1f * x1f + y1f * y1f);
  if (!(rdist1 <= p.rdistcutoff)) {
    bp[e] = make_float2(0.f, 0.f);
    if (two) bp2[e] = make_float2(0.f, 0.f);
    return;
  }
  float2 re = bp[e];
  float2 im = two ? bp2[e] : make_float2(0.f, 0.f);
  // first pass (x1 in [1, nx/2]) or second pass x1 == 0: the element itself
  filter_visit(p, jin, y1n, z0n, false, re, im);
  // second pass, x1 = -jin < 0: the conjugate visit of the mirrored frequency
  if (two && jin >= 1 && jin <= p.nx / 2 - 1) {
    const int yc = -((iin <= p.ny / 2) ? iin : iin - p.ny);
    filter_visit(p, -jin, yc, -z0n, true, re, im);
  }
  bp[e] = re;
  if (two) bp2[e] = im;
}

void launch_filter(const FilterParams& p, hipStream_t s) {
  const long long total = (long long)p.nxh * p.ny * p.nz;
  hipLaunchKernelGGL(k_filter, dim3(ceil_div(total, 256)), dim3(256), 0, s, p);
  SIMR_LAUNCH_CHECK();
}

__global__ void k_zero(float2* __restrict__ p, long long n) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e < n) p[e] = make_float2(0.f, 0.f);
}

void launch_zero(float2* p, long long n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, n);
  SIMR_LAUNCH_CHECK();
}

}  // namespace simr
