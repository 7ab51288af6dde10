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
// prof: this plane's radial OTF profiles at kz = z0, [order][nxotf].
__device__ void filter_visit(const FilterParams& p, const float2* prof, int x1, int y1, int z0, bool conj, float2& re,
                             float2& im) {
#pragma clang fp contract(off)
  const int order = p.order;
  const float kx = (float)order * p.k0[p.dir].x;
  const float ky = (float)order * p.k0[p.dir].y;
  const float x1f = (float)x1 * p.dkx, y1f = (float)y1 * p.dky;
  const float rdist1 = sqrtf(x1f * x1f + y1f * y1f);
  const float xabs = x1f + kx, yabs = y1f + ky;
  const float rdistabs = sqrtf(xabs * xabs + yabs * yabs);
  const int nxotf = p.otf[0].nxotf;
  const float dkr = p.otf[0].dkrotf;
  const float2 otf1 = profile_lookup(prof + order * nxotf, nxotf, dkr, x1f, y1f);
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
        const float2 otf2 = profile_lookup(prof + ao2 * nxotf, nxotf, dkr, x2, y2);
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

// Workgroup (blockIdx.x, iz = blockIdx.y): 256 stored elements (jin, iin) of
// plane iz of the band pair.  The plane's radial OTF profiles at kz = +z0 and
// kz = -z0 (the conjugate visit) are built once in LDS.  Planes whose z lies
// outside +-zd are zeroed (kernel3).
__global__ __launch_bounds__(256) void k_filter(FilterParams p) {
#pragma clang fp contract(off)
  extern __shared__ float2 prof[];  // [2][norders][nxotf]
  const long long nplane = (long long)p.nxh * p.ny;
  const int iz = blockIdx.y;
  const long long ep = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long e = (long long)iz * nplane + ep;
  const int zd = p.zd[p.order];
  float2* bp = p.bp;
  float2* bp2 = p.bp2;
  const bool two = p.order != 0;
  int z0n = 0;
  if (iz <= zd)
    z0n = iz;
  else if (iz >= p.nz - zd)
    z0n = iz - p.nz;
  else {
    if (ep < nplane) {
      bp[e] = make_float2(0.f, 0.f);
      if (two) bp2[e] = make_float2(0.f, 0.f);
    }
    return;  // whole block: uniform
  }
  const int nxotf = p.otf[0].nxotf;
  const int nprof = p.norders * nxotf;
  for (int i = threadIdx.x; i < 2 * nprof; i += 256) {
    const int side = i / nprof, rem = i - side * nprof;
    const int o = rem / nxotf, ir = rem - o * nxotf;
    prof[i] = otf_profile_entry(p.otf[o], ir, side ? -z0n : z0n, p.kzscale);
  }
  __syncthreads();
  if (ep >= nplane) return;
  const int iin = (int)(ep / p.nxh), jin = (int)(ep - (long long)iin * p.nxh);
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
  filter_visit(p, prof, jin, y1n, z0n, false, re, im);
  // second pass, x1 = -jin < 0: the conjugate visit of the mirrored frequency
  if (two && jin >= 1 && jin <= p.nx / 2 - 1) {
    const int yc = -((iin <= p.ny / 2) ? iin : iin - p.ny);
    filter_visit(p, prof + nprof, -jin, yc, -z0n, true, re, im);
  }
  bp[e] = re;
  if (two) bp2[e] = im;
}

void launch_filter(const FilterParams& p, hipStream_t s) {
  const long long nplane = (long long)p.nxh * p.ny;
  const size_t lds = (size_t)2 * p.norders * p.otf[0].nxotf * sizeof(float2);
  if (lds > 64 * 1024) throw ArgError("OTF table too long for the filter's LDS profile");
  hipLaunchKernelGGL(k_filter, dim3(ceil_div(nplane, 256), p.nz), dim3(256), lds, s, p);
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
