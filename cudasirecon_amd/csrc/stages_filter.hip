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

// OTF profile value at radius index krindex, bilinear as profile_lookup.  The
// index is a square root, so never negative: truncation is the floor (the
// reference's clamp at 0 cannot act), and the interpolation is
// t0 + ar (t1 - t0) in fused form (within an ulp of (1 - ar) t0 + ar t1).
__device__ __forceinline__ float2 prof_pos(const float2* prof, int nxotf, float krindex) {
  int irindex = (int)krindex;
  float ar = __builtin_amdgcn_fractf(krindex);
  if (irindex > nxotf - 2) {
    irindex = nxotf - 2;
    ar = 1.f;
  }
  const float2 t0 = prof[irindex], t1 = prof[irindex + 1];
  return make_float2(__builtin_fmaf(ar, t1.x - t0.x, t0.x), __builtin_fmaf(ar, t1.y - t0.y, t0.y));
}

// OTF table of order o for the band being filtered
__device__ __forceinline__ OtfTab otf_at(const FilterParams& p, int o, long long off) {
  OtfTab t = p.otf[o];
  t.t += off;
  return t;
}

// One visit of filterbands_kernel1 at band coordinates (x1, y1, z0) whose
// rdist1 <= rdistcutoff; updates (re, im) of the stored element.
// prof: this plane's radial OTF profiles at kz = z0, [order][nxotf].
// The arithmetic is the reference's; what is hoisted out of the (dir2,
// order2) loop (the z terms, 1/noiseVar, order2*k0) is computed by the same
// expressions once, and terms whose squared distance already exceeds the
// cut-off are skipped before the square root.
// One (dir2, order2) term of the Wiener denominator at absolute frequency
// (xabs, yabs), plane z0; 0 for the band's own term and for terms outside the
// support.
// UNV: every noiseVar is 1 (known at compile time; the division would
// otherwise be computed and discarded by a select in the unrolled loop).
// |modamp of (dir, order)|^2: the host's value, or formed from the device
// modamps with the host's float expression
__device__ __forceinline__ float amp_mag2(const FilterParams& p, int dir, int order) {
#pragma clang fp contract(off)
  if (p.amps == nullptr) return p.ampmag2_all[dir * p.norders + order];
  const float2 a = p.amps[dir * p.norders + order];
  return a.x * a.x + a.y * a.y;
}

template <bool UNV, bool RA>
__device__ __forceinline__ float filter_term(const FilterParams& p, const float2* prof, int nxotf, int dir, int order,
                                             long long otf_off, int dir2, int order2, float xabs, float yabs, int z0,
                                             float zf3) {
#pragma clang fp contract(off)
  if (dir2 == dir && order2 == order) return 0.f;
  if (!p.bFilteroverlaps && !(order2 == 0 && order == 0)) return 0.f;
  // x2 = xabs - order2 k0x: order2 k0x is exact for |order2| <= 2, so the
  // fused form rounds exactly as the reference's subtraction (k0 is read from
  // the direction's two values rather than a per-term table, which keeps the
  // unrolled loop's constants in scalar registers)
  const float2 k0 = p.k0[dir2];
  const float x2 = __builtin_fmaf(-(float)order2, k0.x, xabs), y2 = __builtin_fmaf(-(float)order2, k0.y, yabs);
  const float d2 = __builtin_fmaf(x2, x2, y2 * y2);
  // rdist2 < rdistcutoff, tested on the squared distance; the OTF radius
  // index sqrt(d2)/dkr and the damping term (rdist2/rdistcutoff)^2 are
  // formed from d2 as well (a few ulps from the reference's sqrtf and
  // divisions, far inside the output tolerance)
  if (!(d2 < p.cut2)) return 0.f;
  const int ao2 = order2 < 0 ? -order2 : order2;
  float2 otf2;
  if constexpr (RA)
    otf2 = prof_pos(prof + ao2 * nxotf, nxotf, __builtin_amdgcn_sqrtf(d2 * p.inv_dkr2));
  else
    otf2 = otf3d_interp(otf_at(p, ao2, otf_off), x2, y2, z0, p.kzscale);
  float w = __builtin_fmaf(otf2.x, otf2.x, otf2.y * otf2.y);
  if (!UNV && !p.unit_nv) w = w / p.noiseVar[dir2 * p.norders + ao2];
  if (order2 != 0) w *= amp_mag2(p, dir2, ao2);
  if (p.bSuppress && order2 != 0 && d2 <= p.supp2)
    w *= dev_suppress(__builtin_amdgcn_sqrtf(d2 * p.inv_min_dkr2));
  else if (!p.bDampenOrder0 && p.bSuppress && order2 == 0 && d2 <= p.supp2)
    w *= dev_suppress(__builtin_amdgcn_sqrtf(d2 * p.inv_min_dkr2));
  else if (p.bDampenOrder0 && order2 == 0)
    w *= d2 * p.inv_rcut2 + zf3;
  if (p.bNoKz0 && order2 == 0 && z0 == 0) w = 0.0f;
  return w;
}

// ND, NO > 0: the direction and order counts as compile-time constants (the
// (dir2, order2) loop is unrolled, so its per-term constants are read once
// and the OTF lookups of different terms overlap); 0: taken from p.
template <int ND, int NO, bool UNV, bool RA>
__device__ __forceinline__ void filter_visit(const FilterParams& p, const FilterBand& fb, const float2* prof, int x1,
                                             int y1, int z0, bool conj, float2& re, float2& im) {
#pragma clang fp contract(off)
  const int dir = fb.dir, order = fb.order;
  const long long otf_off = fb.otf_off;
  const float kx = (float)order * p.k0[dir].x;
  const float ky = (float)order * p.k0[dir].y;
  const float x1f = (float)x1 * p.dkx, y1f = (float)y1 * p.dky;
  // divisions by per-launch constants are products with their reciprocals and
  // square roots are the hardware's (each within an ulp or two of the
  // reference's correctly rounded forms, far inside the output tolerance)
  const float d1 = x1f * x1f + y1f * y1f;
  const float xabs = x1f + kx, yabs = y1f + ky;
  const int nxotf = p.otf[0].nxotf;
  // dev_order0damping's z term: fabsf(zindex / zlimit) cubed
  float zf3 = 0.f;
  if (p.bDampenOrder0) {
    const float zf = fabsf((float)z0 * p.inv_zd0);
    zf3 = zf * zf * zf;
  }
  float2 otf1;
  if constexpr (RA)
    otf1 = prof_pos(prof + order * nxotf, nxotf, __builtin_amdgcn_sqrtf(d1 * p.inv_dkr2));
  else
    otf1 = otf3d_interp(otf_at(p, order, otf_off), x1f, y1f, z0, p.kzscale);
  float weight = otf1.x * otf1.x + otf1.y * otf1.y;
  if (order != 0) weight *= amp_mag2(p, dir, order);
  float dampfact = p.inv_nv[dir * p.norders + order];
  if (p.bSuppress && order != 0 && d1 <= p.supp2)
    dampfact *= dev_suppress(__builtin_amdgcn_sqrtf(d1 * p.inv_min_dkr2));
  else if (!p.bDampenOrder0 && p.bSuppress && order == 0 && d1 <= p.supp2)
    dampfact *= dev_suppress(__builtin_amdgcn_sqrtf(d1 * p.inv_min_dkr2));
  else if (p.bDampenOrder0 && order == 0)
    dampfact *= d1 * p.inv_rcut2 + zf3;
  if (order == 0 && z0 == 0 && p.bNoKz0) dampfact = 0.f;
  weight *= dampfact;
  float sumweight = weight;
  if constexpr (ND > 0) {
    // With a radially averaged OTF and unit noise variances the order-0 term
    // is the same for every direction (x2 = xabs, the band's own profile,
    // the same damping): it is formed once and added at each direction's
    // place in the sum, so the summation order stays the reference's
    float w0 = 0.f;
    if constexpr (RA && UNV) w0 = filter_term<UNV, RA>(p, prof, nxotf, -1, order, otf_off, 0, 0, xabs, yabs, z0, zf3);
#pragma unroll
    for (int dir2 = 0; dir2 < ND; dir2++)
#pragma unroll
      for (int order2 = -(NO - 1); order2 < NO; order2++) {
        if (RA && UNV && order2 == 0)
          sumweight += (dir2 == dir && order == 0) ? 0.f : w0;
        else
          sumweight += filter_term<UNV, RA>(p, prof, nxotf, dir, order, otf_off, dir2, order2, xabs, yabs, z0, zf3);
      }
  } else {
    for (int dir2 = 0; dir2 < p.ndirs; dir2++)
      for (int order2 = -(p.norders - 1); order2 < p.norders; order2++)
        sumweight += filter_term<false, RA>(p, prof, nxotf, dir, order, otf_off, dir2, order2, xabs, yabs, z0, zf3);
  }
  sumweight += p.wiener;
  const float rs = __builtin_amdgcn_rcpf(sumweight);
  float2 scale = make_float2((dampfact * otf1.x) * rs, (dampfact * (-otf1.y)) * rs);
  if (p.apodizeoutput) {
    float rho;
    const float zdistabs = (float)abs(z0);
    if (p.zapocutoff > 0 && p.bessel) {
      const float rx = xabs * p.inv_rcut, ry = yabs * p.inv_apo, rz = zdistabs * p.inv_zapo;
      rho = __builtin_amdgcn_sqrtf(rx * rx + ry * ry + rz * rz);
    } else {
      // (r / apocutoff)^2 + (z / zapocutoff)^2, the z term absent in 2D
      const float ra2 = (xabs * xabs + yabs * yabs) * p.inv_apo2;
      const float za = p.zapocutoff > 0 ? zdistabs * p.inv_zapo : 0.f;
      rho = __builtin_amdgcn_sqrtf(ra2 + za * za);
    }
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
    float2 ca = fb.conjamp;
    if (p.amps != nullptr) {
      const float2 a = p.amps[dir * p.norders + order];
      ca = make_float2(a.x, -a.y);
    }
    scale = cmul_nc(scale, ca);
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

// Workgroup (blockIdx.x, iz = blockIdx.y, band blockIdx.z of the batch):
// per*256 stored elements (jin, iin) of plane iz of the band pair (per = 4,
// or 1 when the batch is too small to fill the device that way).  The plane's radial OTF profiles at
// kz = +z0 and kz = -z0 (the conjugate visit) are built once in LDS.  Planes
// whose z lies outside +-zd are zeroed (kernel3).
constexpr int kFiltPer = 4;

template <int ND, int NO, bool UNV, bool RA>
__global__ __launch_bounds__(256) void k_filter(FilterParams p, FilterBatch batch, int per) {
#pragma clang fp contract(off)
  extern __shared__ float2 prof[];  // [2][norders][nxotf]
  const FilterBand fb = batch.b[blockIdx.z];
  const long long nplane = (long long)p.nxh * p.ny;
  const int iz = blockIdx.y;
  const long long ep0 = (long long)blockIdx.x * (256 * per) + threadIdx.x;
  const int zd = p.zd[fb.order];
  float2* bp = fb.bp;
  float2* bp2 = fb.bp2;
  const bool two = fb.order != 0;
  int z0n = 0;
  if (iz <= zd)
    z0n = iz;
  else if (iz >= p.nz - zd)
    z0n = iz - p.nz;
  else {
    for (int k = 0; k < per; ++k) {
      const long long ep = ep0 + k * 256;
      if (ep < nplane) {
        const long long e = (long long)iz * nplane + ep;
        bp[e] = make_float2(0.f, 0.f);
        if (two) bp2[e] = make_float2(0.f, 0.f);
      }
    }
    return;  // whole block: uniform
  }
  const int nxotf = p.otf[0].nxotf;
  const int nprof = p.norders * nxotf;
  if constexpr (RA) {
    for (int i = threadIdx.x; i < 2 * nprof; i += 256) {
      const int side = i / nprof, rem = i - side * nprof;
      const int o = rem / nxotf, ir = rem - o * nxotf;
      prof[i] = otf_profile_entry(otf_at(p, o, fb.otf_off), ir, side ? -z0n : z0n, p.kzscale);
    }
    __syncthreads();
  }
  for (int k = 0; k < per; ++k) {
    const long long ep = ep0 + k * 256;
    if (ep >= nplane) break;
    const long long e = (long long)iz * nplane + ep;
    // (iin, jin) = divmod(ep, nxh) without a 64-bit division (ep < 2^24)
    int iin = (int)((float)ep * p.inv_nxh);
    int jin = (int)ep - iin * p.nxh;
    if (jin < 0) {
      iin--;
      jin += p.nxh;
    } else if (jin >= p.nxh) {
      iin++;
      jin -= p.nxh;
    }
    const int y1n = (iin < p.ny - p.ny / 2) ? iin : iin - p.ny;
    const float x1f = (float)jin * p.dkx, y1f = (float)y1n * p.dky;
    const float rdist1 = sqrtf(x1f * x1f + y1f * y1f);
    if (!(rdist1 <= p.rdistcutoff)) {
      bp[e] = make_float2(0.f, 0.f);
      if (two) bp2[e] = make_float2(0.f, 0.f);
      continue;
    }
    float2 re = bp[e];
    float2 im = two ? bp2[e] : make_float2(0.f, 0.f);
    // visit 0: first pass (x1 in [1, nx/2]) or second pass x1 == 0, the
    // element itself; visit 1: second pass, x1 = -jin < 0, the conjugate
    // visit of the mirrored frequency.  One loop body for both keeps one copy
    // of the unrolled denominator (and its constants) in the kernel.
    const int nvis = (two && jin >= 1 && jin <= p.nx / 2 - 1) ? 2 : 1;
    const int yc = -((iin <= p.ny / 2) ? iin : iin - p.ny);
    for (int v = 0; v < nvis; ++v)
      filter_visit<ND, NO, UNV, RA>(p, fb, v ? prof + nprof : prof, v ? -jin : jin, v ? yc : y1n, v ? -z0n : z0n,
                                    v != 0, re, im);
    bp[e] = re;
    if (two) bp2[e] = im;
  }
}

template <bool RA>
static void launch_filter_kind(const FilterParams& p, const FilterBatch& b, size_t lds, dim3 grid, int per,
                               hipStream_t s) {
  const bool unv = p.unit_nv != 0;
  if (p.ndirs == 3 && p.norders == 3) {  // the 3-beam 3D case
    if (unv)
      hipLaunchKernelGGL((k_filter<3, 3, true, RA>), grid, dim3(256), lds, s, p, b, per);
    else
      hipLaunchKernelGGL((k_filter<3, 3, false, RA>), grid, dim3(256), lds, s, p, b, per);
  } else if (p.ndirs == 3 && p.norders == 2) {  // 2D, 3 phases
    if (unv)
      hipLaunchKernelGGL((k_filter<3, 2, true, RA>), grid, dim3(256), lds, s, p, b, per);
    else
      hipLaunchKernelGGL((k_filter<3, 2, false, RA>), grid, dim3(256), lds, s, p, b, per);
  } else {
    hipLaunchKernelGGL((k_filter<0, 0, false, RA>), grid, dim3(256), lds, s, p, b, per);
  }
}

void launch_filter_batch(const FilterParams& p, const FilterBatch& b, hipStream_t s) {
  if (b.n < 1 || b.n > kMaxDirs * kMaxOrders) throw ArgError("filter: batch of " + std::to_string(b.n) + " bands");
  for (int k = 0; k < b.n; ++k) {
    const FilterBand& e = b.b[k];
    if (e.dir < 0 || e.dir >= p.ndirs || e.order < 0 || e.order >= p.norders || e.bp == nullptr ||
        (e.order != 0 && e.bp2 == nullptr))
      throw ArgError("filter: band entry out of range");
  }
  const long long nplane = (long long)p.nxh * p.ny;
  // four elements per thread amortise the plane's profile build; a batch
  // that would then give fewer than 2048 workgroups takes one per thread
  const int per = ceil_div(nplane, 256 * kFiltPer) * p.nz * b.n >= 2048 ? kFiltPer : 1;
  const dim3 grid(ceil_div(nplane, 256 * per), p.nz, b.n);
  if (p.otf[0].ra) {
    const size_t lds = (size_t)2 * p.norders * p.otf[0].nxotf * sizeof(float2);
    if (lds > 64 * 1024) throw ArgError("OTF table too long for the filter's LDS profile");
    launch_filter_kind<true>(p, b, lds, grid, per, s);
  } else {
    launch_filter_kind<false>(p, b, 0, grid, per, s);
  }
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
