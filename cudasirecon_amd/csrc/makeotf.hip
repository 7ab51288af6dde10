// makeotf on the GPU: PSF stack -> rotationally averaged OTF.
//
// Follows the reference's makeotf program (src/otf/radialft.cpp:82-434),
// which runs on the CPU with FFTW.  Device passes:
//   k_psf_prep        camera-pixel fix, border background, copy, edge apodization
//                     (radialft.cpp:207-260, 487-503, 702-744)
//   k_center_*        phase-mean peak and its six neighbours (determine_center, 507-665)
//   k_psf_separate    background subtraction + 1/n separation matrix (280-300, 440-485)
//   fft3d_any         3D forward transform of every band (FFTW r2c, 339-348)
//   k_shift_compensate  bead-centre phase ramp and bead-size compensation
//                     (shift_center 846-874, beadsize_compensate 784-841, sphereFFT 1155-1167)
//   k_radial_avg      radial binning in the reference's summation order (radialft 876-936)
//   k_otf_finish      symmetrisation, modify, cleanup, fixorigin, rescale,
//                     combine_reim and conjugation (radialft.cpp:393-420, 746-782,
//                     938-1034, 1169-1242) in one workgroup
// The host computes the three parabola fits of the bead centre and the
// (pixel -> radial bin) lists, both from the reference's float expressions.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/simr.h"
#include "common.hpp"
#include "fft_api.hpp"

namespace simr {

namespace {

constexpr double kPId = 3.14159265358979323846;
constexpr int kFinishThreads = 1024;
constexpr int kMaxFixorigin = 1024;

struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) { SIMR_HIP(hipMalloc(&p, bytes < 256 ? 256 : bytes)); }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
};

struct PrepParams {
  int nx, ny, nz, nphases, napod, estimate;
  int fix_minus1;  // camera fix when the last pixel is exactly -1 (the I2M sections), else when it is < 0
};

// One workgroup per section.
__global__ __launch_bounds__(1024) void k_psf_prep(PrepParams p, const float* __restrict__ psf, float* __restrict__ vol,
                                                   float* __restrict__ bg, const float* __restrict__ apof,
                                                   const float* __restrict__ cosx, const float* __restrict__ cosy) {
#pragma clang fp contract(off)
  __shared__ double red[1024];
  __shared__ unsigned cnt[1024];
  const int s = blockIdx.x;
  const int z = s / p.nphases, phase = s - z * p.nphases;
  const long long nxy = (long long)p.nx * p.ny;
  const float* src = psf + (long long)s * nxy;
  float* dst = vol + ((long long)phase * p.nz + z) * nxy;
  const float last = src[nxy - 1];
  const bool fix = p.fix_minus1 ? last == -1.0f : last < 0;
  const float lastv = fix ? src[nxy - 2] : last;  // camera fix of the last pixel
  const int tid = threadIdx.x, nt = blockDim.x;
  if (p.estimate) {
    const int border = 20;
    double sum = 0.0;
    unsigned c = 0;
    for (long long i = tid; i < nxy; i += nt) {
      const int y = (int)(i / p.nx), x = (int)(i - (long long)y * p.nx);
      if (y < border || y > p.ny - border || x < border || x > p.nx - border) {
        sum += (double)(i == nxy - 1 ? lastv : src[i]);
        c++;
      }
    }
    red[tid] = sum;
    cnt[tid] = c;
    __syncthreads();
    for (int w = nt / 2; w > 0; w >>= 1) {
      if (tid < w) {
        red[tid] += red[tid + w];
        cnt[tid] += cnt[tid + w];
      }
      __syncthreads();
    }
    if (tid == 0) bg[s] = (float)(red[0] / cnt[0]);
  }
  for (long long i = tid; i < nxy; i += nt) dst[i] = (i == nxy - 1) ? lastv : src[i];
  __syncthreads();
  const int nx = p.nx, ny = p.ny;
  if (p.napod > 0) {
    for (int k = tid; k < nx; k += nt) {
      const float diff = (dst[(long long)(ny - 1) * nx + k] - dst[k]) / 2;
      for (int l = 0; l < p.napod; l++) {
        const float f = apof[l];
        dst[(long long)l * nx + k] += diff * f;
        dst[(long long)(ny - 1 - l) * nx + k] -= diff * f;
      }
    }
    __syncthreads();
    for (int l = tid; l < ny; l += nt) {
      float* row = dst + (long long)l * nx;
      const float diff = (row[nx - 1] - row[0]) / 2;
      for (int k = 0; k < p.napod; k++) {
        const float f = apof[k];
        row[k] += diff * f;
        row[nx - 1 - k] -= diff * f;
      }
    }
  } else if (p.napod == -1) {
    for (long long i = tid; i < nxy; i += nt) {
      const int y = (int)(i / nx), x = (int)(i - (long long)y * nx);
      dst[i] *= (cosx[x] * cosy[y]);
    }
  }
}

__device__ __forceinline__ bool better(float v, long long i, float bv, long long bi) {
  return v > bv || (v == bv && i < bi);
}

__device__ __forceinline__ float phase_mean(const float* vol, long long nvol, int nphases, long long i) {
#pragma clang fp contract(off)
  float s = 0.f;
  for (int ph = 0; ph < nphases; ph++) s += vol[ph * nvol + i];
  return s / nphases;
}

__global__ __launch_bounds__(256) void k_center_partial(const float* __restrict__ vol, long long nvol, int nphases,
                                                        float* __restrict__ pv, long long* __restrict__ pi) {
  __shared__ float sv[256];
  __shared__ long long si[256];
  float bv = 0.f;
  long long bi = 0x7fffffffffffffffLL;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nvol; i += (long long)gridDim.x * 256) {
    const float v = phase_mean(vol, nvol, nphases, i);
    if (better(v, i, bv, bi)) {
      bv = v;
      bi = i;
    }
  }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w && better(sv[threadIdx.x + w], si[threadIdx.x + w], sv[threadIdx.x], si[threadIdx.x])) {
      sv[threadIdx.x] = sv[threadIdx.x + w];
      si[threadIdx.x] = si[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = sv[0];
    pi[blockIdx.x] = si[0];
  }
}

// out: [0] maxval, [1..2] z-/z+, [3..4] y-/y+, [5..6] x-/x+ ; iout: maxk, maxi, maxj
__global__ __launch_bounds__(256) void k_center_final(const float* __restrict__ vol, int nx, int ny, int nz,
                                                      int nphases, const float* __restrict__ pv,
                                                      const long long* __restrict__ pi, int nparts,
                                                      float* __restrict__ out, int* __restrict__ iout) {
  __shared__ float sv[256];
  __shared__ long long si[256];
  float bv = 0.f;
  long long bi = 0x7fffffffffffffffLL;
  for (int k = threadIdx.x; k < nparts; k += 256)
    if (better(pv[k], pi[k], bv, bi)) {
      bv = pv[k];
      bi = pi[k];
    }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w && better(sv[threadIdx.x + w], si[threadIdx.x + w], sv[threadIdx.x], si[threadIdx.x])) {
      sv[threadIdx.x] = sv[threadIdx.x + w];
      si[threadIdx.x] = si[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const long long nxy = (long long)nx * ny, nvol = nxy * nz;
  float maxval = 0.f;
  int maxi = 0, maxj = 0, maxk = 0;
  if (sv[0] > 0.f) {  // strict '>' against the initial 0 (determine_center)
    maxval = sv[0];
    const long long b = si[0];
    maxk = (int)(b / nxy);
    maxi = (int)((b - (long long)maxk * nxy) / nx);
    maxj = (int)(b - (long long)maxk * nxy - (long long)maxi * nx);
  }
  // the reference's stack is padded to an even width + 2 (or nx+1): x wraps
  // with that width and the padded columns read as 0
  const int nxe = (nx % 2) ? nx + 1 : nx + 2;
  auto at = [&](int x, int y, int z) -> float {
    if (x >= nx) return 0.f;
    return phase_mean(vol, nvol, nphases, ((long long)z * ny + y) * nx + x);
  };
  int iminus = maxi - 1, iplus = maxi + 1, jminus = maxj - 1, jplus = maxj + 1, kminus = maxk - 1, kplus = maxk + 1;
  if (iminus < 0) iminus += ny;
  if (iplus >= ny) iplus -= ny;
  if (jminus < 0) jminus += nxe;
  if (jplus >= nxe) jplus -= nxe;
  if (kminus < 0) kminus += nz;
  if (kplus >= nz) kplus -= nz;
  out[0] = maxval;
  out[1] = at(maxj, maxi, kminus);
  out[2] = at(maxj, maxi, kplus);
  out[3] = at(maxj, iminus, maxk);
  out[4] = at(maxj, iplus, maxk);
  out[5] = at(jminus, maxi, maxk);
  out[6] = at(jplus, maxi, maxk);
  iout[0] = maxk;
  iout[1] = maxi;
  iout[2] = maxj;
}

struct SepParams {
  long long nxy, nvox;
  int nphases;
  float M[SIMR_MAX_PHASES * SIMR_MAX_PHASES];  // M[j*nphases + i] = sepMatrix(j, i)
};

__global__ __launch_bounds__(256) void k_psf_separate(SepParams p, const float* __restrict__ vol,
                                                      const float* __restrict__ bg, float2* __restrict__ spec,
                                                      float* __restrict__ band0) {
#pragma clang fp contract(off)
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= p.nvox) return;
  const int z = (int)(i / p.nxy);
  float v[SIMR_MAX_PHASES];
  for (int j = 0; j < p.nphases; j++) v[j] = vol[j * p.nvox + i] - bg[z * p.nphases + j];
  for (int b = 0; b < p.nphases; b++) {
    float o;
    if (p.nphases > 1) {
      o = 0.0f;
      for (int j = 0; j < p.nphases; j++) o += v[j] * p.M[j * p.nphases + b];
    } else {
      o = v[0];
    }
    spec[b * p.nvox + i] = make_float2(o, 0.f);
    if (b == 0 && band0) band0[i] = o;
  }
}

struct ShiftParams {
  int nx, ny, nz, nphases, do_compen;
  float dphix, dphiy, dphiz;
  float dkx, dky, dkz, radius, limit;
  float k0x[SIMR_MAX_ORDERS], k0y[SIMR_MAX_ORDERS];
};

__device__ double sphere_ft(float k, float radius) {
  if (k > 0 || k < 0) {
    const double x = 2 * kPId * radius * k;
    return radius / (kPId * k * k) * (sin(x) / x - cos(x));
  }
  return 4 * kPId * ((double)radius * radius * radius) / 3;
}

// In place on every band's full spectrum (pitch nx), x <= nx/2 only.
__global__ __launch_bounds__(256) void k_shift_compensate(ShiftParams p, float2* __restrict__ spec) {
#pragma clang fp contract(off)
  const int nxh = p.nx / 2 + 1;
  const long long per = (long long)nxh * p.ny * p.nz;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= per * p.nphases) return;
  const int b = (int)(t / per);
  const long long e = t - (long long)b * per;
  const int jin = (int)(e % nxh);
  const int iin = (int)((e / nxh) % p.ny);
  const int kin = (int)(e / ((long long)nxh * p.ny));
  float2* v = spec + (long long)b * p.nx * p.ny * p.nz + ((long long)kin * p.ny + iin) * p.nx + jin;
  int kz = kin;
  if (kz > p.nz / 2) kz -= p.nz;
  int ky = iin;
  if (iin > p.ny / 2) ky -= p.ny;
  const float phi1 = p.dphiz * kz;
  const float phi2 = p.dphiy * ky;
  const float phi = phi1 + phi2 + p.dphix * jin;
  const float c = (float)cos((double)phi), sn = (float)sin((double)phi);
  float2 a = *v;
  a = make_float2(a.x * c - a.y * sn, a.x * sn + a.y * c);
  if (p.do_compen) {
    const int order = (b + 1) / 2;
    float ratio;
    if (!(order == 0 && kin == 0 && iin == 0 && jin == 0)) {
      const float kzf = kz * p.dkz;
      const float kyf = ky * p.dky + p.k0y[order];
      const float kxf = jin * p.dkx + p.k0x[order];
      const float rho = (float)sqrt((double)(kzf * kzf + kyf * kyf + kxf * kxf));
      ratio = (float)(sphere_ft(rho, p.radius) / (double)p.limit);
    } else {
      ratio = 1;
    }
    a = make_float2(a.x / ratio, a.y / ratio);
  }
  *v = a;
}

// One thread per (band, radial bin, kz): sums the bin's pixels in row-major
// order, as the reference's loop does for each output element.
__global__ __launch_bounds__(256) void k_radial_avg(const float2* __restrict__ spec, int nx, int ny, int nz, int nkr,
                                                    int nbands, const int* __restrict__ off,
                                                    const int* __restrict__ pix, float2* __restrict__ avg) {
#pragma clang fp contract(off)
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long long)nbands * nkr * nz) return;
  const int kz = (int)(t % nz);
  const int r = (int)((t / nz) % nkr);
  const int b = (int)(t / ((long long)nz * nkr));
  const float2* plane = spec + (long long)b * nx * ny * nz + (long long)kz * nx * ny;
  float2 acc = make_float2(0.f, 0.f);
  const int a0 = off[r], a1 = off[r + 1];
  for (int k = a0; k < a1; k++) {
    const float2 v = plane[pix[k]];
    acc.x = acc.x + v.x;
    acc.y = acc.y + v.y;
  }
  if (a1 > a0) {
    const float c = (float)(a1 - a0);
    acc = make_float2(acc.x / c, acc.y / c);
  }
  avg[((long long)b * nkr + r) * nz + kz] = acc;
}

struct FinishParams {
  int nphases, norders, nkr, nz, nx;
  int ifixz, ifixr;
  int cleanup, twolens, leave1_1, leave1_2, leave2, icleanup;
  float na, nimm, linespacing, dr, dz;
  unsigned lamdanm;
  int kx1, kx2;
  int dorescale, combine, forced_pi, conjugate;
};

// cleanup (radialft.cpp:938-1034) of one radial row ix; dkr = 1/(nr*dr).
__device__ void cleanup_row(float2* otf, int ix, int order, const FinishParams& p, float dkr) {
#pragma clang fp contract(off)
  const int nz = p.nz;
  const float lamda = p.lamdanm * 0.001;
  const float sinalpha = p.na / p.nimm;
  const float cosalpha = (float)cos(asin((double)sinalpha));
  const float krmax = 2 * p.na / lamda;
  const float k0mag = 1.0 / p.linespacing;
  const float dkz = 1. / (nz * p.dz);
  const float2 zero = make_float2(0.f, 0.f);
  float2* row = otf + (long long)ix * nz;
  const float kr = ix * dkr;
  if (!p.twolens) {
    if (kr <= krmax) {
      const double beta = asin((double)((p.na - kr * lamda) / p.nimm));
      const double kzedge = (double)(p.nimm / lamda) * (cos(beta) - (double)cosalpha);
      if (order == 0) {
        const int kzstart = (int)rint((kzedge / dkz) + 1);
        const int kzend = nz - kzstart;
        for (int iz = max(kzstart, 0); iz <= kzend && iz < nz; iz++) row[iz] = zero;
      } else if (order == 1) {
        const int jotfshape = (int)rint((kzedge / dkz) + 0.999);
        const int kzend = min(p.leave1_1 - jotfshape, nz - 1);
        if (kzend >= 0)
          for (int iz = 0; iz <= kzend; iz++) {
            if (iz == 0) {
              row[0] = zero;
            } else {
              row[iz] = zero;
              row[nz - iz] = zero;
            }
          }
        const int kzstart = p.leave1_2 + jotfshape;
        for (int iz = max(kzstart + 1, 1); iz < nz / 2 + 1; iz++) {
          row[iz] = zero;
          row[nz - iz] = zero;
        }
      } else {
        const int jotfshape = (int)rint((kzedge / dkz) + 0.999);
        const int kzstart = p.leave2 + jotfshape;
        for (int iz = max(kzstart, 1); iz < nz / 2 + 1; iz++) {
          row[iz] = zero;
          row[nz - iz] = zero;
        }
      }
    } else {
      for (int iz = 0; iz < nz; iz++) row[iz] = zero;
    }
  } else {
    const float lambdaem = lamda / p.nimm;
    const float lambdaexc = 0.88 * lambdaem;
    const float two_over_lambdaem = 2 / lambdaem;
    const float two_over_lambdaexc = 2 / lambdaexc;
    const float beta = (float)asin((double)(k0mag / two_over_lambdaexc));
    int center_of_arc;
    if (order == 0)
      center_of_arc = (int)ceil((double)(two_over_lambdaexc / dkz)) + 3;
    else if (order == 1)
      center_of_arc = (int)ceil(1. / lambdaexc * (cos((double)beta) + 1) / dkz) + 3;
    else
      center_of_arc = (int)ceil((double)two_over_lambdaexc * cos((double)beta) / dkz) + 3;
    if (kr <= krmax) {
      const int kzstart =
          (int)ceil(center_of_arc + sqrt((double)(two_over_lambdaem * two_over_lambdaem - kr * kr)) / dkz);
      const int kzend = nz - kzstart;
      for (int iz = max(kzstart, 0); iz <= kzend && iz < nz; iz++) row[iz] = zero;
    } else {
      for (int iz = 0; iz < nz; iz++) row[iz] = zero;
    }
  }
}

__device__ __forceinline__ void block_sum2(double a, double b, double* ra, double* rb, double& oa, double& ob) {
  const int tid = threadIdx.x;
  ra[tid] = a;
  rb[tid] = b;
  __syncthreads();
  for (int w = kFinishThreads / 2; w > 0; w >>= 1) {
    if (tid < w) {
      ra[tid] += ra[tid + w];
      rb[tid] += rb[tid + w];
    }
    __syncthreads();
  }
  oa = ra[0];
  ob = rb[0];
  __syncthreads();
}

// Post-processing of the radial averages, in the reference's order, by one workgroup.
__global__ __launch_bounds__(kFinishThreads) void k_otf_finish(FinishParams p, float2* __restrict__ avg,
                                                               float dkr_clean, float* __restrict__ info_out) {
#pragma clang fp contract(off)
  __shared__ double ra[kFinishThreads], rb[kFinishThreads];
  __shared__ double fsum[kMaxFixorigin];
  __shared__ float scalefactor, slope_s, avg_s, meani_s;
  const int tid = threadIdx.x;
  const int nz = p.nz, nkr = p.nkr;
  const long long per = (long long)nkr * nz;
  if (tid == 0) scalefactor = 1.0f;
  // force a real origin and Hermitian symmetry along kz (radialft 918-932)
  for (int row = tid; row < p.nphases * nkr; row += kFinishThreads) {
    float2* r = avg + (long long)row * nz;
    r[0].y = 0.f;
    for (int kz = 1; kz <= nz / 2; kz++) {
      const float2 a = r[kz], b = r[nz - kz];
      const float2 m = make_float2((a.x + b.x) / 2.f, (a.y + (-b.y)) / 2.f);
      r[kz] = m;
      r[nz - kz] = make_float2(m.x, -m.y);
    }
  }
  __syncthreads();
  for (int phase = 0; phase < p.nphases; phase++) {
    float2* otf = avg + (long long)phase * per;
    const int order = (phase + 1) / 2;
    // modify (radialft.cpp:1169-1186)
    if (p.ifixz)
      for (int kz = tid; kz < nz; kz += kFinishThreads)
        if (kz != 0 || order > 0) otf[kz] = otf[nz + kz];
    __syncthreads();
    if (p.ifixr)
      for (int kx = p.ifixr + tid; kx < nkr; kx += kFinishThreads) {
        const float2 a = otf[(long long)kx * nz + 1], b = otf[(long long)kx * nz + nz - 1];
        otf[(long long)kx * nz] = make_float2((a.x + b.x) / 2.f, (a.y + b.y) / 2.f);
      }
    __syncthreads();
    if (p.cleanup)
      for (int ix = tid; ix < p.icleanup; ix += kFinishThreads) cleanup_row(otf, ix, order, p, dkr_clean);
    __syncthreads();
    // fixorigin on the order-0 band (radialft.cpp:1188-1224)
    if (phase == 0 && p.kx1 > 0) {
      for (int i = tid; i <= p.kx2; i += kFinishThreads) {
        double s = 0;
        if (i == 0)
          s = otf[0].x;
        else
          for (int j = 0; j < nz; j++) s += otf[(long long)i * nz + j].x;
        fsum[i] = s;
      }
      __syncthreads();
      if (tid == 0) {
        const float meani = 0.5 * (p.kx1 + p.kx2);
        const int numvals = p.kx2 - p.kx1 + 1;
        double totsum = 0, ysum = 0, sqsum = 0;
        for (int i = p.kx1; i <= p.kx2; i++) {
          totsum += fsum[i];
          ysum += fsum[i] * (i - meani);
          sqsum += (i - meani) * (i - meani);
        }
        slope_s = ysum / sqsum;
        avg_s = totsum / numvals;
        meani_s = meani;
      }
      __syncthreads();
      for (int i = tid; i < p.kx1; i += kFinishThreads) {
        const float lineval = avg_s + (i - meani_s) * slope_s;
        otf[(long long)i * nz].x = (float)(otf[(long long)i * nz].x - (fsum[i] - lineval));
      }
      __syncthreads();
    }
    // rescale (radialft.cpp:1226-1242)
    if (tid == 0 && (order == 0 || p.dorescale)) scalefactor = 1 / otf[0].x;
    __syncthreads();
    const float sf = scalefactor;
    for (long long i = tid; i < per; i += kFinishThreads) otf[i] = make_float2(otf[i].x * sf, otf[i].y * sf);
    __syncthreads();
  }
  if (tid == 0) info_out[0] = scalefactor;
  // combine_reim (radialft.cpp:746-782)
  if (p.combine) {
    const long long probe = (long long)min(p.nx / 4, nkr - 1) * nz;
    for (int order = 1; order < p.norders; order++) {
      float2* re = avg + (long long)(2 * order - 1) * per;
      float2* im = avg + (long long)(2 * order) * per;
      double a = 0, b = 0;
      for (long long i = tid; i < per; i += kFinishThreads) {
        a += (double)hypotf(re[i].x, re[i].y);
        b += (double)hypotf(im[i].x, im[i].y);
      }
      double sa, sb;
      block_sum2(a, b, ra, rb, sa, sb);
      double phi = atan(sb / sa);
      const float rr = re[probe].x, ii = im[probe].x;
      if (rr < 0 && ii > 0)
        phi += kPId;
      else if (rr < 0 && ii < 0)
        phi = kPId - phi;
      else if (rr > 0 && ii > 0)
        phi = -phi;
      if (order == 1 && p.forced_pi) phi = kPId - phi;
      if (tid == 0) info_out[order] = (float)phi;
      const float c = (float)cos(phi), s = (float)sin(-phi);
      __syncthreads();
      for (long long i = tid; i < per; i += kFinishThreads) {
        const float2 x = re[i], y = im[i];
        re[i] = make_float2(x.x * c + y.x * s, x.y * c + y.y * s);
        im[i] = make_float2(0.f, 0.f);
      }
      __syncthreads();
    }
  }
  if (p.conjugate)
    for (long long i = tid; i < per * p.nphases; i += kFinishThreads) avg[i].y = -avg[i].y;
}

// fitparabola (radialft.cpp:677-695)
float fit_parabola(float a1, float a2, float a3) {
  const float slope = 0.5 * (a3 - a1);
  const float curve = (a3 + a1) - 2 * a2;
  if (curve == 0) return 0.0f;
  const float peak = -slope / curve;
  if (peak > 1.5 || peak < -1.5) return 0.0f;
  return peak;
}

}  // namespace

void otf_params_default(simr_otf_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->nphases = 5;
  p->k0angle = 1.57f;
  p->linespacing = 0.2f;
  p->bead_diameter = 0.12f;
  p->na = 1.4f;
  p->nimm = 1.515f;
  p->wavelength = 525;
  p->dr = 0.106f;
  p->dz = 0.1f;
  p->background = -1.f;
  p->napodize = 10;
  p->do_compen = 1;
  p->ifixz = 1;
  p->ifixr = 0;
}

size_t otf_elems(const simr_otf_params& p, int nx, int ny, int nsections) {
  if (p.nphases < 1 || nx < 1 || ny < 1) return 0;
  const int nz = nsections / p.nphases;
  const int norders = (p.nphases + 1) / 2;
  const int nbo = p.five_bands ? 2 * norders - 1 : norders;
  return (size_t)nbo * (std::min(nx, ny) / 2 + 1) * (size_t)nz;
}

// i2m: the I2M OTF of radialft.cpp:267-286, 289-307, 345-357, 383-386,
// 425-430 from one I2M section per z (nsec = nz): camera fix only for an
// exact -1, no apodization, its own bead centre, one background (the border
// estimate of section nz/2) for every section, no bead-size compensation, no
// modify or cleanup, fixorigin and the order-0 rescale.
void make_otf(const simr_otf_params& p0, const float* psf, int nx, int ny, int nsec, const float* section_bg,
              float* otf_out, float* band0_out, simr_otf_info* info, bool i2m) {
  simr_otf_params p = p0;
  if (i2m) {
    p.nphases = 1;
    p.napodize = 0;
    p.background = -1.f;
    p.do_compen = 0;
    p.ifixz = p.ifixr = 0;
    p.leavekz[0] = p.leavekz[1] = p.leavekz[2] = 0;
    p.five_bands = 0;
    p.conjugate = 0;
    section_bg = nullptr;
    band0_out = nullptr;
  }
  SIMR_REQUIRE(psf != nullptr && otf_out != nullptr, "null PSF or output pointer");
  SIMR_REQUIRE(p.nphases >= 1 && p.nphases <= SIMR_MAX_PHASES, "makeotf: nphases must be 1..5");
  SIMR_REQUIRE(nx >= 4 && ny >= 4, "makeotf: PSF sections must be at least 4x4");
  const int nphases = p.nphases, nz = nsec / nphases;
  SIMR_REQUIRE(nz >= 1, "makeotf: fewer sections than nphases");
  SIMR_REQUIRE(p.napodize >= -1 && p.napodize <= 64 && 2 * p.napodize <= std::min(nx, ny),
               "makeotf: napodize must be -1..64 and fit the section");
  SIMR_REQUIRE(p.dr > 0 && p.dz > 0, "makeotf: pixel sizes must be positive");
  SIMR_REQUIRE(p.interpkr[0] <= 0 || (p.interpkr[1] >= p.interpkr[0] && p.interpkr[1] < kMaxFixorigin &&
                                      p.interpkr[1] < std::min(nx, ny) / 2 + 1),
               "makeotf: -fixorigin range out of bounds");
  SIMR_REQUIRE(p.ifixr >= 0 && (p.ifixr == 0 || nz >= 2), "makeotf: -ifixkr must be >= 0 and needs a 3D PSF");
  for (int i = 0; i < 3; i++)
    SIMR_REQUIRE(p.leavekz[0] <= 0 || p.leavekz[i] >= 0, "makeotf: -leavekz values must be >= 0");
  const int norders = (nphases + 1) / 2;
  const int nr = std::min(nx, ny), nkr = nr / 2 + 1;
  const long long nxy = (long long)nx * ny, nvox = nxy * nz;
  if (p.leavekz[0] > 0 && !p.coherent_bsim) SIMR_REQUIRE(nz >= 2, "makeotf: -leavekz needs a 3D PSF");

  hipStream_t s;
  SIMR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  } sg{s};

  DevBuf d_psf((size_t)nsec * nxy * 4), d_vol((size_t)nphases * nvox * 4), d_bg((size_t)nsec * 4);
  DevBuf d_spec((size_t)nphases * nvox * 8), d_scratch((size_t)nvox * 8);
  DevBuf d_tab((size_t)(64 + nx + ny) * 4);
  SIMR_HIP(hipMemcpyAsync(d_psf.p, psf, (size_t)nsec * nxy * 4, hipMemcpyHostToDevice, s));

  // edge-blend factors: 1 - sin(((l+0.5)/napodize)*pi/2) in double (radialft.cpp:709,718);
  // cosine apodization factors sin(pi*(k+0.5)/n)
  std::vector<float> tab(64 + nx + ny, 0.f);
  for (int l = 0; l < p.napodize && l < 64; l++)
    tab[l] = (float)(1 - std::sin((((float)l + 0.5) / p.napodize) * kPId * 0.5));
  for (int k = 0; k < nx; k++) tab[64 + k] = (float)std::sin(kPId * ((float)k + 0.5) / nx);
  for (int l = 0; l < ny; l++) tab[64 + nx + l] = (float)std::sin(kPId * ((float)l + 0.5) / ny);
  SIMR_HIP(hipMemcpyAsync(d_tab.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, s));

  int estimate = 0;
  std::vector<float> bgh(nsec, 0.f);
  if (section_bg)
    std::memcpy(bgh.data(), section_bg, (size_t)nsec * 4);
  else if (p.background >= 0)
    std::fill(bgh.begin(), bgh.end(), p.background);
  else
    estimate = 1;
  SIMR_HIP(hipMemcpyAsync(d_bg.p, bgh.data(), (size_t)nsec * 4, hipMemcpyHostToDevice, s));

  const int nused = nz * nphases;  // trailing sections beyond whole z planes are ignored
  const PrepParams pp{nx, ny, nz, nphases, p.napodize, estimate, i2m ? 1 : 0};
  hipLaunchKernelGGL(k_psf_prep, dim3(nused), dim3(1024), 0, s, pp, d_psf.as<float>(), d_vol.as<float>(),
                     d_bg.as<float>(), d_tab.as<float>(), d_tab.as<float>() + 64, d_tab.as<float>() + 64 + nx);
  SIMR_LAUNCH_CHECK();
  if (i2m) {
    // background_i2m: section nz/2's border estimate, subtracted from every section
    float b;
    SIMR_HIP(hipMemcpyAsync(&b, d_bg.as<float>() + nz / 2, 4, hipMemcpyDeviceToHost, s));
    SIMR_HIP(hipStreamSynchronize(s));
    std::fill(bgh.begin(), bgh.end(), b);
    SIMR_HIP(hipMemcpyAsync(d_bg.p, bgh.data(), (size_t)nsec * 4, hipMemcpyHostToDevice, s));
    SIMR_HIP(hipStreamSynchronize(s));
  }

  // bead centre
  const int nparts = (int)std::min<long long>(1024, (nvox + 255) / 256);
  DevBuf d_pv((size_t)nparts * 4), d_pi((size_t)nparts * 8), d_cout(64), d_ciout(64);
  hipLaunchKernelGGL(k_center_partial, dim3(nparts), dim3(256), 0, s, d_vol.as<float>(), nvox, nphases,
                     d_pv.as<float>(), d_pi.as<long long>());
  SIMR_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_center_final, dim3(1), dim3(256), 0, s, d_vol.as<float>(), nx, ny, nz, nphases,
                     d_pv.as<float>(), d_pi.as<long long>(), nparts, d_cout.as<float>(), d_ciout.as<int>());
  SIMR_LAUNCH_CHECK();
  float cv[7];
  int ci[3];
  SIMR_HIP(hipMemcpyAsync(cv, d_cout.p, sizeof(cv), hipMemcpyDeviceToHost, s));
  SIMR_HIP(hipMemcpyAsync(ci, d_ciout.p, sizeof(ci), hipMemcpyDeviceToHost, s));
  SIMR_HIP(hipStreamSynchronize(s));
  const float maxval = cv[0];
  const float zc = nz > 1 ? ci[0] + fit_parabola(cv[1], maxval, cv[2]) : (float)ci[0];
  const float yc = ci[1] + fit_parabola(cv[3], maxval, cv[4]);
  const float xc = ci[2] + fit_parabola(cv[5], maxval, cv[6]);

  // separation (1/nphases-normalised makematrix, radialft.cpp:440-452)
  SepParams sp{};
  sp.nxy = nxy;
  sp.nvox = nvox;
  sp.nphases = nphases;
  const double phi = 2 * kPId / nphases;
  for (int j = 0; j < nphases; j++) {
    sp.M[j * nphases + 0] = (float)(1.0 / nphases);
    for (int order = 1; order < norders; order++) {
      sp.M[j * nphases + 2 * order - 1] = (float)(std::cos(j * order * phi) / nphases);
      sp.M[j * nphases + 2 * order] = (float)(std::sin(j * order * phi) / nphases);
    }
  }
  DevBuf d_band0(band0_out ? (size_t)nvox * 4 : 256);
  hipLaunchKernelGGL(k_psf_separate, dim3(ceil_div(nvox, 256)), dim3(256), 0, s, sp, d_vol.as<float>(),
                     d_bg.as<float>(), d_spec.as<float2>(), band0_out ? d_band0.as<float>() : nullptr);
  SIMR_LAUNCH_CHECK();

  for (int b = 0; b < nphases; b++)
    fft3d_any(d_spec.as<float2>() + (long long)b * nvox, d_scratch.as<float2>(), nx, ny, nz, false, s);

  ShiftParams sh{};
  sh.nx = nx;
  sh.ny = ny;
  sh.nz = nz;
  sh.nphases = nphases;
  sh.do_compen = p.do_compen;
  sh.dphiz = (float)(2 * kPId * zc / nz);
  sh.dphiy = (float)(2 * kPId * yc / ny);
  sh.dphix = (float)(2 * kPId * xc / nx);
  sh.dkx = (float)(1. / (nx * p.dr));
  sh.dky = (float)(1. / (ny * p.dr));
  sh.dkz = (float)(1. / (nz * p.dz));
  sh.radius = (float)(p.bead_diameter * 0.5);
  sh.limit = (float)(4 * kPId * sh.radius * sh.radius * sh.radius / 3);
  const float k0mag = 1 / p.linespacing;
  for (int order = 0; order < norders; order++) {
    const float frac = ((float)order) / (norders - 1) * k0mag;
    sh.k0x[order] = (float)((double)frac * std::cos((double)p.k0angle));
    sh.k0y[order] = (float)((double)frac * std::sin((double)p.k0angle));
  }
  const long long nxh = nx / 2 + 1;
  hipLaunchKernelGGL(k_shift_compensate, dim3(ceil_div(nxh * ny * nz * nphases, 256)), dim3(256), 0, s, sh,
                     d_spec.as<float2>());
  SIMR_LAUNCH_CHECK();

  // radial bins of the half-spectrum pixels, in row-major order (radialft 893-914)
  std::vector<std::vector<int>> bins(nkr);
  {
    const float dkx = 1.f / (nx * p.dr), dky = 1.f / (ny * p.dr), maxK = .5f / p.dr, dkr = 1.f / (nr * p.dr);
    for (int iin = 0; iin < ny; iin++) {
      int ky = iin;
      if (iin > ny / 2) ky -= ny;
      const float ky_f = ky * dky;
      for (int jin = 0; jin < nx / 2 + 1; jin++) {
        const float kx_f = jin * dkx;
        const float rdist = (float)std::sqrt((double)(kx_f * kx_f + ky_f * ky_f));
        if (rdist < maxK) {
          const unsigned indr = (unsigned)std::rint((double)(rdist / dkr));
          SIMR_REQUIRE(indr < (unsigned)nkr, "makeotf: radial bin out of range");
          bins[indr].push_back(iin * nx + jin);
        }
      }
    }
  }
  std::vector<int> off(nkr + 1, 0), pix;
  for (int r = 0; r < nkr; r++) {
    off[r + 1] = off[r] + (int)bins[r].size();
    pix.insert(pix.end(), bins[r].begin(), bins[r].end());
  }
  DevBuf d_off((size_t)(nkr + 1) * 4), d_pix(std::max<size_t>(pix.size(), 1) * 4);
  DevBuf d_avg((size_t)nphases * nkr * nz * 8), d_info(64);
  SIMR_HIP(hipMemcpyAsync(d_off.p, off.data(), off.size() * 4, hipMemcpyHostToDevice, s));
  if (!pix.empty()) SIMR_HIP(hipMemcpyAsync(d_pix.p, pix.data(), pix.size() * 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_radial_avg, dim3(ceil_div((long long)nphases * nkr * nz, 256)), dim3(256), 0, s,
                     d_spec.as<float2>(), nx, ny, nz, nkr, nphases, d_off.as<int>(), d_pix.as<int>(),
                     d_avg.as<float2>());
  SIMR_LAUNCH_CHECK();

  FinishParams fp{};
  fp.nphases = nphases;
  fp.norders = norders;
  fp.nkr = nkr;
  fp.nz = nz;
  fp.nx = nx;
  fp.ifixz = p.ifixz;
  fp.ifixr = p.ifixr;
  fp.cleanup = (!p.coherent_bsim && p.leavekz[0] > 0) ? 1 : 0;
  fp.twolens = p.twolens;
  fp.leave1_1 = p.leavekz[0];
  fp.leave1_2 = p.leavekz[1];
  fp.leave2 = p.leavekz[2];
  fp.icleanup = std::min(nx / 2 + 1, nkr);
  fp.na = p.na;
  fp.nimm = p.nimm;
  fp.linespacing = p.linespacing;
  fp.dr = p.dr;
  fp.dz = p.dz;
  fp.lamdanm = (unsigned)p.wavelength;
  fp.kx1 = p.interpkr[0];
  fp.kx2 = p.interpkr[1];
  fp.dorescale = p.dorescale;
  fp.combine = (!p.five_bands && nphases > 1) ? 1 : 0;
  fp.forced_pi = p.forced_pi_shift;
  fp.conjugate = p.conjugate;
  const float dkr_clean = (float)(1. / (nr * p.dr));
  hipLaunchKernelGGL(k_otf_finish, dim3(1), dim3(kFinishThreads), 0, s, fp, d_avg.as<float2>(),
                     dkr_clean, d_info.as<float>());
  SIMR_LAUNCH_CHECK();

  // output bands: order 0, then each order's combined band (and its im band with five_bands)
  const long long per = (long long)nkr * nz;
  int slot = 0;
  auto take = [&](int band) {
    SIMR_HIP(hipMemcpyAsync(otf_out + 2 * per * slot++, d_avg.as<float2>() + band * per, per * 8,
                            hipMemcpyDeviceToHost, s));
  };
  for (int i = 0; i < norders; i++) {
    if (i == 0) {
      take(0);
    } else {
      take(2 * i - 1);
      if (p.five_bands) take(2 * i);
    }
  }
  float finfo[SIMR_MAX_ORDERS + 1] = {0, 0, 0, 0};
  SIMR_HIP(hipMemcpyAsync(finfo, d_info.p, sizeof(finfo), hipMemcpyDeviceToHost, s));
  if (band0_out) SIMR_HIP(hipMemcpyAsync(band0_out, d_band0.p, (size_t)nvox * 4, hipMemcpyDeviceToHost, s));
  SIMR_HIP(hipStreamSynchronize(s));
  if (info) {
    std::memset(info, 0, sizeof(*info));
    info->nx = nx;
    info->ny = ny;
    info->nz = nz;
    info->norders = norders;
    info->nbands_out = p.five_bands ? 2 * norders - 1 : norders;
    info->nkr = nkr;
    info->dkr = 1 / (nx * p.dr);
    info->dkz = 1 / (nz * p.dz);
    info->xc = xc;
    info->yc = yc;
    info->zc = zc;
    info->scalefactor = finfo[0];
    for (int o = 1; o < norders; o++) info->phi[o] = fp.combine ? finfo[o] : 0.f;
  }
}

}  // namespace simr
