// Forward path: flat-field, bleach correction, apodisation, phase separation
// and the 3D forward FFT of every direction's bands.
//
// Reference stages replaced (src/cudaSirecon): load_and_flatfield
// cudaSirecon.cpp:1715-1731, rescale gpuFunctionsImpl.cu:110-195, apodize
// 39-82, transformXYSlice cudaSirecon.cpp:542-579, separate
// gpuFunctionsImpl.cu:214-291, z-FFT cudaSirecon.cpp:238-257.
//
// The reference runs these as ~3000 launches per stack with a synchronous
// host round trip per plane.  Here: one reduction kernel for all section
// sums, one single-thread ratio kernel, and one fused row kernel that reads
// the raw uint16/float sections once and writes the separated bands' x
// transforms; the y and z transforms follow as batched column passes.
// Separation is linear and applied per pixel, so it commutes with the FFT
// and the apodisation; it is applied in real space before the row FFT.
#include "device_math.hpp"
#include "fft.hpp"
#include "fft_api.hpp"
#include "sizes.hpp"
#include "stages.hpp"

namespace simr {

template <typename T>
__device__ __forceinline__ float flat_px(const T* raw, long long sec, int nx, int ny, int y, int x, const float* bg,
                                         const float* slope, float bg0, float inscale) {
#pragma clang fp contract(off)
  const long long e = (long long)y * nx + x;
  float v = (float)raw[sec * (long long)nx * ny + e];
  float b = bg ? bg[e] : bg0;
  float s = slope ? slope[e] : 1.0f;
  return ((v - b) - 0.0f) * s * inscale;
}

// ---------------------------------------------------------------- section sums
template <typename T>
__global__ __launch_bounds__(256) void k_section_sums(const T* __restrict__ raw, const float* __restrict__ bg,
                                                      const float* __restrict__ slope, float bg0, float inscale,
                                                      int nx, int ny, double* __restrict__ sums) {
  __shared__ double red[256];
  const long long sec = blockIdx.x;
  const long long npix = (long long)nx * ny;
  double acc = 0.0;
  for (long long e = threadIdx.x; e < npix; e += 256) {
    const int y = (int)(e / nx), x = (int)(e % nx);
    acc += (double)flat_px(raw, sec, nx, ny, y, x, bg, slope, bg0, inscale);
  }
  double s = block_sum_d<256>(acc, red);
  if (threadIdx.x == 0) sums[sec] = s;
}

void launch_section_sums(const void* raw, bool u16, const float* bg, const float* slope, float inscale, int nx,
                         int ny, int nimg, double* sums, hipStream_t s) {
  const float bg0 = 0.f;  // used only when bg is null
  if (u16)
    hipLaunchKernelGGL(k_section_sums<uint16_t>, dim3(nimg), dim3(256), 0, s, (const uint16_t*)raw, bg, slope, bg0,
                       inscale, nx, ny, sums);
  else
    hipLaunchKernelGGL(k_section_sums<float>, dim3(nimg), dim3(256), 0, s, (const float*)raw, bg, slope, bg0,
                       inscale, nx, ny, sums);
  SIMR_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- ratios
__device__ __forceinline__ int section_of(int dir, int z, int ph, int ndirs, int nz, int nphases, int fastSI) {
  return fastSI ? (z * ndirs * nphases + dir * nphases + ph) : (dir * nz * nphases + z * nphases + ph);
}

// rescale's ratio logic (gpuFunctionsImpl.cu:134-156): per (dir, z) in the
// reference's loop order; sdp persists across time points (equalizet).
__global__ void k_ratios(const double* __restrict__ sums, double* __restrict__ sdp, float* __restrict__ ratio,
                         int ndirs, int nz, int nphases, int fastSI, int it, int equalizez, int equalizet,
                         int do_rescale) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int dir = 0; dir < ndirs; ++dir)
    for (int z = 0; z < nz; ++z) {
      if (!do_rescale) {
        for (int ph = 0; ph < nphases; ++ph) ratio[section_of(dir, z, ph, ndirs, nz, nphases, fastSI)] = 1.0f;
        continue;
      }
      const float s0 = (float)sums[section_of(dir, z, 0, ndirs, nz, nphases, fastSI)];
      if (dir == 0 && !(equalizet && it != 0)) sdp[z] = (double)s0;
      const float ref = equalizez ? (float)sdp[0] : (float)sdp[z];
      for (int ph = 0; ph < nphases; ++ph) {
        const int sec = section_of(dir, z, ph, ndirs, nz, nphases, fastSI);
        ratio[sec] = ref / (float)sums[sec];
      }
    }
}

void launch_ratios(const double* sums, double* sdp, float* ratio, int ndirs, int nz, int nphases, int fastSI,
                   int it, int equalizez, int equalizet, int do_rescale, hipStream_t s) {
  hipLaunchKernelGGL(k_ratios, dim3(1), dim3(64), 0, s, sums, sdp, ratio, ndirs, nz, nphases, fastSI, it,
                     equalizez, equalizet, do_rescale);
  SIMR_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- fused rows
// One workgroup: direction blockIdx.z, plane blockIdx.y, rows y0 = 2*blockIdx.x
// and y0+1.  The two real rows of each band are packed as one complex row
// (row0 + i row1), transformed, and split with
//   X0[k] = (Z[k] + conj Z[N-k]) / 2,   X1[k] = (Z[k] - conj Z[N-k]) / 2i.
template <int N, typename T>
__global__ __launch_bounds__(kMaxBands* fft_tpt(N)) void k_rows_fwd(FwdParams p, const T* __restrict__ raw,
                                                                     const float* __restrict__ bg,
                                                                     const float* __restrict__ slope,
                                                                     const float* __restrict__ ratio,
                                                                     const float* __restrict__ apofact,
                                                                     float2* __restrict__ bands,
                                                                     const float2* __restrict__ tw) {
#pragma clang fp contract(off)
  constexpr int TPT = fft_tpt(N);
  constexpr int NT = kMaxBands * TPT;
  constexpr int P = lds_pitch(N);
  __shared__ float2 cz[kMaxBands * P];
  __shared__ float rv[kMaxBands * 2 * N];
  __shared__ float dy[kMaxBands * 2];
  const int tid = threadIdx.x;
  const int y0 = 2 * blockIdx.x;
  const int z = blockIdx.y, dir = blockIdx.z;
  const int ny = p.ny;
  const int napod = p.napod;
  // 1. flat-field, bleach ratio, x-apodisation (apodize_x_kernel 52-66)
  for (int idx = tid; idx < p.nphases * 2 * N; idx += NT) {
    const int ph = idx / (2 * N);
    const int rem = idx - ph * 2 * N;
    const int r = rem / N, x = rem - (rem / N) * N;
    const int y = y0 + r;
    float v = 0.f;
    if (y < ny) {
      const int sec = section_of(dir, z, ph, p.ndirs, p.nz, p.nphases, p.fastSI);
      const float ra = ratio[sec];
      v = flat_px(raw, sec, N, ny, y, x, bg, slope, 0.f, p.inscale) * ra;
      if (napod > 0 && (y < napod || y >= ny - napod)) {
        const float top = flat_px(raw, sec, N, ny, 0, x, bg, slope, 0.f, p.inscale) * ra;
        const float bot = flat_px(raw, sec, N, ny, ny - 1, x, bg, slope, 0.f, p.inscale) * ra;
        const float diff = (bot - top) * 0.5f;
        for (int l = 0; l < napod; ++l) {
          const float f = apofact[l];
          if (y == l) v += diff * f;
          if (y == ny - 1 - l) v -= diff * f;
        }
      }
    }
    rv[(ph * 2 + r) * N + x] = v;
  }
  __syncthreads();
  // 2. y-apodisation of each row from its own end values (apodize_y_kernel 68-82)
  if (napod > 0) {
    if (tid < p.nphases * 2) dy[tid] = (rv[tid * N + N - 1] - rv[tid * N]) * 0.5f;
    __syncthreads();
    for (int idx = tid; idx < p.nphases * 2 * N; idx += NT) {
      const int row = idx / N, x = idx - (idx / N) * N;
      if (x < napod || x >= N - napod) {
        float v = rv[row * N + x];
        const float d = dy[row];
        for (int k = 0; k < napod; ++k) {
          const float f = apofact[k];
          if (x == k) v += d * f;
          if (x == N - 1 - k) v -= d * f;
        }
        rv[row * N + x] = v;
      }
    }
    __syncthreads();
  }
  // 3. phase separation (separate_kernel 267-291) into packed complex rows
  for (int idx = tid; idx < kMaxBands * N; idx += NT) {
    const int b = idx / N, x = idx - (idx / N) * N;
    float a0 = 0.f, a1 = 0.f;
    if (b < p.nbands) {
      for (int ph = 0; ph < p.nphases; ++ph) {
        const float m = p.sep[b * p.nphases + ph];
        a0 += m * rv[(ph * 2 + 0) * N + x];
        a1 += m * rv[(ph * 2 + 1) * N + x];
      }
    }
    cz[b * P + x] = make_float2(a0, a1);
  }
  __syncthreads();
  fft_lds<N, false>(cz + (tid / TPT) * P, tid % TPT, tw);
  // 4. split the packed transforms into the two rows' half spectra
  const int nxh = N / 2 + 1;
  const bool has1 = (y0 + 1) < ny;
  for (int idx = tid; idx < p.nbands * nxh; idx += NT) {
    const int b = idx / nxh, k = idx - (idx / nxh) * nxh;
    const float2 Z = cz[b * P + k];
    const float2 Zm = cz[b * P + ((N - k) % N)];
    const float2 X0 = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
    const float2 X1 = make_float2(0.5f * (Z.y + Zm.y), 0.5f * (Zm.x - Z.x));
    const long long base = ((long long)(dir * p.nbands + b) * p.nz + z) * ny;
    bands[(base + y0) * nxh + k] = X0;
    if (has1) bands[(base + y0 + 1) * nxh + k] = X1;
  }
}

template <int N>
static void launch_rows_fwd_n(const FwdParams& p, const void* raw, bool u16, const float* bg, const float* slope,
                              const float* ratio, const float* apofact, float2* bands, hipStream_t s) {
  dim3 grid((p.ny + 1) / 2, p.nz, p.ndirs);
  dim3 block(kMaxBands * fft_tpt(N));
  if (u16)
    hipLaunchKernelGGL((k_rows_fwd<N, uint16_t>), grid, block, 0, s, p, (const uint16_t*)raw, bg, slope, ratio,
                       apofact, bands, twiddles(N));
  else
    hipLaunchKernelGGL((k_rows_fwd<N, float>), grid, block, 0, s, p, (const float*)raw, bg, slope, ratio, apofact,
                       bands, twiddles(N));
  SIMR_LAUNCH_CHECK();
}

void launch_rows_fwd(const FwdParams& p, const void* raw, bool u16, const float* bg, const float* slope,
                     const float* ratio, const float* apofact, float2* bands, hipStream_t s) {
  switch (p.nx) {
#define SIMR_CASE(N)                                                                  \
  case N:                                                                             \
    launch_rows_fwd_n<N>(p, raw, u16, bg, slope, ratio, apofact, bands, s);           \
    return;
    SIMR_XY_SIZES(SIMR_CASE)
#undef SIMR_CASE
    default:
      throw ArgError("image width " + std::to_string(p.nx) + " is not a supported size");
  }
}

}  // namespace simr
