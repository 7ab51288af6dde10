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

// (raw - background - 0) * slope * inscale (load_and_flatfield, cudaSirecon.cpp:1726)
__device__ __forceinline__ float flat1(float v, float b, float s, float inscale) {
#pragma clang fp contract(off)
  return ((v - b) - 0.0f) * s * inscale;
}

// Four consecutive raw pixels starting at element offset e (e % 4 == 0).
__device__ __forceinline__ float4 load4(const uint16_t* raw, size_t e) {
  const ushort4 u = *reinterpret_cast<const ushort4*>(raw + e);
  return make_float4((float)u.x, (float)u.y, (float)u.z, (float)u.w);
}
__device__ __forceinline__ float4 load4(const float* raw, size_t e) { return *reinterpret_cast<const float4*>(raw + e); }

template <typename T>
__device__ __forceinline__ float4 flat4(const T* raw, size_t sec_off, int pix, const float* bg, const float* slope,
                                        float bg0, float inscale) {
  const float4 v = load4(raw, sec_off + pix);
  float4 b = make_float4(bg0, bg0, bg0, bg0), s = make_float4(1.f, 1.f, 1.f, 1.f);
  if (bg) {
    b = *reinterpret_cast<const float4*>(bg + pix);
    s = *reinterpret_cast<const float4*>(slope + pix);
  }
  return make_float4(flat1(v.x, b.x, s.x, inscale), flat1(v.y, b.y, s.y, inscale), flat1(v.z, b.z, s.z, inscale),
                     flat1(v.w, b.w, s.w, inscale));
}

// ---------------------------------------------------------------- resampling
// One output pixel (xout, y) of section sec: the z index of the section
// (Angle->Z->Phase, or Z->Angle->Phase with fastSI) sets the shift; xin and
// the interpolation weight follow the reference's double/float expression
// order.  in[indin + 1] past the last pixel of the section (the reference
// reads beyond its buffer there) is taken as 0.
template <typename T>
__device__ __forceinline__ float flat_px(const T* raw, size_t off, long long i, const float* bg, const float* slope,
                                         float bg0, float inscale) {
  const float b = bg ? bg[i] : bg0, s = slope ? slope[i] : 1.f;
  return flat1((float)raw[off + i], b, s, inscale);
}

template <typename T>
__global__ __launch_bounds__(256) void k_resample(const T* __restrict__ raw, const float* __restrict__ bg,
                                                  const float* __restrict__ slope, float bg0, float inscale,
                                                  int nx_raw, int nx, int ny, int nz, int ndirs, int nphases,
                                                  int fastSI, float factor, int shift, float* __restrict__ out) {
#pragma clang fp contract(off)
  const int xout = blockIdx.x * 256 + threadIdx.x;
  const int y = blockIdx.y, sec = blockIdx.z;
  if (xout >= nx) return;
  const int z = fastSI ? sec / (ndirs * nphases) : (sec / nphases) % nz;
  float xin = (float)xout;
  if (fabsf(factor) > 0)
    xin = (float)((double)xout - (double)nx / 2.0 + (double)shift - (double)factor * ((double)z - (double)nz / 2.0) +
                  (double)nx_raw / 2.0);
  float v = 0.f;
  if (xin >= 0.f && xin < (float)nx_raw) {
    const float fl = floorf(xin);
    const long long i0 = (long long)y * nx_raw + (long long)fl;
    const float offset = xin - fl;
    const size_t off = (size_t)sec * (size_t)nx_raw * ny;
    const float a = flat_px(raw, off, i0, bg, slope, bg0, inscale);
    const float b = i0 + 1 < (long long)nx_raw * ny ? flat_px(raw, off, i0 + 1, bg, slope, bg0, inscale) : 0.f;
    v = (1 - offset) * a + offset * b;
  }
  out[((size_t)sec * ny + y) * nx + xout] = v;
}

void launch_resample(const void* raw, bool u16, const float* bg, const float* slope, float bg0, float inscale,
                     int nx_raw, int nx, int ny, int nz, int nimg, int ndirs, int nphases, int fastSI, float factor,
                     int shift, float* out, hipStream_t s) {
  if (nimg > 65535 || ny > 65535) throw ArgError("resampling: too many sections or rows");
  const dim3 grid(ceil_div(nx, 256), ny, nimg);
  if (u16)
    hipLaunchKernelGGL(k_resample<uint16_t>, grid, dim3(256), 0, s, (const uint16_t*)raw, bg, slope, bg0, inscale,
                       nx_raw, nx, ny, nz, ndirs, nphases, fastSI, factor, shift, out);
  else
    hipLaunchKernelGGL(k_resample<float>, grid, dim3(256), 0, s, (const float*)raw, bg, slope, bg0, inscale, nx_raw,
                       nx, ny, nz, ndirs, nphases, fastSI, factor, shift, out);
  SIMR_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- section sums
// Block (sec, part) sums part `part` of kSumSplit of section sec in double.
template <typename T>
__global__ __launch_bounds__(256) void k_section_sums(const T* __restrict__ raw, const float* __restrict__ bg,
                                                      const float* __restrict__ slope, float bg0, float inscale,
                                                      int npix, double* __restrict__ partials) {
  __shared__ double red[256];
  const int sec = blockIdx.x, part = blockIdx.y;
  const int nq = npix >> 2;
  const int q0 = (int)((long long)nq * part / kSumSplit), q1 = (int)((long long)nq * (part + 1) / kSumSplit);
  const size_t off = (size_t)sec * npix;
  double acc = 0.0;
#pragma unroll 4
  for (int q = q0 + (int)threadIdx.x; q < q1; q += 256) {
    const float4 v = flat4(raw, off, q << 2, bg, slope, bg0, inscale);
    acc += (double)v.x + (double)v.y + (double)v.z + (double)v.w;
  }
  const double s = block_sum_d<256>(acc, red);
  if (threadIdx.x == 0) partials[sec * kSumSplit + part] = s;
}

void launch_section_sums(const void* raw, bool u16, const float* bg, const float* slope, float bg0, float inscale,
                         int nx, int ny, int nimg, double* partials, hipStream_t s) {
  if ((nx * ny) % 4) throw ArgError("section size must be a multiple of 4 pixels");
  dim3 grid(nimg, kSumSplit);
  if (u16)
    hipLaunchKernelGGL(k_section_sums<uint16_t>, grid, dim3(256), 0, s, (const uint16_t*)raw, bg, slope, bg0,
                       inscale, nx * ny, partials);
  else
    hipLaunchKernelGGL(k_section_sums<float>, grid, dim3(256), 0, s, (const float*)raw, bg, slope, bg0, inscale,
                       nx * ny, partials);
  SIMR_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- ratios
__device__ __forceinline__ int section_of(int dir, int z, int ph, int ndirs, int nz, int nphases, int fastSI) {
  return fastSI ? (z * ndirs * nphases + dir * nphases + ph) : (dir * nz * nphases + z * nphases + ph);
}

__device__ __forceinline__ float section_sum(const double* partials, int sec) {
  double s = 0.0;
  for (int k = 0; k < kSumSplit; ++k) s += partials[sec * kSumSplit + k];
  return (float)s;  // the reference keeps plane sums in float
}

// rescale's ratio logic (gpuFunctionsImpl.cu:134-156).  Each z is independent
// except through sdp[0] under equalizez, so direction 0 sets sdp first (as the
// reference's dir-0 loop does), then every (dir, z) forms its ratios.
__global__ __launch_bounds__(256) void k_ratios(const double* __restrict__ partials, double* __restrict__ sdp,
                                                float* __restrict__ ratio, int ndirs, int nz, int nphases,
                                                int fastSI, int it, int equalizez, int equalizet, int do_rescale) {
  if (do_rescale && !(equalizet && it != 0))
    for (int z = threadIdx.x; z < nz; z += 256)
      sdp[z] = (double)section_sum(partials, section_of(0, z, 0, ndirs, nz, nphases, fastSI));
  __syncthreads();
  for (int idx = threadIdx.x; idx < ndirs * nz; idx += 256) {
    const int dir = idx / nz, z = idx - dir * nz;
    const float ref = equalizez ? (float)sdp[0] : (float)sdp[z];
    for (int ph = 0; ph < nphases; ++ph) {
      const int sec = section_of(dir, z, ph, ndirs, nz, nphases, fastSI);
      ratio[sec] = do_rescale ? ref / section_sum(partials, sec) : 1.0f;
    }
  }
}

void launch_ratios(const double* partials, double* sdp, float* ratio, int ndirs, int nz, int nphases, int fastSI,
                   int it, int equalizez, int equalizet, int do_rescale, hipStream_t s) {
  hipLaunchKernelGGL(k_ratios, dim3(1), dim3(256), 0, s, partials, sdp, ratio, ndirs, nz, nphases, fastSI, it,
                     equalizez, equalizet, do_rescale);
  SIMR_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- fused rows
// One workgroup: direction blockIdx.z, plane blockIdx.y, rows y0 = 2*blockIdx.x
// and y0+1.  The two real rows of each band are packed as one complex row
// (row0 + i row1), transformed, and split with
//   X0[k] = (Z[k] + conj Z[N-k]) / 2,   X1[k] = (Z[k] - conj Z[N-k]) / 2i.
// All raw loads of a workgroup are 4-pixel vector loads with compile-time
// trip counts, so they are all in flight before the first use.
// Threads per row transform: the FFT core's count, halved until the five
// band rows fit one workgroup (2048-point rows take 128 threads, not 256).
constexpr int rows_tpt(int N) {
  int t = fft_tpt(N);
  while (kMaxBands * t > 1024) t /= 2;
  return t;
}

template <int N, typename T>
__global__ __launch_bounds__(kMaxBands* rows_tpt(N)) void k_rows_fwd(FwdParams p, const T* __restrict__ raw,
                                                                     const float* __restrict__ bg,
                                                                     const float* __restrict__ slope,
                                                                     const float* __restrict__ ratio,
                                                                     const float* __restrict__ apofact,
                                                                     float2* __restrict__ bands,
                                                                     const float2* __restrict__ tw) {
#pragma clang fp contract(off)
  constexpr int TPT = rows_tpt(N);
  constexpr int NT = kMaxBands * TPT;
  constexpr int P = lds_pitch(N);
  constexpr int Q = N / 4;
  static_assert((kMaxBands * P + N) * 8 <= 160 * 1024, "band rows and twiddles fit the LDS");
  constexpr int NQ = kMaxBands * 2 * Q;
  constexpr int QPT = (NQ + NT - 1) / NT;
  // the flat-fielded phase rows (rv, 2 kMaxBands rows of N floats) share
  // their LDS with the packed band rows (cz) that the separation writes:
  // every thread reads its columns' phase values before any band is written
  __shared__ __align__(16) float2 cz[kMaxBands * P];
  static_assert(2 * kMaxBands * N <= 2 * kMaxBands * P, "phase rows fit in the band rows");
  float* const rv = reinterpret_cast<float*>(cz);
  __shared__ float dy[kMaxBands * 2];
  const int tid = threadIdx.x;
  FftTw<N, TPT> twr;
  __shared__ float2 twstage[N > 1 ? N : 1];
  load_twiddles(twr, twstage, tid % TPT, tw, threadIdx.x, blockDim.x);
  const int y0 = 2 * blockIdx.x;
  const int z = blockIdx.y, dir = blockIdx.z;
  const int ny = p.ny;
  const int napod = p.napod;
  const size_t npix = (size_t)N * ny;
  const bool edge_block = napod > 0 && (y0 < napod || y0 + 1 >= ny - napod);
  // 1. flat-field, bleach ratio, x-apodisation (apodize_x_kernel 52-66)
#pragma unroll
  for (int i = 0; i < QPT; ++i) {
    const int q = tid + i * NT;
    if (NQ % NT == 0 || q < NQ) {
      const int pr = q / Q, x = (q - (q / Q) * Q) * 4;
      const int ph = pr >> 1, r = pr & 1;
      const int phc = ph < p.nphases ? ph : p.nphases - 1;
      const int y = y0 + r;
      const int yc = y < ny ? y : ny - 1;
      const int sec = section_of(dir, z, phc, p.ndirs, p.nz, p.nphases, p.fastSI);
      const float ra = ratio[sec];
      float4 v = flat4(raw, sec * npix, yc * N + x, bg, slope, p.bg0, p.inscale);
      v = make_float4(v.x * ra, v.y * ra, v.z * ra, v.w * ra);
      if (edge_block && (y < napod || y >= ny - napod)) {
        float4 top = flat4(raw, sec * npix, x, bg, slope, p.bg0, p.inscale);
        float4 bot = flat4(raw, sec * npix, (ny - 1) * N + x, bg, slope, p.bg0, p.inscale);
        const float4 d = make_float4((bot.x * ra - top.x * ra) * 0.5f, (bot.y * ra - top.y * ra) * 0.5f,
                                     (bot.z * ra - top.z * ra) * 0.5f, (bot.w * ra - top.w * ra) * 0.5f);
        for (int l = 0; l < napod; ++l) {
          const float f = apofact[l];
          if (y == l) {
            v.x += d.x * f;
            v.y += d.y * f;
            v.z += d.z * f;
            v.w += d.w * f;
          }
          if (y == ny - 1 - l) {
            v.x -= d.x * f;
            v.y -= d.y * f;
            v.z -= d.z * f;
            v.w -= d.w * f;
          }
        }
      }
      if (ph >= p.nphases || y >= ny) v = make_float4(0.f, 0.f, 0.f, 0.f);
      *reinterpret_cast<float4*>(rv + pr * N + x) = v;
    }
  }
  __syncthreads();
  // 2. y-apodisation of each row from its own end values (apodize_y_kernel 68-82)
  if (napod > 0) {
    if (tid < kMaxBands * 2) dy[tid] = (rv[tid * N + N - 1] - rv[tid * N]) * 0.5f;
    __syncthreads();
    const int ne = 2 * napod;
    for (int idx = tid; idx < kMaxBands * 2 * ne; idx += NT) {
      const int row = idx / ne, kk = idx - (idx / ne) * ne;
      const int x = kk < napod ? kk : N - ne + kk;
      float v = rv[row * N + x];
      const float d = dy[row];
      for (int k = 0; k < napod; ++k) {
        const float f = apofact[k];
        if (x == k) v += d * f;
        if (x == N - 1 - k) v -= d * f;
      }
      rv[row * N + x] = v;  // each element is read and written by this thread only
    }
    __syncthreads();
  }
  // 3. phase separation (separate_kernel 267-291) into packed complex rows:
  // a thread reads the ten phase values of its column x once and forms all
  // bands there
  constexpr int XPT = (N + NT - 1) / NT;
  float r0[XPT][kMaxBands], r1[XPT][kMaxBands];
#pragma unroll
  for (int i = 0; i < XPT; ++i) {
    const int x = tid + i * NT < N ? tid + i * NT : N - 1;
#pragma unroll
    for (int ph = 0; ph < kMaxBands; ++ph) {
      r0[i][ph] = rv[(ph * 2 + 0) * N + x];
      r1[i][ph] = rv[(ph * 2 + 1) * N + x];
    }
  }
  __syncthreads();  // rv is overwritten by cz from here
#pragma unroll
  for (int i = 0; i < XPT; ++i) {
    const int x = tid + i * NT;
    if (N % NT == 0 || x < N) {
#pragma unroll
      for (int b = 0; b < kMaxBands; ++b) {
        float a0 = 0.f, a1 = 0.f;
#pragma unroll
        for (int ph = 0; ph < kMaxBands; ++ph) {
          const float m = p.sep[b * kMaxBands + ph];  // zero outside nbands x nphases
          a0 += m * r0[i][ph];
          a1 += m * r1[i][ph];
        }
        cz[b * P + lds_swz<N>(x)] = make_float2(a0, a1);
      }
    }
  }
  __syncthreads();
  fft_lds<N, false, TPT>(cz + (tid / TPT) * P, tid % TPT, twr);
  // 4. split the packed transforms into the two rows' half spectra
  constexpr int nxh = N / 2 + 1;
  constexpr int UPT = (kMaxBands * nxh + NT - 1) / NT;
  const bool has1 = (y0 + 1) < ny;
#pragma unroll
  for (int i = 0; i < UPT; ++i) {
    const int idx = tid + i * NT;
    const int b = idx / nxh, k = idx - (idx / nxh) * nxh;
    if (b < p.nbands) {
      const float2 Zv = cz[b * P + lds_swz<N>(k)];
      const float2 Zm = cz[b * P + lds_swz<N>((N - k) % N)];
      const float2 X0 = make_float2(0.5f * (Zv.x + Zm.x), 0.5f * (Zv.y - Zm.y));
      const float2 X1 = make_float2(0.5f * (Zv.y + Zm.y), 0.5f * (Zm.x - Zv.x));
      const size_t base = ((size_t)(dir * p.nbands + b) * p.nz + z) * ny;
      bands[(base + y0) * nxh + k] = X0;
      if (has1) bands[(base + y0 + 1) * nxh + k] = X1;
    }
  }
}

template <int N>
static void launch_rows_fwd_n(const FwdParams& p, const void* raw, bool u16, const float* bg, const float* slope,
                              const float* ratio, const float* apofact, float2* bands, hipStream_t s) {
  dim3 grid((p.ny + 1) / 2, p.nz, p.ndirs);
  dim3 block(kMaxBands * rows_tpt(N));
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
