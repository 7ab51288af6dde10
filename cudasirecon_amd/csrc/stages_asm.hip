// Zoomed band assembly (reference: assemblerealspacebands,
// src/cudaSirecon/gpuFunctionsImpl.cu:1741-1851; move_kernel 1853-1913;
// write_outbuffer_kernel1/2 1914-1940; cos_sin_kernel 1942-1954).
//
// The reference, per direction and order, zero-fills a (zoom nx)(zoom ny)(z_zoom nz)
// complex volume, scatters the half-complex band into it (move_kernel),
// inverse-transforms the whole volume and accumulates Re(.) or
// 2 Re(. e^{i phi}) into the output.  Here the inverse transform is split into
// its three 1D passes and pruned to where data can be non-zero:
//   z pass  : reads the band through move_kernel's index map, only for the
//             nx*ny columns the band occupies;
//   y pass  : zero-padded to zoom*ny, only for the nx occupied x columns;
//   x pass  : zero-padded to zoom*nx, all bands of a row at once, with the
//             modulation and the accumulation into the output fused in, so the
//             output is written once per stack instead of read-modified-written
//             per band.
// The order-0 bands of all directions share one transform (linearity).
#include "device_math.hpp"
#include "fft.hpp"
#include "fft_api.hpp"
#include "sizes.hpp"
#include "stages.hpp"

namespace simr {

// move_kernel's map: occupied output coordinate x = xc - (nx/2 - 1), xc < nx.
__device__ __forceinline__ int z_of_zout(int zout, int nz, int zdim, bool& valid) {
  valid = true;
  if (nz == 1) {
    valid = zout == 0;
    return 0;
  }
  if (zout <= nz - nz / 2) return zout;
  if (zout >= zdim - (nz / 2 - 1)) return zout - zdim;
  valid = false;
  return 0;
}

// Band value placed by move_kernel at lateral (x, y) and axial z (centred coords).
__device__ __forceinline__ float2 moved_value(const AsmParams& p, int a, int x, int y, int z) {
  bool conj = false;
  if (x < 0) {
    x = -x;
    y = -y;
    z = -z;
    conj = true;
  }
  if (y < 0) y += p.ny;
  if (z < 0) z += p.nz;
  const long long indin = ((long long)z * p.ny + y) * p.nxh + x;
  const long long dirstride = (long long)p.nbands * p.band_stride;
  if (a == 0) {
    float2 v = make_float2(0.f, 0.f);
    for (int d = 0; d < p.ndirs; ++d) {
      const float2 b = p.bands[d * dirstride + indin];
      v.x += b.x;
      v.y += b.y;
    }
    if (conj) v.y = -v.y;
    return v;
  }
  const int no = p.norders - 1;
  const int dir = (a - 1) / no, order = 1 + (a - 1) % no;
  const float2* base = p.bands + dir * dirstride;
  float2 re = base[(2 * order - 1) * p.band_stride + indin];
  float2 im = base[(2 * order) * p.band_stride + indin];
  if (conj) {
    re.y = -re.y;
    im.y = -im.y;
  }
  return make_float2(re.x - im.y, re.y + im.x);
}

// ---- z pass: W adjacent xc columns of row yc (blockIdx.y) of band a (blockIdx.z)
template <int N>
__global__ __launch_bounds__(256) void k_asm_z(AsmParams p, const float2* __restrict__ tw) {
  constexpr int TPT = fft_tpt(N);
  constexpr int W = (256 / TPT) > 0 ? 256 / TPT : 1;
  constexpr int NT = W * TPT;
  constexpr int LP = lds_pitch(N);
  __shared__ float2 lds[W * LP];
  FftTw<N> twr;
  __shared__ float2 twstage[N > 1 ? N : 1];
  load_twiddles(twr, twstage, threadIdx.x % TPT, tw, threadIdx.x, blockDim.x);
  const int a = blockIdx.z, yc = blockIdx.y;
  const int xc0 = blockIdx.x * W;
  const int y = yc - (p.ny / 2 - 1);
  for (int idx = threadIdx.x; idx < N * W; idx += NT) {
    const int zout = idx / W, w = idx - (idx / W) * W;
    const int xc = xc0 + w;
    bool valid;
    const int z = z_of_zout(zout, p.nz, N, valid);
    float2 v = make_float2(0.f, 0.f);
    if (valid && xc < p.nx) v = moved_value(p, a, xc - (p.nx / 2 - 1), y, z);
    lds[w * LP + lds_swz<N>(zout)] = v;
  }
  __syncthreads();
  fft_lds<N, true>(lds + (threadIdx.x / TPT) * LP, threadIdx.x % TPT, twr);
  for (int idx = threadIdx.x; idx < N * W; idx += NT) {
    const int zout = idx / W, w = idx - (idx / W) * W;
    const int xc = xc0 + w;
    if (xc < p.nx) p.stageZ[(((long long)a * N + zout) * p.ny + yc) * p.nx + xc] = lds[w * LP + lds_swz<N>(zout)];
  }
}

// ---- y pass: W adjacent xc columns of plane zout (blockIdx.y) of band a (blockIdx.z)
template <int N>
constexpr int asm_y_w() {
  constexpr int tpt = fft_tpt(N);
  int w = 1024 / tpt;
  if (w > 16) w = 16;
  while (w > 1 && w * lds_pitch(N) * 8 > 80 * 1024) w /= 2;
  return w;
}

template <int N>
__global__ __launch_bounds__(asm_y_w<N>() * fft_tpt(N)) void k_asm_y(AsmParams p, const float2* __restrict__ tw) {
  constexpr int TPT = fft_tpt(N);
  constexpr int W = asm_y_w<N>();
  constexpr int NT = W * TPT;
  constexpr int LP = lds_pitch(N);
  __shared__ float2 lds[W * LP];
  FftTw<N> twr;
  __shared__ float2 twstage[N > 1 ? N : 1];
  load_twiddles(twr, twstage, threadIdx.x % TPT, tw, threadIdx.x, blockDim.x);
  const int a = blockIdx.z, zout = blockIdx.y;
  const int xc0 = blockIdx.x * W;
  const float2* src = p.stageZ + ((long long)a * p.zdim + zout) * p.ny * p.nx;
  for (int idx = threadIdx.x; idx < N * W; idx += NT) {
    const int Y = idx / W, w = idx - (idx / W) * W;
    const int xc = xc0 + w;
    int y;
    bool valid = xc < p.nx;
    if (Y <= p.ny - p.ny / 2)
      y = Y;
    else if (Y >= N - (p.ny / 2 - 1))
      y = Y - N;
    else {
      y = 0;
      valid = false;
    }
    lds[w * LP + lds_swz<N>(Y)] = valid ? src[(long long)(y + (p.ny / 2 - 1)) * p.nx + xc] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lds<N, true>(lds + (threadIdx.x / TPT) * LP, threadIdx.x % TPT, twr);
  float2* dst = p.stageY + ((long long)a * p.zdim + zout) * (long long)N * p.nx;
  for (int idx = threadIdx.x; idx < N * W; idx += NT) {
    const int Y = idx / W, w = idx - (idx / W) * W;
    const int xc = xc0 + w;
    if (xc < p.nx) dst[(long long)Y * p.nx + xc] = lds[w * LP + lds_swz<N>(Y)];
  }
}

// ---- x pass: row Y (blockIdx.x) of every output plane.
// A 256-thread workgroup transforms G bands of the row at a time (TPT
// threads each) and walks all bands, so every output element is accumulated
// in registers and stored once.  The modulation phasors of the row (E per
// thread for bands 1..kAsmLutBands) are computed once per workgroup and kept
// in registers for all planes; the FFT twiddles are read from an LDS table.
// Several workgroups share a CU, so the row loads of one overlap the
// transforms of another.
constexpr int kAsmXThreads = 512;
constexpr int kAsmLutBands = 6;  // 1 + ndirs*(norders-1) <= 7 keeps every phasor in registers

template <int N>
constexpr int asm_x_tpt() {
  return fft_tpt(N) >= 128 ? 128 : fft_tpt(N);
}

// modulation phasor of assembly band a (>= 1) at output column X of row Y
// (cos_sin_kernel, gpuFunctionsImpl.cu:1942-1954: float angle, then sin/cos)
template <int N>
__device__ __forceinline__ float2 asm_phasor(const AsmParams& p, int a, int X, int Y) {
#pragma clang fp contract(off)
  const int no = p.norders - 1;
  const int dir = (a - 1) / no, order = 1 + (a - 1) % no;
  const float k0x = p.k0[dir].x * (float)order, k0y = p.k0[dir].y * (float)order;
  const float s1 = (float)(X - N / 2) * k0x;
  const float s2 = (float)(Y - p.ydim / 2) * k0y;
  const float sum = s1 + s2;
  const float angle = (float)((double)p.fact * 3.14159265358979323846 * (double)sum * (double)p.dxyo);
  float sn, cs;
  sincosf(angle, &sn, &cs);
  return make_float2(cs, sn);
}

template <int N>
__global__ __launch_bounds__(kAsmXThreads, 4) void k_asm_x(AsmParams p, const float2* __restrict__ tw) {
#pragma clang fp contract(off)
  constexpr int TPT = asm_x_tpt<N>();
  constexpr int G = kAsmXThreads / TPT;
  constexpr int LP = lds_pitch(N);
  constexpr int E = (N + kAsmXThreads - 1) / kAsmXThreads;
  constexpr bool LUT = E <= 4;
  __shared__ float2 lds[G * LP];
  __shared__ float2 twt[N];
  for (int i = threadIdx.x; i < N; i += kAsmXThreads) twt[i] = tw[i];
  const int Y = blockIdx.x;
  const int g = threadIdx.x / TPT, t = threadIdx.x % TPT;
  const int half = p.nx / 2;
  float2 lut[LUT ? E : 1][LUT ? kAsmLutBands : 1];
  if constexpr (LUT) {
#pragma unroll
    for (int e = 0; e < E; ++e)
#pragma unroll
      for (int b = 0; b < kAsmLutBands; ++b) {
        const int X = threadIdx.x + e * kAsmXThreads;
        lut[e][b] = (X < N && b + 1 < p.nasm) ? asm_phasor<N>(p, b + 1, X, Y) : make_float2(1.f, 0.f);
      }
  }
  __syncthreads();
  for (int zout = 0; zout < p.zdim; ++zout) {
    float acc[E];
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = 0.f;
#pragma nounroll
    for (int a0 = 0; a0 < p.nasm; a0 += G) {
      // G bands of row Y of plane zout, zero-padded to N, into LDS
      for (int idx = threadIdx.x; idx < G * N; idx += kAsmXThreads) {
        const int sl = idx / N, X = idx - sl * N;
        const int a = a0 + sl;
        float2 v = make_float2(0.f, 0.f);
        if (a < p.nasm) {
          int x;
          bool valid = true;
          if (X <= p.nx - half)
            x = X;
          else if (X >= N - (half - 1))
            x = X - N;
          else {
            x = 0;
            valid = false;
          }
          if (valid) v = p.stageY[(((long long)a * p.zdim + zout) * p.ydim + Y) * p.nx + (x + (half - 1))];
        }
        lds[sl * LP + lds_swz<N>(X)] = v;
      }
      __syncthreads();
      fft_lds_t<N, true, TPT>(lds + g * LP, t, twt);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int X = threadIdx.x + e * kAsmXThreads;
        if (X < N) {
#pragma unroll
          for (int sl = 0; sl < G; ++sl) {
            const int a = a0 + sl;
            if (a < p.nasm) {
              const float2 v = lds[sl * LP + lds_swz<N>(X)];
              if (a == 0) {
                acc[e] += v.x;  // write_outbuffer_kernel1
              } else {          // write_outbuffer_kernel2: double intermediate, as in the reference
                float2 ph;
                if (LUT && a <= kAsmLutBands) {
                  ph = lut[LUT ? e : 0][0];
#pragma unroll
                  for (int b = 1; b < kAsmLutBands; ++b)
                    if (a - 1 == b) ph = lut[LUT ? e : 0][LUT ? b : 0];
                } else {
                  ph = asm_phasor<N>(p, a, X, Y);
                }
                const double d = (double)v.x * 2.0 * (double)ph.x - (double)v.y * 2.0 * (double)ph.y;
                acc[e] = (float)((double)acc[e] + d);
              }
            }
          }
        }
      }
      __syncthreads();
    }
    float* orow = p.out + ((long long)zout * p.ydim + Y) * N;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int X = threadIdx.x + e * kAsmXThreads;
      if (X < N) orow[X] = acc[e];
    }
  }
}

void launch_asm_z(const AsmParams& p, hipStream_t s) {
  switch (p.zdim) {
#define SIMR_CASE(N)                                                                              \
  case N: {                                                                                       \
    constexpr int W = (256 / fft_tpt(N)) > 0 ? 256 / fft_tpt(N) : 1;                              \
    hipLaunchKernelGGL(k_asm_z<N>, dim3(ceil_div(p.nx, W), p.ny, p.nasm), dim3(256), 0, s, p,     \
                       twiddles(N));                                                              \
    SIMR_LAUNCH_CHECK();                                                                          \
    return;                                                                                       \
  }
    SIMR_Z_SIZES(SIMR_CASE)
#undef SIMR_CASE
    default:
      throw ArgError("zoomed stack depth " + std::to_string(p.zdim) + " is not a supported size");
  }
}

void launch_asm_y(const AsmParams& p, hipStream_t s) {
  switch (p.ydim) {
#define SIMR_CASE(N)                                                                                \
  case N: {                                                                                         \
    constexpr int W = asm_y_w<N>();                                                                 \
    hipLaunchKernelGGL(k_asm_y<N>, dim3(ceil_div(p.nx, W), p.zdim, p.nasm), dim3(W * fft_tpt(N)),   \
                       0, s, p, twiddles(N));                                                       \
    SIMR_LAUNCH_CHECK();                                                                            \
    return;                                                                                         \
  }
    SIMR_ZOOM_SIZES(SIMR_CASE)
#undef SIMR_CASE
    default:
      throw ArgError("zoomed height " + std::to_string(p.ydim) + " is not a supported size");
  }
}

void launch_asm_x(const AsmParams& p, hipStream_t s) {
  switch (p.xdim) {
#define SIMR_CASE(N)                                                                                       \
  case N:                                                                                                  \
    hipLaunchKernelGGL(k_asm_x<N>, dim3(p.ydim), dim3(kAsmXThreads), 0, s, p,                          \
                       twiddles(N));                                                                       \
    SIMR_LAUNCH_CHECK();                                                                                   \
    return;
    SIMR_ZOOM_SIZES(SIMR_CASE)
#undef SIMR_CASE
    default:
      throw ArgError("zoomed width " + std::to_string(p.xdim) + " is not a supported size");
  }
}

}  // namespace simr
