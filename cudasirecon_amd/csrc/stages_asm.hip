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

// ---- x pass: row Y (blockIdx.x) of planes [kAsmZc*blockIdx.y, +kAsmZc), bands
// [a0, a1) (<= 8 per launch).  The modulation phasors of the row are computed
// once per workgroup and kept in registers across its planes.
constexpr int kAsmSlots = 8;
constexpr int kAsmZc = 4;

template <int N>
constexpr int asm_x_tpt() {
  return fft_tpt(N) < 128 ? fft_tpt(N) : 128;
}

template <int N>
__global__ __launch_bounds__(kAsmSlots* asm_x_tpt<N>()) void k_asm_x(AsmParams p, int a0, int a1, int accumulate,
                                                                     const float2* __restrict__ tw) {
#pragma clang fp contract(off)
  constexpr int TPT = asm_x_tpt<N>();
  constexpr int NT = kAsmSlots * TPT;
  constexpr int LP = lds_pitch(N);
  constexpr int E = (N + NT - 1) / NT;
  __shared__ float2 lds[kAsmSlots * LP];
  FftTw<N, TPT> twr;
  __shared__ float2 twstage[N > 1 ? N : 1];
  load_twiddles(twr, twstage, threadIdx.x % TPT, tw, threadIdx.x, blockDim.x);
  const int Y = blockIdx.x;
  const int nb = a1 - a0;
  // modulation phasors of the sideband slots for this row (cos_sin_kernel 1942-1954)
  float2 lut[E][kAsmSlots];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int X = threadIdx.x + e * NT;
#pragma unroll
    for (int sl = 0; sl < kAsmSlots; ++sl) {
      lut[e][sl] = make_float2(1.f, 0.f);
      const int a = a0 + sl;
      if (X < N && sl < nb && a > 0) {
        const int no = p.norders - 1;
        const int dir = (a - 1) / no, order = 1 + (a - 1) % no;
        const float k0x = p.k0[dir].x * (float)order, k0y = p.k0[dir].y * (float)order;
        const float s1 = (float)(X - N / 2) * k0x;
        const float s2 = (float)(Y - p.ydim / 2) * k0y;
        const float sum = s1 + s2;
        const float angle = (float)((double)p.fact * 3.14159265358979323846 * (double)sum * (double)p.dxyo);
        float sn, cs;
        sincosf(angle, &sn, &cs);
        lut[e][sl] = make_float2(cs, sn);
      }
    }
  }
  const int zbeg = blockIdx.y * kAsmZc;
  const int zend = min(p.zdim, zbeg + kAsmZc);
  for (int zout = zbeg; zout < zend; ++zout) {
    for (int idx = threadIdx.x; idx < kAsmSlots * N; idx += NT) {
      const int sl = idx / N, X = idx - (idx / N) * N;
      float2 v = make_float2(0.f, 0.f);
      if (sl < nb) {
        int x;
        bool valid = true;
        if (X <= p.nx - p.nx / 2)
          x = X;
        else if (X >= N - (p.nx / 2 - 1))
          x = X - N;
        else {
          x = 0;
          valid = false;
        }
        if (valid)
          v = p.stageY[(((long long)(a0 + sl) * p.zdim + zout) * p.ydim + Y) * p.nx + (x + (p.nx / 2 - 1))];
      }
      lds[sl * LP + lds_swz<N>(X)] = v;
    }
    __syncthreads();
    fft_lds<N, true, TPT>(lds + (threadIdx.x / TPT) * LP, threadIdx.x % TPT, twr);
    float* orow = p.out + ((long long)zout * p.ydim + Y) * N;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int X = threadIdx.x + e * NT;
      if (X < N) {
        float acc = accumulate ? orow[X] : 0.f;
#pragma unroll
        for (int sl = 0; sl < kAsmSlots; ++sl) {
          if (sl < nb) {
            const float2 v = lds[sl * LP + lds_swz<N>(X)];
            if (a0 + sl == 0) {
              acc += v.x;  // write_outbuffer_kernel1
            } else {       // write_outbuffer_kernel2: double intermediate, as in the reference
              const double d = (double)v.x * 2.0 * (double)lut[e][sl].x - (double)v.y * 2.0 * (double)lut[e][sl].y;
              acc = (float)((double)acc + d);
            }
          }
        }
        orow[X] = acc;
      }
    }
    __syncthreads();
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

void launch_asm_x(const AsmParams& p, int a0, int a1, bool accumulate, hipStream_t s) {
  if (a1 - a0 > kAsmSlots) throw ArgError("launch_asm_x: at most 8 bands per launch");
  switch (p.xdim) {
#define SIMR_CASE(N)                                                                                  \
  case N:                                                                                             \
    hipLaunchKernelGGL(k_asm_x<N>, dim3(p.ydim, ceil_div(p.zdim, kAsmZc)), dim3(kAsmSlots * asm_x_tpt<N>()),     \
                       0, s, p, a0, a1,                                                               \
                       (int)accumulate, twiddles(N));                                                 \
    SIMR_LAUNCH_CHECK();                                                                              \
    return;
    SIMR_ZOOM_SIZES(SIMR_CASE)
#undef SIMR_CASE
    default:
      throw ArgError("zoomed width " + std::to_string(p.xdim) + " is not a supported size");
  }
}

}  // namespace simr
