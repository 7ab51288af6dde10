// Batched strided FFT kernels built on the LDS Stockham core (fft.hpp), and
// the per-device twiddle tables.
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "fft.hpp"
#include "fft2.hpp"
#include "fft_api.hpp"

namespace simr {

// ------------------------------------------------------------------ twiddles
namespace {
std::mutex g_tw_mu;
std::map<std::pair<int, int>, float2*> g_tw;  // (device, N) -> table, kept for the process lifetime
}  // namespace

const float2* twiddles(int N) {
  int dev = 0;
  SIMR_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto key = std::make_pair(dev, N);
  auto it = g_tw.find(key);
  if (it != g_tw.end()) return it->second;
  std::vector<float2> h(N > 0 ? N : 1);
  const long double twopi = 6.283185307179586476925286766559L;
  for (int m = 0; m < N; ++m) {
    long double a = -twopi * (long double)m / (long double)N;
    h[m] = make_float2((float)cosl(a), (float)sinl(a));
  }
  float2* d = nullptr;
  SIMR_HIP(hipMalloc(&d, sizeof(float2) * h.size()));
  SIMR_HIP(hipMemcpy(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice));
  g_tw[key] = d;
  return d;
}

// exp(-2 pi i m / n) in double (sincospi: the argument is an exact multiple
// of pi up to the rounding of 2m/n), rounded to float.
__global__ __launch_bounds__(256) void k_twiddle_table(float2* __restrict__ out, int n) {
  const int m = blockIdx.x * 256 + threadIdx.x;
  if (m >= n) return;
  double sn, cs;
  sincospi(-2.0 * (double)m / (double)n, &sn, &cs);
  out[m] = make_float2((float)cs, (float)sn);
}

void twiddles_into(float2* out, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_twiddle_table, dim3(ceil_div(n, 256)), dim3(256), 0, s, out, n);
  SIMR_LAUNCH_CHECK();
}

// W transforms per workgroup so that a workgroup has at most 1024 threads and
// at most 80 KB of LDS (two workgroups per CU within gfx950's 160 KB).
template <int N>
constexpr int tile_w() {
  constexpr int tpt = fft_tpt(N);
  int w = 1024 / tpt;
  if (w > 128) w = 128;
  while (w > 1 && w * lds_pitch(N) * 8 > 80 * 1024) w /= 2;
  return w;
}

// ------------------------------------------------------------------ kernels
// Column transforms: element i of column c of plane p at
//   data[p * plane_stride + i * pitch + c],  c < ncols.
// A workgroup transforms W adjacent columns, so each of its global row
// accesses is one contiguous W-element segment.
template <int N, bool INV, int W>
__global__ __launch_bounds__(W* fft_tpt(N)) void k_fft_cols(float2* __restrict__ data, long long pitch,
                                                            long long ncols, long long plane_stride,
                                                            const float2* __restrict__ tw) {
  constexpr int TPT = fft_tpt(N);
  constexpr int NT = W * TPT;
  constexpr int P = lds_pitch(N);
  __shared__ float2 lds[W * P];
  const int tid = threadIdx.x;
  FftTw<N> twr;
  __shared__ float2 twstage[N > 1 ? N : 1];
  load_twiddles(twr, twstage, tid % TPT, tw, threadIdx.x, blockDim.x);
  unsigned bx, by;
  xcd_tile(bx, by);
  const long long c0 = (long long)bx * W;
  float2* base = data + (long long)by * plane_stride;
  // all loads of a thread are issued before the first LDS write (clamped
  // addresses, no branches), so they are in flight together
  constexpr int EPT = (N * W + NT - 1) / NT;
  float2 v[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int idx = tid + k * NT;
    const int i = (idx < N * W ? idx : N * W - 1) / W, w = idx % W;
    const long long c = c0 + w < ncols ? c0 + w : ncols - 1;
    v[k] = base[i * pitch + c];
  }
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int idx = tid + k * NT;
    const int i = idx / W, w = idx % W;
    if (idx < N * W) lds[w * P + lds_swz<N>(i)] = (c0 + w < ncols) ? v[k] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lds<N, INV>(lds + (tid / TPT) * P, tid % TPT, twr);
  for (int idx = tid; idx < N * W; idx += NT) {
    const int i = idx / W, w = idx % W;
    const long long c = c0 + w;
    if (c < ncols) base[i * pitch + c] = lds[w * P + lds_swz<N>(i)];
  }
}

// Row transforms: row r at data[r * pitch], N contiguous elements, r < nrows.
template <int N, bool INV, int W>
__global__ __launch_bounds__(W* fft_tpt(N)) void k_fft_rows(float2* __restrict__ data, long long pitch,
                                                            long long nrows, const float2* __restrict__ tw) {
  constexpr int TPT = fft_tpt(N);
  constexpr int NT = W * TPT;
  constexpr int P = lds_pitch(N);
  __shared__ float2 lds[W * P];
  const int tid = threadIdx.x;
  FftTw<N> twr;
  __shared__ float2 twstage[N > 1 ? N : 1];
  load_twiddles(twr, twstage, tid % TPT, tw, threadIdx.x, blockDim.x);
  const long long r0 = (long long)blockIdx.x * W;
  constexpr int EPT = (N * W + NT - 1) / NT;
  float2 v[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int idx = tid + k * NT < N * W ? tid + k * NT : N * W - 1;
    const int w = idx / N, i = idx % N;
    const long long r = r0 + w < nrows ? r0 + w : nrows - 1;
    v[k] = data[r * pitch + i];
  }
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int idx = tid + k * NT;
    const int w = idx / N, i = idx % N;
    if (idx < N * W) lds[w * P + lds_swz<N>(i)] = (r0 + w < nrows) ? v[k] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lds<N, INV>(lds + (tid / TPT) * P, tid % TPT, twr);
  for (int idx = tid; idx < N * W; idx += NT) {
    const int w = idx / N, i = idx % N;
    const long long r = r0 + w;
    if (r < nrows) data[r * pitch + i] = lds[w * P + lds_swz<N>(i)];
  }
}

// ---------------------------------------------------------- two-pass kernels
// Lengths with a two-pass plan whose transforms do not straddle a wavefront
// (use_fft2) use the register core: each transform's row is kept in natural
// order in LDS at fft2_pitch, the twiddles of each thread are built once in
// registers, and the two exchanges inside a transform synchronise its
// wavefront only.
// W transforms per workgroup: at most 512 threads and 80 KB of LDS, whole waves.
template <int N>
constexpr int tile_w2() {
  constexpr int tpt = fft2_tpt(N) > 0 ? fft2_tpt(N) : 1;
  int w = 512 / tpt;
  if (w > 64) w = 64;
  while (w > 1 && w * fft2_pitch<N>() * 8 > 80 * 1024) w /= 2;
  return w;
}

template <int N, bool INV, int W>
__global__ __launch_bounds__(W* fft2_tpt(N)) void k_fft2_cols(float2* __restrict__ data, long long pitch,
                                                             long long ncols, long long plane_stride,
                                                             const float2* __restrict__ tw) {
  constexpr int TPT = fft2_tpt(N);
  constexpr int NT = W * TPT;
  constexpr int SP = fft2_pitch<N>();
  static_assert(NT % 64 == 0, "whole wavefronts");
  __shared__ float2 lds[W * SP];
  const int tid = threadIdx.x;
  Fft2Tw<N> tw2;
  tw2.load(tid % TPT, tw);
  unsigned bx, by;
  xcd_tile(bx, by);
  const long long c0 = (long long)bx * W;
  float2* base = data + (long long)by * plane_stride;
  constexpr int EPT = N * W / NT;
  float2 v[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int idx = tid + k * NT;
    const int i = idx / W, w = idx % W;
    const long long c = c0 + w < ncols ? c0 + w : ncols - 1;
    v[k] = base[i * pitch + c];
  }
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int idx = tid + k * NT;
    const int i = idx / W, w = idx % W;
    lds[w * SP + i] = (c0 + w < ncols) ? v[k] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft2_lds<N, INV, true>(lds + (tid / TPT) * SP, tid % TPT, tw2);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int idx = tid + k * NT;
    const int i = idx / W, w = idx % W;
    if (c0 + w < ncols) base[i * pitch + c0 + w] = lds[w * SP + i];
  }
}

template <int N, bool INV, int W>
__global__ __launch_bounds__(W* fft2_tpt(N)) void k_fft2_rows(float2* __restrict__ data, long long pitch,
                                                             long long nrows, const float2* __restrict__ tw) {
  constexpr int TPT = fft2_tpt(N);
  constexpr int NT = W * TPT;
  constexpr int SP = fft2_pitch<N>();
  static_assert(NT % 64 == 0, "whole wavefronts");
  __shared__ float2 lds[W * SP];
  const int tid = threadIdx.x;
  Fft2Tw<N> tw2;
  tw2.load(tid % TPT, tw);
  const long long r0 = (long long)blockIdx.x * W;
  constexpr int EPT = N * W / NT;
  float2 v[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int idx = tid + k * NT;
    const int w = idx / N, i = idx % N;
    const long long r = r0 + w < nrows ? r0 + w : nrows - 1;
    v[k] = data[r * pitch + i];
  }
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int idx = tid + k * NT;
    const int w = idx / N, i = idx % N;
    lds[w * SP + i] = (r0 + w < nrows) ? v[k] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft2_lds<N, INV, true>(lds + (tid / TPT) * SP, tid % TPT, tw2);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int idx = tid + k * NT;
    const int w = idx / N, i = idx % N;
    if (r0 + w < nrows) data[(r0 + w) * pitch + i] = lds[w * SP + i];
  }
}

template <int N>
static void launch_cols(float2* data, long long pitch, long long ncols, long long nplanes,
                        long long plane_stride, bool inverse, hipStream_t s, const float2* tw) {
  if (!tw) tw = twiddles(N);
  if constexpr (use_fft2<N>()) {
    constexpr int W = tile_w2<N>();
    if constexpr (W * fft2_tpt(N) % 64 == 0) {
      dim3 grid(ceil_div(ncols, W), (unsigned)nplanes);
      dim3 block(W * fft2_tpt(N));
      if (inverse)
        hipLaunchKernelGGL((k_fft2_cols<N, true, W>), grid, block, 0, s, data, pitch, ncols, plane_stride, tw);
      else
        hipLaunchKernelGGL((k_fft2_cols<N, false, W>), grid, block, 0, s, data, pitch, ncols, plane_stride, tw);
      SIMR_LAUNCH_CHECK();
      return;
    }
  }
  constexpr int W = tile_w<N>();
  dim3 grid(ceil_div(ncols, W), (unsigned)nplanes);
  dim3 block(W * fft_tpt(N));
  if (inverse)
    hipLaunchKernelGGL((k_fft_cols<N, true, W>), grid, block, 0, s, data, pitch, ncols, plane_stride, tw);
  else
    hipLaunchKernelGGL((k_fft_cols<N, false, W>), grid, block, 0, s, data, pitch, ncols, plane_stride, tw);
  SIMR_LAUNCH_CHECK();
}

template <int N>
static void launch_rows(float2* data, long long pitch, long long nrows, bool inverse, hipStream_t s,
                        const float2* tw) {
  if (!tw) tw = twiddles(N);
  if constexpr (use_fft2<N>()) {
    constexpr int W = tile_w2<N>();
    if constexpr (W * fft2_tpt(N) % 64 == 0) {
      dim3 grid(ceil_div(nrows, W));
      dim3 block(W * fft2_tpt(N));
      if (inverse)
        hipLaunchKernelGGL((k_fft2_rows<N, true, W>), grid, block, 0, s, data, pitch, nrows, tw);
      else
        hipLaunchKernelGGL((k_fft2_rows<N, false, W>), grid, block, 0, s, data, pitch, nrows, tw);
      SIMR_LAUNCH_CHECK();
      return;
    }
  }
  constexpr int W = tile_w<N>();
  dim3 grid(ceil_div(nrows, W));
  dim3 block(W * fft_tpt(N));
  if (inverse)
    hipLaunchKernelGGL((k_fft_rows<N, true, W>), grid, block, 0, s, data, pitch, nrows, tw);
  else
    hipLaunchKernelGGL((k_fft_rows<N, false, W>), grid, block, 0, s, data, pitch, nrows, tw);
  SIMR_LAUNCH_CHECK();
}

bool fft_size_supported(int n) {
  switch (n) {
#define SIMR_CASE(N) \
  case N:            \
    return true;
    SIMR_FFT_SIZES(SIMR_CASE)
#undef SIMR_CASE
    case 1:
      return true;
    default:
      return false;
  }
}

void fft_cols(int n, float2* data, long long pitch, long long ncols, long long nplanes, long long plane_stride,
              bool inverse, hipStream_t s, const float2* tw) {
  if (n == 1 || ncols <= 0 || nplanes <= 0) return;
  if (nplanes > 65535) throw ArgError("fft_cols: too many planes");
  switch (n) {
#define SIMR_CASE(N)                                                          \
  case N:                                                                     \
    launch_cols<N>(data, pitch, ncols, nplanes, plane_stride, inverse, s, tw); \
    return;
    SIMR_FFT_SIZES(SIMR_CASE)
#undef SIMR_CASE
    default:
      throw ArgError("FFT length " + std::to_string(n) + " is not supported (must be 2^a 3^b 5^c, <= 4096)");
  }
}

void fft_rows(int n, float2* data, long long pitch, long long nrows, bool inverse, hipStream_t s,
              const float2* tw) {
  if (n == 1 || nrows <= 0) return;
  switch (n) {
#define SIMR_CASE(N)                                     \
  case N:                                                \
    launch_rows<N>(data, pitch, nrows, inverse, s, tw);  \
    return;
    SIMR_FFT_SIZES(SIMR_CASE)
#undef SIMR_CASE
    default:
      throw ArgError("FFT length " + std::to_string(n) + " is not supported (must be 2^a 3^b 5^c, <= 4096)");
  }
}

// Direct DFT: one output element per thread, accumulated in double,
// twiddle index n*k mod N stepped incrementally.
__global__ __launch_bounds__(256) void k_dft_lines(const float2* __restrict__ in, float2* __restrict__ out, int N,
                                                   long long sn, long long ninner, long long so, long long nouter,
                                                   const float2* __restrict__ tw, int inverse) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long total = ninner * N * nouter;
  if (t >= total) return;
  const long long i = t % ninner;
  const int k = (int)((t / ninner) % N);
  const long long o = t / (ninner * N);
  const float2* src = in + o * so + i;
  double re = 0.0, im = 0.0;
  int m = 0;
  for (int n = 0; n < N; ++n) {
    const float2 x = src[(long long)n * sn];
    float2 w = tw[m];
    if (inverse) w.y = -w.y;
    re += (double)x.x * w.x - (double)x.y * w.y;
    im += (double)x.x * w.y + (double)x.y * w.x;
    m += k;
    if (m >= N) m -= N;
  }
  out[o * so + (long long)k * sn + i] = make_float2((float)re, (float)im);
}

void dft_lines(const float2* in, float2* out, int n, long long n_stride, long long ninner, long long outer_stride,
               long long nouter, bool inverse, hipStream_t s, const float2* tw) {
  const long long total = ninner * n * nouter;
  if (total <= 0) return;
  if (ceil_div(total, 256) > 2147483647LL) throw ArgError("dft_lines: problem too large");
  hipLaunchKernelGGL(k_dft_lines, dim3(ceil_div(total, 256)), dim3(256), 0, s, in, out, n, n_stride, ninner,
                     outer_stride, nouter, tw ? tw : twiddles(n), inverse ? 1 : 0);
  SIMR_LAUNCH_CHECK();
}

// Direct DFT along a strided axis for lengths without a transform kernel,
// over contiguous columns: thread i of batch o transforms column i, KB
// outputs at a time with the accumulators in registers, so each input is read
// once per block of KB outputs; the twiddles exp(-+2 pi i m / N) come from an
// LDS table, indexed by k z mod N stepped incrementally (the same index for
// every lane: a broadcast read).  Out of place.
constexpr int kDftColsKB = 16;
__global__ __launch_bounds__(256) void k_dft_cols(const float2* __restrict__ in, float2* __restrict__ out, int N,
                                                  long long sn, long long ninner, long long so,
                                                  const float2* __restrict__ tw, int inverse) {
  extern __shared__ float2 tws[];
  for (int m = threadIdx.x; m < N; m += 256) {
    float2 w = tw[m];
    if (inverse) w.y = -w.y;
    tws[m] = w;
  }
  __syncthreads();
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= ninner) return;
  const float2* src = in + (long long)blockIdx.y * so + i;
  float2* dst = out + (long long)blockIdx.y * so + i;
  for (int k0 = 0; k0 < N; k0 += kDftColsKB) {
    float2 acc[kDftColsKB];
    int m[kDftColsKB], step[kDftColsKB];
#pragma unroll
    for (int j = 0; j < kDftColsKB; ++j) {
      acc[j] = make_float2(0.f, 0.f);
      m[j] = 0;
      step[j] = k0 + j < N ? k0 + j : 0;  // outputs past N stay at index 0 and are not stored
    }
    for (int z = 0; z < N; ++z) {
      const float2 x = src[(long long)z * sn];
#pragma unroll
      for (int j = 0; j < kDftColsKB; ++j) {
        const float2 w = tws[m[j]];
        acc[j].x += x.x * w.x - x.y * w.y;
        acc[j].y += x.x * w.y + x.y * w.x;
        m[j] += step[j];
        if (m[j] >= N) m[j] -= N;
      }
    }
#pragma unroll
    for (int j = 0; j < kDftColsKB; ++j)
      if (k0 + j < N) dst[(long long)(k0 + j) * sn] = acc[j];
  }
}

void dft_cols(const float2* in, float2* out, int n, long long n_stride, long long ninner, long long outer_stride,
              long long nouter, bool inverse, hipStream_t s) {
  if (ninner <= 0 || nouter <= 0 || n <= 0) return;
  if (nouter > 65535 || ceil_div(ninner, 256) > 2147483647LL) throw ArgError("dft_cols: problem too large");
  hipLaunchKernelGGL(k_dft_cols, dim3((unsigned)ceil_div(ninner, 256), (unsigned)nouter), dim3(256),
                     (size_t)n * sizeof(float2), s, in, out, n, n_stride, ninner, outer_stride, twiddles(n),
                     inverse ? 1 : 0);
  SIMR_LAUNCH_CHECK();
}

bool fft3d_needs_scratch(int nx, int ny, int nz) {
  return (nx > 1 && !fft_size_supported(nx)) || (ny > 1 && !fft_size_supported(ny)) ||
         (nz > 1 && !fft_size_supported(nz));
}

void fft3d_any(float2* data, float2* scratch, int nx, int ny, int nz, bool inverse, hipStream_t s,
               const float2* const* tw) {
  const long long nxy = (long long)nx * ny, n = nxy * nz;
  const float2* twx = tw ? tw[0] : nullptr;
  const float2* twy = tw ? tw[1] : nullptr;
  const float2* twz = tw ? tw[2] : nullptr;
  auto dft_pass = [&](int len, long long sn, long long ninner, long long so, long long nouter, const float2* t) {
    if (!scratch) throw ArgError("fft3d_any: a direct-DFT pass needs scratch memory");
    dft_lines(data, scratch, len, sn, ninner, so, nouter, inverse, s, t);
    SIMR_HIP(hipMemcpyAsync(data, scratch, (size_t)n * sizeof(float2), hipMemcpyDeviceToDevice, s));
  };
  if (nx > 1) {
    if (fft_size_supported(nx))
      fft_rows(nx, data, nx, (long long)ny * nz, inverse, s, twx);
    else
      dft_pass(nx, 1, 1, nx, (long long)ny * nz, twx);
  }
  if (ny > 1) {
    if (fft_size_supported(ny))
      fft_cols(ny, data, nx, nx, nz, nxy, inverse, s, twy);
    else
      dft_pass(ny, nx, nx, nxy, nz, twy);
  }
  if (nz > 1) {
    if (fft_size_supported(nz))
      fft_cols(nz, data, nxy, nxy, 1, 0, inverse, s, twz);
    else
      dft_pass(nz, nxy, nxy, 0, 1, twz);
  }
}

void fft3d(float2* data, int nx, int ny, int nz, bool inverse, hipStream_t s) {
  const long long nxy = (long long)nx * ny;
  if (nx > 1) fft_rows(nx, data, nx, (long long)ny * nz, inverse, s);
  if (ny > 1) fft_cols(ny, data, nx, nx, nz, nxy, inverse, s);
  if (nz > 1) fft_cols(nz, data, nxy, nxy, 1, 0, inverse, s);
}

}  // namespace simr
