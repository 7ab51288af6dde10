// Host-side entry points of the batched FFT kernels (fft_kernels.hip).
#pragma once
#include "common.hpp"

namespace simr {

// Per-device table exp(-2 pi i m / n), m < n (float2), cached for the process
// lifetime.  The first request for a length on a device allocates and fills
// the table (hipMalloc and a blocking copy); contexts request every length
// they use when they are created, so their launch paths only look tables up.
const float2* twiddles(int n);
// The same table computed on the device into caller memory (n float2), on stream s.
void twiddles_into(float2* out, int n, hipStream_t s);
bool fft_size_supported(int n);

// In-place column transforms of length n: element i of column c of plane p at
// data[p * plane_stride + i * pitch + c] for c < ncols, p < nplanes.
// tw: the length's twiddle table (nullptr: the cached twiddles(n)).
void fft_cols(int n, float2* data, long long pitch, long long ncols, long long nplanes, long long plane_stride,
              bool inverse, hipStream_t s, const float2* tw = nullptr);
// In-place row transforms: row r (< nrows) at data[r * pitch], n contiguous elements.
void fft_rows(int n, float2* data, long long pitch, long long nrows, bool inverse, hipStream_t s,
              const float2* tw = nullptr);
// In-place 3D complex transform of a [nz][ny][nx] array.
void fft3d(float2* data, int nx, int ny, int nz, bool inverse, hipStream_t s);

// Direct DFT along one axis for any length (lengths the Stockham kernels do
// not cover).  Line element n of (outer o, inner i) is at
// o*outer_stride + n*n_stride + i, i < ninner, o < nouter; out-of-place.
void dft_lines(const float2* in, float2* out, int n, long long n_stride, long long ninner, long long outer_stride,
               long long nouter, bool inverse, hipStream_t s, const float2* tw = nullptr);
// Direct DFT along a strided axis over ninner contiguous columns per batch
// (nouter batches <= 65535), out of place, float accumulation, any length:
// the z passes for depths without a transform kernel.
void dft_cols(const float2* in, float2* out, int n, long long n_stride, long long ninner, long long outer_stride,
              long long nouter, bool inverse, hipStream_t s);
// 3D transform of any size: Stockham passes where the length is supported,
// direct DFT passes otherwise (scratch: nx*ny*nz elements, used only when
// some length is not supported).  tw[3]: the twiddle tables of nx, ny, nz
// (nullptr: the cached twiddles()).
void fft3d_any(float2* data, float2* scratch, int nx, int ny, int nz, bool inverse, hipStream_t s,
               const float2* const* tw = nullptr);
// True when some length of the 3D transform has no transform kernel, so that
// fft3d_any runs a direct-DFT pass and needs its scratch.
bool fft3d_needs_scratch(int nx, int ny, int nz);

}  // namespace simr
