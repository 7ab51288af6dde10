// Host-side entry points of the batched FFT kernels (fft_kernels.hip).
#pragma once
#include "common.hpp"

namespace simr {

// Per-device table exp(-2 pi i m / n), m < n (float2), cached for the process lifetime.
const float2* twiddles(int n);
bool fft_size_supported(int n);

// In-place column transforms of length n: element i of column c of plane p at
// data[p * plane_stride + i * pitch + c] for c < ncols, p < nplanes.
void fft_cols(int n, float2* data, long long pitch, long long ncols, long long nplanes, long long plane_stride,
              bool inverse, hipStream_t s);
// In-place row transforms: row r (< nrows) at data[r * pitch], n contiguous elements.
void fft_rows(int n, float2* data, long long pitch, long long nrows, bool inverse, hipStream_t s);
// In-place 3D complex transform of a [nz][ny][nx] array.
void fft3d(float2* data, int nx, int ny, int nz, bool inverse, hipStream_t s);

}  // namespace simr
