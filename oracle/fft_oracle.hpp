// TEST INFRASTRUCTURE ONLY - the CPU oracle's FFT.  Never linked into the product.
//
// The reference calls cuFFT (call sites src/cudaSirecon/cudaSirecon.cpp:238-257,
// 542-579; src/cudaSirecon/gpuFunctionsImpl.cu:333-349, 639-669, 1781-1849) and
// makeotf calls FFTW (src/otf/radialft.cpp:339-348).  Neither library exists
// here, so the oracle uses this self-contained mixed-radix Stockham FFT written
// from the textbook definition X[k] = sum_n x[n] exp(-+2 pi i n k / N)
// (forward = minus sign, inverse = plus sign, both unnormalised, as cuFFT).
// Twiddles are evaluated in long double and rounded once to T.
// Checked against numpy.fft and a brute-force DFT by tests/test_oracle.py.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace orc {

template <typename T>
struct C2 {
  T x, y;
};

template <typename T>
inline C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
inline C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
inline C2<T> cmulc(C2<T> a, C2<T> b) {
  T re = a.x * b.x - a.y * b.y;
  T im = a.x * b.y + a.y * b.x;
  return {re, im};
}

template <typename T>
struct FFTPlan {
  int n = 0;
  int max_radix = 1;
  std::vector<int> radix;
  std::vector<C2<T>> tw;  // tw[m] = exp(-2 pi i m / n)

  explicit FFTPlan(int n_) : n(n_) {
    if (n < 1) throw std::runtime_error("fft size must be >= 1");
    int m = n;
    const int pref[] = {4, 2, 3, 5, 7};
    while (m > 1) {
      bool found = false;
      for (int r : pref) {
        if (m % r == 0) {
          radix.push_back(r);
          m /= r;
          found = true;
          break;
        }
      }
      if (!found) {  // generic prime radix
        int p = 11;
        while (m % p) p += 2;
        radix.push_back(p);
        m /= p;
      }
    }
    for (int r : radix) max_radix = r > max_radix ? r : max_radix;
    tw.resize(n);
    const long double twopi = 6.283185307179586476925286766559L;
    for (int k = 0; k < n; ++k) {
      long double a = -twopi * (long double)k / (long double)n;
      tw[k].x = (T)cosl(a);
      tw[k].y = (T)sinl(a);
    }
  }

  // In-place transform of one contiguous line; work must hold n elements.
  void exec(C2<T>* data, C2<T>* work, bool inverse) const {
    if (n == 1) return;
    C2<T>* src = data;
    C2<T>* dst = work;
    int Ns = 1;
    // one butterfly's inputs and outputs (a prime radix can be as long as n)
    std::vector<C2<T>> v(max_radix), o(max_radix);
    for (int R : radix) {
      const int m = n / R;
      const int twstep = n / (Ns * R);
      const int rstep = n / R;
      for (int j = 0; j < m; ++j) {
        const int k = j % Ns;
        for (int r = 0; r < R; ++r) {
          C2<T> a = src[j + r * m];
          if (r && k) {
            C2<T> w = tw[(r * k * twstep) % n];
            if (inverse) w.y = -w.y;
            a = cmulc(a, w);
          }
          v[r] = a;
        }
        // DFT of length R
        for (int q = 0; q < R; ++q) {
          C2<T> acc = v[0];
          for (int r = 1; r < R; ++r) {
            C2<T> w = tw[((r * q) % R) * rstep];
            if (inverse) w.y = -w.y;
            acc = cadd(acc, cmulc(v[r], w));
          }
          o[q] = acc;
        }
        const int idxD = (j / Ns) * Ns * R + k;
        for (int q = 0; q < R; ++q) dst[idxD + q * Ns] = o[q];
      }
      C2<T>* t = src;
      src = dst;
      dst = t;
      Ns *= R;
    }
    if (src != data) std::memcpy(data, src, sizeof(C2<T>) * n);
  }
};

// Transform every line of length n along one axis of a strided array.
// lines: number of lines; line_start(l) gives the offset of line l; stride: element stride.
template <typename T, typename StartFn>
void fft_lines(const FFTPlan<T>& plan, C2<T>* base, long long nlines, long long stride,
               StartFn line_start, bool inverse) {
#pragma omp parallel
  {
    std::vector<C2<T>> line(plan.n), work(plan.n);
#pragma omp for schedule(static)
    for (long long l = 0; l < nlines; ++l) {
      C2<T>* p = base + line_start(l);
      for (int i = 0; i < plan.n; ++i) line[i] = p[(long long)i * stride];
      plan.exec(line.data(), work.data(), inverse);
      for (int i = 0; i < plan.n; ++i) p[(long long)i * stride] = line[i];
    }
  }
}

// Full 3D complex transform of a [nz][ny][nx] array (any of the sizes may be 1).
template <typename T>
void fft3d(C2<T>* a, int nx, int ny, int nz, bool inverse) {
  const long long nxy = (long long)nx * ny;
  if (nx > 1) {
    FFTPlan<T> px(nx);
    fft_lines(px, a, (long long)ny * nz, 1, [&](long long l) { return l * nx; }, inverse);
  }
  if (ny > 1) {
    FFTPlan<T> py(ny);
    fft_lines(py, a, (long long)nx * nz, nx,
              [&](long long l) { return (l / nx) * nxy + (l % nx); }, inverse);
  }
  if (nz > 1) {
    FFTPlan<T> pz(nz);
    fft_lines(pz, a, nxy, nxy, [&](long long l) { return l; }, inverse);
  }
}

}  // namespace orc
