// LDS-staged Stockham autosort FFT for gfx950 (replaces the reference's cuFFT
// calls, src/cudaSirecon/cudaSirecon.cpp:238-257,542-579 and
// src/cudaSirecon/gpuFunctionsImpl.cu:333-349,639-669,1781-1849).
//
// Formulation (self-sorting Stockham, radix R, sub-transform length Ns):
//   for every butterfly j in [0, N/R):  k = j mod Ns
//     v[r] = a[j + r N/R] * w_N^{r k N/(Ns R)},  r = 0..R-1
//     v    = DFT_R(v)
//     b[(j/Ns) Ns R + k + q Ns] = v[q]
// run for the radix sequence of N (8s first, then 4, 2, 5, 3, 7), Ns = product of
// the radices already applied.  Forward uses w_N = exp(-2 pi i / N), inverse
// its conjugate; both unnormalised (as cuFFT).
//
// One workgroup holds W transforms of length N in LDS; TPT threads cooperate
// on each transform.  Everything about N (radices, butterflies per thread,
// twiddle strides) is a compile-time constant, so all loops unroll and the
// butterfly operands stay in VGPRs; each stage is one LDS read sweep, a
// barrier and one LDS write sweep (in place).  Twiddles come from a per-N
// table of exp(-2 pi i m / N), m < N, computed in long double on the host.
#pragma once
#include "common.hpp"

namespace simr {

struct Radices {
  int n;
  int r[16];
};

constexpr Radices make_radices(int N) {
  Radices out{0, {}};
  int m = N;
  while (m % 8 == 0 && m > 8) {
    out.r[out.n++] = 8;
    m /= 8;
  }
  if (m == 8) {
    out.r[out.n++] = 8;
    m = 1;
  }
  if (m % 4 == 0) {
    out.r[out.n++] = 4;
    m /= 4;
  }
  if (m % 2 == 0) {
    out.r[out.n++] = 2;
    m /= 2;
  }
  while (m % 5 == 0) {
    out.r[out.n++] = 5;
    m /= 5;
  }
  while (m % 3 == 0) {
    out.r[out.n++] = 3;
    m /= 3;
  }
  while (m % 7 == 0) {
    out.r[out.n++] = 7;
    m /= 7;
  }
  if (m != 1) out.n = -1;  // unsupported factor
  return out;
}

constexpr int radix_prod_before(const Radices& rd, int s) {
  int p = 1;
  for (int i = 0; i < s; ++i) p *= rd.r[i];
  return p;
}

constexpr int max_radix(const Radices& rd) {
  int m = 1;
  for (int i = 0; i < rd.n; ++i) m = rd.r[i] > m ? rd.r[i] : m;
  return m;
}
constexpr int min_radix(const Radices& rd) {
  int m = 64;
  for (int i = 0; i < rd.n; ++i) m = rd.r[i] < m ? rd.r[i] : m;
  return m;
}

// Threads per transform: a power of two close to N / max_radix, at most 256.
constexpr int fft_tpt(int N) {
  Radices rd = make_radices(N);
  int target = N / max_radix(rd);
  int t = 1;
  while (t * 2 <= target && t * 2 <= 256) t *= 2;
  return t;
}

constexpr bool fft_supported(int N) { return N >= 1 && make_radices(N).n >= 0; }

// ---------------------------------------------------------------- butterflies
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {  // a * (-i) forward, a * (+i) inverse
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2* v);

template <>
__device__ __forceinline__ void dft<1, false>(float2*) {}
template <>
__device__ __forceinline__ void dft<1, true>(float2*) {}

template <int R, bool INV>
__device__ __forceinline__ void dft2(float2* v) {
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
template <>
__device__ __forceinline__ void dft<2, false>(float2* v) { dft2<2, false>(v); }
template <>
__device__ __forceinline__ void dft<2, true>(float2* v) { dft2<2, true>(v); }

template <bool INV>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
  float2 t2 = cadd(a1, a3), t3 = mul_mi<INV>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  float2 x1 = cadd(t1, t3), x3 = csub(t1, t3);
  a1 = x1;
  a3 = x3;
}
template <>
__device__ __forceinline__ void dft<4, false>(float2* v) { dft4<false>(v[0], v[1], v[2], v[3]); }
template <>
__device__ __forceinline__ void dft<4, true>(float2* v) { dft4<true>(v[0], v[1], v[2], v[3]); }

template <bool INV>
__device__ __forceinline__ void dft8(float2* v) {
  // even/odd split: E = DFT4(v0,v2,v4,v6), O = DFT4(v1,v3,v5,v7)
  float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<INV>(e0, e1, e2, e3);
  dft4<INV>(o0, o1, o2, o3);
  const float h = 0.70710678118654752440f;
  // w8^1 = h(1 -+ i), w8^2 = -+i, w8^3 = h(-1 -+ i)
  float2 w1o1 = INV ? make_float2(h * (o1.x - o1.y), h * (o1.x + o1.y))
                    : make_float2(h * (o1.x + o1.y), h * (o1.y - o1.x));
  float2 w2o2 = mul_mi<INV>(o2);
  float2 w3o3 = INV ? make_float2(-h * (o3.x + o3.y), h * (o3.x - o3.y))
                    : make_float2(h * (o3.y - o3.x), -h * (o3.x + o3.y));
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, w1o1);
  v[5] = csub(e1, w1o1);
  v[2] = cadd(e2, w2o2);
  v[6] = csub(e2, w2o2);
  v[3] = cadd(e3, w3o3);
  v[7] = csub(e3, w3o3);
}
template <>
__device__ __forceinline__ void dft<8, false>(float2* v) { dft8<false>(v); }
template <>
__device__ __forceinline__ void dft<8, true>(float2* v) { dft8<true>(v); }

template <bool INV>
__device__ __forceinline__ void dft3(float2* v) {
  const float c = -0.5f;
  const float s = 0.86602540378443864676f;  // sin(2 pi / 3)
  float2 a0 = v[0];
  float2 t1 = cadd(v[1], v[2]);
  float2 t2 = csub(v[1], v[2]);
  float2 m = make_float2(a0.x + c * t1.x, a0.y + c * t1.y);
  // forward: X1 = m - i s t2, X2 = m + i s t2
  float2 ist2 = make_float2(-s * t2.y, s * t2.x);  // i s t2
  v[0] = cadd(a0, t1);
  if (INV) {
    v[1] = cadd(m, ist2);
    v[2] = csub(m, ist2);
  } else {
    v[1] = csub(m, ist2);
    v[2] = cadd(m, ist2);
  }
}
template <>
__device__ __forceinline__ void dft<3, false>(float2* v) { dft3<false>(v); }
template <>
__device__ __forceinline__ void dft<3, true>(float2* v) { dft3<true>(v); }

template <bool INV>
__device__ __forceinline__ void dft5(float2* v) {
  const float c1 = 0.30901699437494742410f;   // cos(2 pi/5)
  const float c2 = -0.80901699437494742410f;  // cos(4 pi/5)
  const float s1 = 0.95105651629515357212f;   // sin(2 pi/5)
  const float s2 = 0.58778525229247312917f;   // sin(4 pi/5)
  float2 a0 = v[0];
  float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  float2 b1 = make_float2(a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y);
  float2 b2 = make_float2(a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y);
  float2 d1 = make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
  float2 d2 = make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
  float2 id1 = make_float2(-d1.y, d1.x), id2 = make_float2(-d2.y, d2.x);
  v[0] = make_float2(a0.x + t1.x + t2.x, a0.y + t1.y + t2.y);
  if (INV) {
    v[1] = cadd(b1, id1);
    v[4] = csub(b1, id1);
    v[2] = cadd(b2, id2);
    v[3] = csub(b2, id2);
  } else {
    v[1] = csub(b1, id1);
    v[4] = cadd(b1, id1);
    v[2] = csub(b2, id2);
    v[3] = cadd(b2, id2);
  }
}
template <>
__device__ __forceinline__ void dft<5, false>(float2* v) { dft5<false>(v); }
template <>
__device__ __forceinline__ void dft<5, true>(float2* v) { dft5<true>(v); }

// 7-point DFT by the symmetric pairs t_k = v_k + v_{7-k}, u_k = v_k - v_{7-k}
// (k = 1..3): X_0 = v_0 + t_1 + t_2 + t_3 and, for m = 1..3,
// a_m = v_0 + sum_k cos(2 pi k m / 7) t_k, b_m = sum_k sin(2 pi k m / 7) u_k,
// X_m = a_m - i b_m, X_{7-m} = a_m + i b_m (forward; the inverse swaps them).
// k m mod 7 folds every angle onto 2 pi j / 7, j = 1..3.
template <bool INV>
__device__ __forceinline__ void dft7(float2* v) {
  const float c1 = 0.62348980185873353053f;   // cos(2 pi/7)
  const float c2 = -0.22252093395631440429f;  // cos(4 pi/7)
  const float c3 = -0.90096886790241912624f;  // cos(6 pi/7)
  const float s1 = 0.78183148246802980871f;   // sin(2 pi/7)
  const float s2 = 0.97492791218182360702f;   // sin(4 pi/7)
  const float s3 = 0.43388373911755812048f;   // sin(6 pi/7)
  const float2 a0 = v[0];
  const float2 t1 = cadd(v[1], v[6]), t2 = cadd(v[2], v[5]), t3 = cadd(v[3], v[4]);
  const float2 u1 = csub(v[1], v[6]), u2 = csub(v[2], v[5]), u3 = csub(v[3], v[4]);
  const float2 a1 = make_float2(a0.x + c1 * t1.x + c2 * t2.x + c3 * t3.x, a0.y + c1 * t1.y + c2 * t2.y + c3 * t3.y);
  const float2 a2 = make_float2(a0.x + c2 * t1.x + c3 * t2.x + c1 * t3.x, a0.y + c2 * t1.y + c3 * t2.y + c1 * t3.y);
  const float2 a3 = make_float2(a0.x + c3 * t1.x + c1 * t2.x + c2 * t3.x, a0.y + c3 * t1.y + c1 * t2.y + c2 * t3.y);
  const float2 b1 = make_float2(s1 * u1.x + s2 * u2.x + s3 * u3.x, s1 * u1.y + s2 * u2.y + s3 * u3.y);
  const float2 b2 = make_float2(s2 * u1.x - s3 * u2.x - s1 * u3.x, s2 * u1.y - s3 * u2.y - s1 * u3.y);
  const float2 b3 = make_float2(s3 * u1.x - s1 * u2.x + s2 * u3.x, s3 * u1.y - s1 * u2.y + s2 * u3.y);
  // i b
  const float2 ib1 = make_float2(-b1.y, b1.x), ib2 = make_float2(-b2.y, b2.x), ib3 = make_float2(-b3.y, b3.x);
  v[0] = make_float2(a0.x + t1.x + t2.x + t3.x, a0.y + t1.y + t2.y + t3.y);
  if (INV) {
    v[1] = cadd(a1, ib1);
    v[6] = csub(a1, ib1);
    v[2] = cadd(a2, ib2);
    v[5] = csub(a2, ib2);
    v[3] = cadd(a3, ib3);
    v[4] = csub(a3, ib3);
  } else {
    v[1] = csub(a1, ib1);
    v[6] = cadd(a1, ib1);
    v[2] = csub(a2, ib2);
    v[5] = cadd(a2, ib2);
    v[3] = csub(a3, ib3);
    v[4] = cadd(a3, ib3);
  }
}
template <>
__device__ __forceinline__ void dft<7, false>(float2* v) { dft7<false>(v); }
template <>
__device__ __forceinline__ void dft<7, true>(float2* v) { dft7<true>(v); }

// ------------------------------------------------------------- LDS layout
// Element i of a transform lives at slot lds_swz<N>(i) of its LDS row.  The
// XOR swizzle permutes within aligned 16-element groups; it makes the
// stride-R writes of the early Stockham stages bank-conflict free (for
// N >= 128 every read and write sweep of every stage is conflict free; a
// plain layout averages 2x and reaches 8x conflicts on those stages).
template <int N>
__host__ __device__ constexpr int lds_swz(int i) {
  if constexpr (N >= 48)
    return i ^ ((i >> 3) & 15);
  else
    return i;
}

// LDS row pitch (float2 elements) of a transform of length N.
constexpr int lds_pitch(int N) {
  if (N >= 48) return ((N + 15) / 16) * 16 + (N < 128 ? 4 : 2);
  return N + 1;
}

// ------------------------------------------------------------- twiddles
// Every stage with Ns > 1 needs B*(R-1) twiddles per thread; they depend only
// on the thread index, so a kernel loads them once (FftTw::load) and reuses
// them for every transform it performs.
constexpr int tw_stage_count(int N, int TPT, int s) {
  Radices rd = make_radices(N);
  int R = rd.r[s];
  int Ns = radix_prod_before(rd, s);
  int M = N / R;
  int B = (M + TPT - 1) / TPT;
  return Ns > 1 ? B * (R - 1) : 0;
}
constexpr int tw_offset(int N, int TPT, int s) {
  int o = 0;
  for (int i = 0; i < s; ++i) o += tw_stage_count(N, TPT, i);
  return o;
}
constexpr int tw_total(int N, int TPT) {
  Radices rd = make_radices(N);
  return tw_offset(N, TPT, rd.n > 0 ? rd.n : 0);
}

template <int N, int TPT = fft_tpt(N)>
struct FftTw {
  static constexpr int kCount = tw_total(N, TPT);
  float2 w[kCount > 0 ? kCount : 1];

  template <int S>
  __device__ __forceinline__ void load_stage(int t, const float2* __restrict__ tab) {
    constexpr Radices RD = make_radices(N);
    if constexpr (S < RD.n) {
      constexpr int R = RD.r[S];
      constexpr int Ns = radix_prod_before(RD, S);
      constexpr int M = N / R;
      constexpr int B = (M + TPT - 1) / TPT;
      constexpr int off = tw_offset(N, TPT, S);
      if constexpr (Ns > 1) {
        constexpr int step = N / (Ns * R);
#pragma unroll
        for (int b = 0; b < B; ++b) {
          const int j = t + b * TPT;
          const int k = j < M ? j % Ns : 0;
#pragma unroll
          for (int r = 1; r < R; ++r) w[off + b * (R - 1) + r - 1] = tab[r * k * step];
        }
      }
      load_stage<S + 1>(t, tab);
    }
  }
  // forward twiddles exp(-2 pi i m / N); inverse transforms conjugate them on use
  __device__ __forceinline__ void load(int t, const float2* __restrict__ tab) {
    if constexpr (N > 1) load_stage<0>(t, tab);
  }
};

// Stage the N-entry twiddle table in LDS (one coalesced read per workgroup)
// and load this thread's twiddles from there.  stage: __shared__ float2[N];
// contains a barrier, so every thread of the workgroup must call it.
template <int N, int TPT>
__device__ __forceinline__ void load_twiddles(FftTw<N, TPT>& tw, float2* stage, int t, const float2* __restrict__ tab,
                                              int tid, int nthreads) {
  if constexpr (N > 1) {
#pragma unroll 4
    for (int i = tid; i < N; i += nthreads) stage[i] = tab[i];
    __syncthreads();
    tw.load(t, stage);
  }
}

// ------------------------------------------------------------- stage driver
template <int N, bool INV, int TPT, int S>
__device__ __forceinline__ void fft_stage(float2* buf, int t, const FftTw<N, TPT>& tw) {
  constexpr Radices RD = make_radices(N);
  if constexpr (S < RD.n) {
    constexpr int R = RD.r[S];
    constexpr int Ns = radix_prod_before(RD, S);
    constexpr int M = N / R;
    constexpr int B = (M + TPT - 1) / TPT;
    constexpr bool FULL = (M % TPT) == 0;
    constexpr int off = tw_offset(N, TPT, S);
    float2 v[B][R];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int j = t + b * TPT;
      if (FULL || j < M) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = buf[lds_swz<N>(j + r * M)];
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int j = t + b * TPT;
      if (FULL || j < M) {
        const int k = j % Ns;
        if constexpr (Ns > 1) {
#pragma unroll
          for (int r = 1; r < R; ++r) {
            float2 ww = tw.w[off + b * (R - 1) + r - 1];
            if (INV) ww.y = -ww.y;
            v[b][r] = cmul(v[b][r], ww);
          }
        }
        dft<R, INV>(v[b]);
        const int idxD = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) buf[lds_swz<N>(idxD + q * Ns)] = v[b][q];
      }
    }
    __syncthreads();
    fft_stage<N, INV, TPT, S + 1>(buf, t, tw);
  }
}

// Same stage with the twiddles read from an N-entry table (LDS or global)
// at use instead of per-thread registers: for kernels whose register budget
// is needed elsewhere.
template <int N, bool INV, int TPT, int S>
__device__ __forceinline__ void fft_stage_t(float2* buf, int t, const float2* __restrict__ tab) {
  constexpr Radices RD = make_radices(N);
  if constexpr (S < RD.n) {
    constexpr int R = RD.r[S];
    constexpr int Ns = radix_prod_before(RD, S);
    constexpr int M = N / R;
    constexpr int B = (M + TPT - 1) / TPT;
    constexpr bool FULL = (M % TPT) == 0;
    float2 v[B][R];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int j = t + b * TPT;
      if (FULL || j < M) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = buf[lds_swz<N>(j + r * M)];
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int j = t + b * TPT;
      if (FULL || j < M) {
        const int k = j % Ns;
        if constexpr (Ns > 1) {
          constexpr int step = N / (Ns * R);
#pragma unroll
          for (int r = 1; r < R; ++r) {
            float2 ww = tab[r * k * step];
            if (INV) ww.y = -ww.y;
            v[b][r] = cmul(v[b][r], ww);
          }
        }
        dft<R, INV>(v[b]);
        const int idxD = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) buf[lds_swz<N>(idxD + q * Ns)] = v[b][q];
      }
    }
    __syncthreads();
    fft_stage_t<N, INV, TPT, S + 1>(buf, t, tab);
  }
}

template <int N, bool INV, int TPT>
__device__ __forceinline__ void fft_lds_t(float2* buf, int t, const float2* __restrict__ tab) {
  static_assert(fft_supported(N), "FFT length must factor into 2, 3, 5, 7");
  if constexpr (N > 1) fft_stage_t<N, INV, TPT, 0>(buf, t, tab);
}

// In-place FFT of one transform held in LDS (element i at buf[lds_swz<N>(i)])
// by its TPT threads t = 0..TPT-1.  Every thread of the workgroup must call it
// (it contains barriers).
template <int N, bool INV, int TPT = fft_tpt(N)>
__device__ __forceinline__ void fft_lds(float2* buf, int t, const FftTw<N, TPT>& tw) {
  static_assert(fft_supported(N), "FFT length must factor into 2, 3, 5, 7");
  if constexpr (N > 1) fft_stage<N, INV, TPT, 0>(buf, t, tw);
}

// Compile-time list of supported lengths; host code dispatches on it.
#define SIMR_FFT_SIZES(X)                                                                    \
  X(2) X(3) X(4) X(5) X(6) X(8) X(9) X(10) X(12) X(15) X(16) X(18) X(20) X(24) X(25) X(27)   \
  X(30) X(32) X(36) X(40) X(45) X(48) X(50) X(54) X(60) X(64) X(72) X(75) X(80) X(81)        \
  X(90) X(96) X(100) X(120) X(128) X(144) X(150) X(160) X(180) X(192) X(200) X(240) X(256)   \
  X(288) X(300) X(320) X(360) X(384) X(400) X(480) X(500) X(512) X(576) X(600) X(640)        \
  X(720) X(768) X(800) X(960) X(1000) X(1024) X(1152) X(1200) X(1280) X(1440) X(1536)        \
  X(1600) X(1920) X(2000) X(2048) X(2400) X(2560) X(3072) X(4096)                              \
  X(7) X(14) X(21) X(28) X(35) X(42) X(49) X(56) X(63) X(70) X(84) X(98) X(105) X(112) X(126)  \
  X(140) X(168) X(196) X(210) X(224)

}  // namespace simr
