// Two-pass FFT core: N = N1 * N2 with N2 a multiple of N1, TPT = N1 threads
// per transform, N2 values per thread held in registers.
//
//   X[k1 + N1 k2] = sum_{n2} W_N2^{n2 k2} W_N^{n2 k1} sum_{n1} W_N1^{n1 k1} x[N2 n1 + n2]
//
// Pass 1: each thread reads N2/N1 columns n2 of the input (element N2 n1 + n2,
// n1 < N1) from LDS, transforms them with an in-register N1-point DFT and
// multiplies by W_N^{n2 k1} (twiddle table).  The N1 x N2 intermediate goes
// back to LDS transposed, in rows of pitch N2 + 1.  Pass 2: thread k1 reads
// its row and runs an in-register N2-point DFT; X[k1 + N1 k2] lands in
// natural order.
//
// The row is kept in natural order (element i at buf[i]) and the intermediate
// in padded rows, so every LDS address is a per-thread base plus a
// compile-time offset (one register, immediate offsets in the ds
// instructions), and the column reads, transposed writes and row reads are
// all free of bank conflicts (stride N2 + 1 across the threads of a row read).
//
// The in-register DFTs are mixed-radix transforms over compile-time indices;
// their twiddles are compile-time constants (constexpr trigonometry below),
// so they cost no loads and no registers.  Against the one-stage-per-radix LDS
// core this does two LDS round trips instead of one per radix stage, and no
// per-element index arithmetic inside the DFTs.
#pragma once
#include "fft.hpp"

namespace simr {

// ------------------------------------------------------------ constexpr trig
namespace ctrig {
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double sin_poly(double x) {  // |x| <= pi/4
  double x2 = x * x, term = x, sum = x;
  for (int k = 1; k < 14; ++k) {
    term *= -x2 / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cos_poly(double x) {  // |x| <= pi/4
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) {
    term *= -x2 / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
// cos and sin of 2 pi m / n, reduced to the first octant with exact integer
// arithmetic on m.
struct CS {
  double c, s;
};
constexpr CS unit(long long m, long long n) {
  m %= n;
  if (m < 0) m += n;
  // angle = 2 pi m / n = (pi/4) * (8 m / n); octant o = floor(8m/n), rest r in [0,1)
  const long long q = 8 * m;
  const long long o = q / n;
  const double r = (double)(q - o * n) / (double)n;  // fraction of an octant
  const double a = r * (kPi / 4);
  const double ca = cos_poly(a), sa = sin_poly(a);
  const double cb = cos_poly(kPi / 4 - a), sb = sin_poly(kPi / 4 - a);
  switch (o) {
    case 0: return {ca, sa};
    case 1: return {sb, cb};
    case 2: return {-sa, ca};
    case 3: return {-cb, sb};
    case 4: return {-ca, -sa};
    case 5: return {-sb, -cb};
    case 6: return {sa, -ca};
    default: return {cb, -sb};
  }
}
}  // namespace ctrig

// exp(-2 pi i m / n) (forward) or its conjugate, as compile-time floats
template <long long M, long long NN, bool INV>
struct CTw {
  static constexpr ctrig::CS cs = ctrig::unit(M, NN);
  static constexpr float re = (float)cs.c;
  static constexpr float im = INV ? (float)cs.s : (float)-cs.s;
};

template <long long M, long long NN, bool INV>
__device__ __forceinline__ float2 ctw_mul(float2 a) {
  if constexpr (M % NN == 0) {
    return a;
  } else {
    constexpr float wr = CTw<M, NN, INV>::re, wi = CTw<M, NN, INV>::im;
    return make_float2(a.x * wr - a.y * wi, a.x * wi + a.y * wr);
  }
}

// ------------------------------------------------------------ register DFT
// In-place mixed-radix decimation-in-frequency over compile-time indices:
// stage s splits each sub-transform of length m into R = radix(s) parts of
// length L = m / R: a[r] = v[b m + j + r L]; DFT_R; v[b m + j + q L] =
// a[q] * W_m^{j q}.  Every butterfly writes back to the slots it read, so
// the register allocator keeps one value per element; the outputs end in
// digit-reversed slots, undone by renaming (reg_pos).
// v[base + q L] = a[q] * W_m^{J q} (compile-time twiddles)
template <int M, bool INV, int m, int BASE, int L, int J, int R, int q = 0>
__device__ __forceinline__ void dif_store(float2* v, const float2* a) {
  if constexpr (q < R) {
    v[BASE + q * L] = ctw_mul<(long long)J * q, m, INV>(a[q]);
    dif_store<M, INV, m, BASE, L, J, R, q + 1>(v, a);
  }
}

template <int M, bool INV, int S, int B, int J>
__device__ __forceinline__ void dif_bfly(float2* v) {
  constexpr Radices RD = make_radices(M);
  constexpr int R = RD.r[S];
  constexpr int m = M / radix_prod_before(RD, S);
  constexpr int L = m / R;
  if constexpr (B < M / m) {
    if constexpr (J < L) {
      float2 a[R];
#pragma unroll
      for (int r = 0; r < R; ++r) a[r] = v[B * m + J + r * L];
      dft<R, INV>(a);
      dif_store<M, INV, m, B * m + J, L, J, R>(v, a);
      dif_bfly<M, INV, S, B, J + 1>(v);
    } else {
      dif_bfly<M, INV, S, B + 1, 0>(v);
    }
  }
}

// digit-reversed slot of output k for the radix sequence of M from stage S on
constexpr int reg_pos_from(const Radices& rd, int s, int m, int k) {
  if (s >= rd.n) return 0;
  const int R = rd.r[s];
  const int L = m / R;
  return (k % R) * L + reg_pos_from(rd, s + 1, L, k / R);
}
template <int M>
constexpr int reg_pos(int k) {
  return reg_pos_from(make_radices(M), 0, M, k);
}

template <int M, bool INV, int S>
__device__ __forceinline__ void dif_stages(float2* v) {
  constexpr Radices RD = make_radices(M);
  if constexpr (S < RD.n) {
    dif_bfly<M, INV, S, 0, 0>(v);
    dif_stages<M, INV, S + 1>(v);
  }
}

// o[k] = v[reg_pos(k)]; the slot is a template constant so that no part of
// the digit reversal is ever evaluated at run time
template <int M, int K = 0>
__device__ __forceinline__ void reg_gather(float2* o, const float2* v) {
  if constexpr (K < M) {
    constexpr int pos = reg_pos<M>(K);
    o[K] = v[pos];
    reg_gather<M, K + 1>(o, v);
  }
}

// In-place M-point DFT of v[0..M) (natural order in, natural order out).
template <int M, bool INV>
__device__ __forceinline__ void reg_dft(float2* v) {
  if constexpr (M > 1) {
    static_assert(make_radices(M).n > 0, "register DFT length must factor into 2, 3, 5");
    dif_stages<M, INV, 0>(v);
    float2 o[M];
    reg_gather<M>(o, v);
#pragma unroll
    for (int k = 0; k < M; ++k) v[k] = o[k];
  }
}

// ------------------------------------------------------------ plan
struct F2Plan {
  int n1, n2;
};
// N1 = the largest supported factor with N1 | N2 = N/N1 and N2 <= 32.
constexpr F2Plan fft2_plan(int N) {
  if (N % 7 == 0) return {0, 0};  // lengths with a factor 7 run on the radix-stage core
  const int cands[] = {32, 16, 12, 10, 8, 6, 5, 4, 3, 2};
  for (int c : cands) {
    if (N % c) continue;
    const int n2 = N / c;
    if (n2 % c) continue;
    if (n2 > 32) continue;
    if (make_radices(c).n < 0 || make_radices(n2).n < 0) continue;
    return {c, n2};
  }
  return {0, 0};
}
constexpr bool fft2_ok(int N) { return fft2_plan(N).n1 > 0; }
constexpr int fft2_tpt(int N) { return fft2_plan(N).n1; }
// LDS elements one transform needs (the padded intermediate is the larger)
constexpr int fft2_span(int N) { return fft2_plan(N).n1 * (fft2_plan(N).n2 + 1); }
// Whether a length has a plan whose transforms stay inside one wavefront.
template <int N>
constexpr bool use_fft2() {
  return fft2_ok(N) && 64 % fft2_tpt(N) == 0 && fft2_tpt(N) >= 2;
}
// LDS pitch of one transform in a tile of transforms: span rounded up to
// 1 mod 32, so the lanes of a tile row (one element of each of consecutive
// transforms) fall in different banks.
template <int N>
constexpr int fft2_pitch() {
  return fft2_span(N) + ((33 - fft2_span(N) % 32) % 32);
}

// Orders the LDS accesses of one wavefront: its LDS instructions execute in
// issue order, so a fence that the compiler may not move memory accesses
// across is all a wave-private exchange needs.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool WAVE>
__device__ __forceinline__ void fft2_sync() {
  if constexpr (WAVE)
    wave_lds_sync();
  else
    __syncthreads();
}

// Per-thread twiddles of the two-pass core.  Thread t multiplies column
// n2 = t + N1 c by W_N^{n2 k1}, k1 < N1, whatever row it transforms, so the
// factors are built once per kernel: W^{n2 k1} = W^{n2 8a} W^{n2 b} with
// k1 = 8a + b, from 8 + N1/8 table values per column (one product, about one
// ulp more than the table entry, instead of N1 - 1 dependent loads per row).
template <int N>
struct Fft2Tw {
  static constexpr int N1 = fft2_plan(N).n1, MC = fft2_plan(N).n2 / (N1 > 0 ? N1 : 1);
  static constexpr int NB = N1 < 8 ? N1 : 8, NA = (N1 + 7) / 8;
  float2 lo[MC][NB];  // W^{n2 b}
  float2 hi[MC][NA];  // W^{n2 8 a}
  // tab: exp(-2 pi i m / N), m < N (global or LDS)
  __device__ __forceinline__ void load(int t, const float2* __restrict__ tab) {
#pragma unroll
    for (int c = 0; c < MC; ++c) {
      const int n2 = t + N1 * c;
#pragma unroll
      for (int b = 0; b < NB; ++b) lo[c][b] = tab[n2 * b];  // n2 b < N2 N1 = N
#pragma unroll
      for (int a = 0; a < NA; ++a) hi[c][a] = tab[n2 * 8 * a];  // 8a < N1
    }
  }
  template <int C, int K1, bool INV>
  __device__ __forceinline__ float2 get() const {
#pragma clang fp contract(off)
    constexpr int A = K1 / 8, B = K1 % 8;
    float2 w;
    if constexpr (A == 0)
      w = lo[C][B];
    else if constexpr (B == 0)
      w = hi[C][A];
    else
      w = cmul(hi[C][A], lo[C][B]);
    if (INV) w.y = -w.y;
    return w;
  }
};

// v[K1] *= W^{n2 K1} for K1 = 1 .. N1-1 of column slot C
template <int N, bool INV, int C, int K1 = 1>
__device__ __forceinline__ void fft2_twiddle(float2* v, const Fft2Tw<N>& tw) {
  if constexpr (K1 < Fft2Tw<N>::N1) {
    v[K1] = cmul(v[K1], tw.template get<C, K1, INV>());
    fft2_twiddle<N, INV, C, K1 + 1>(v, tw);
  }
}

template <int N, bool INV, int C = 0>
__device__ __forceinline__ void fft2_twiddle_all(float2* v, const Fft2Tw<N>& tw) {
  if constexpr (C < Fft2Tw<N>::MC) {
    fft2_twiddle<N, INV, C>(v + C * Fft2Tw<N>::N1, tw);
    fft2_twiddle_all<N, INV, C + 1>(v, tw);
  }
}

// In-place transform of one length-N row held in LDS in natural order
// (element i at buf[i], fft2_span(N) elements available) by its TPT = N1
// threads t, with the twiddles of thread t (Fft2Tw::load(t, ...)).
// WAVE = false: contains workgroup barriers, every thread of the workgroup
// must call it.  WAVE = true: the caller guarantees that the TPT threads of
// each transform lie in one wavefront and that no other wavefront touches
// buf during the call; the exchanges then synchronise the wave only.
// The same with the twiddles read from the table at use (no twiddle
// registers held across the call; for kernels short of registers).  The
// scheduling fence after each product stops the compiler from hoisting all
// N1 loads ahead of the multiplies, which would double the register
// footprint.
template <int N, bool INV, bool WAVE = false>
__device__ __forceinline__ void fft2_lds(float2* buf, int t, const float2* __restrict__ tab) {
#pragma clang fp contract(off)
  constexpr F2Plan P = fft2_plan(N);
  constexpr int N1 = P.n1, N2 = P.n2, MC = N2 / N1, RP = N2 + 1;
  static_assert(N1 > 0, "no two-pass plan for this length");
  float2 v[N2];
  const float2* col = buf + t;
#pragma unroll
  for (int c = 0; c < MC; ++c)
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) v[c * N1 + n1] = col[N2 * n1 + N1 * c];
  fft2_sync<WAVE>();
  float2* tcol = buf + t;
#pragma unroll
  for (int c = 0; c < MC; ++c) {
    const int n2 = t + N1 * c;
    reg_dft<N1, INV>(v + c * N1);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) {
      float2 w = tab[n2 * k1];  // n2 k1 < N2 N1 = N
      if (INV) w.y = -w.y;
      v[c * N1 + k1] = cmul(v[c * N1 + k1], w);
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) tcol[RP * k1 + N1 * c] = v[c * N1 + k1];
  }
  fft2_sync<WAVE>();
  const float2* row = buf + RP * t;
#pragma unroll
  for (int n2 = 0; n2 < N2; ++n2) v[n2] = row[n2];
  fft2_sync<WAVE>();
  reg_dft<N2, INV>(v);
  float2* out = buf + t;
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2) out[N1 * k2] = v[k2];
  fft2_sync<WAVE>();
}

// As fft2_lds with table twiddles, with pass 1's inputs already in the
// thread's registers, v[c N1 + n1] = x[N2 n1 + N1 c + t] (the row is not read
// from buf; buf only has to be free for this transform when it is called).
template <int N, bool INV, bool WAVE = false>
__device__ __forceinline__ void fft2_lds_from_regs(float2* buf, int t, const float2* __restrict__ tab,
                                                   float2 (&v)[fft2_plan(N).n2]) {
#pragma clang fp contract(off)
  constexpr F2Plan P = fft2_plan(N);
  constexpr int N1 = P.n1, N2 = P.n2, MC = N2 / N1, RP = N2 + 1;
  static_assert(N1 > 0, "no two-pass plan for this length");
  float2* tcol = buf + t;
#pragma unroll
  for (int c = 0; c < MC; ++c) {
    const int n2 = t + N1 * c;
    reg_dft<N1, INV>(v + c * N1);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) {
      float2 w = tab[n2 * k1];
      if (INV) w.y = -w.y;
      v[c * N1 + k1] = cmul(v[c * N1 + k1], w);
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) tcol[RP * k1 + N1 * c] = v[c * N1 + k1];
  }
  fft2_sync<WAVE>();
  const float2* row = buf + RP * t;
#pragma unroll
  for (int n2 = 0; n2 < N2; ++n2) v[n2] = row[n2];
  fft2_sync<WAVE>();
  reg_dft<N2, INV>(v);
  float2* out = buf + t;
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2) out[N1 * k2] = v[k2];
  fft2_sync<WAVE>();
}

template <int N, bool INV, bool WAVE = false>
__device__ __forceinline__ void fft2_lds(float2* buf, int t, const Fft2Tw<N>& tw) {
#pragma clang fp contract(off)
  constexpr F2Plan P = fft2_plan(N);
  constexpr int N1 = P.n1, N2 = P.n2, MC = N2 / N1, RP = N2 + 1;
  static_assert(N1 > 0, "no two-pass plan for this length");
  float2 v[N2];
  // pass 1: columns n2 = t + N1*c (c < MC), N1-point DFTs, twiddles
  const float2* col = buf + t;
#pragma unroll
  for (int c = 0; c < MC; ++c)
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) v[c * N1 + n1] = col[N2 * n1 + N1 * c];
  fft2_sync<WAVE>();
#pragma unroll
  for (int c = 0; c < MC; ++c) reg_dft<N1, INV>(v + c * N1);
  fft2_twiddle_all<N, INV>(v, tw);
  float2* tcol = buf + t;
#pragma unroll
  for (int c = 0; c < MC; ++c)
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) tcol[RP * k1 + N1 * c] = v[c * N1 + k1];
  fft2_sync<WAVE>();
  // pass 2: row k1 = t, N2-point DFT
  const float2* row = buf + RP * t;
#pragma unroll
  for (int n2 = 0; n2 < N2; ++n2) v[n2] = row[n2];
  fft2_sync<WAVE>();
  reg_dft<N2, INV>(v);
  float2* out = buf + t;
#pragma unroll
  for (int k2 = 0; k2 < N2; ++k2) out[N1 * k2] = v[k2];
  fft2_sync<WAVE>();
}

// Register in, register out: pass 1 from the thread's registers
// (v[c N1 + n1] = x[N2 n1 + N1 c + t]), register twiddles, the transposed
// intermediate through buf (fft2_span(N) elements), pass 2 into the
// registers: v[k2] = X[t + N1 k2].  buf is last read by the row reads, so
// the caller passes a barrier before writing it again.
template <int N, bool INV, bool WAVE = false>
__device__ __forceinline__ void fft2_regs(float2* buf, int t, const Fft2Tw<N>& tw, float2 (&v)[fft2_plan(N).n2]) {
#pragma clang fp contract(off)
  constexpr F2Plan P = fft2_plan(N);
  constexpr int N1 = P.n1, N2 = P.n2, MC = N2 / N1, RP = N2 + 1;
  static_assert(N1 > 0, "no two-pass plan for this length");
#pragma unroll
  for (int c = 0; c < MC; ++c) reg_dft<N1, INV>(v + c * N1);
  fft2_twiddle_all<N, INV>(v, tw);
  float2* tcol = buf + t;
#pragma unroll
  for (int c = 0; c < MC; ++c)
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) tcol[RP * k1 + N1 * c] = v[c * N1 + k1];
  fft2_sync<WAVE>();
  const float2* row = buf + RP * t;
#pragma unroll
  for (int n2 = 0; n2 < N2; ++n2) v[n2] = row[n2];
  reg_dft<N2, INV>(v);
}

// As fft2_lds with register twiddles, but the result stays in the caller's
// registers: out[k2] = X[t + N1 k2], k2 < N2 (nothing is written back; buf
// is free again when it returns).
template <int N, bool INV, bool WAVE = false>
__device__ __forceinline__ void fft2_lds_regs(float2* buf, int t, const Fft2Tw<N>& tw,
                                              float2 (&v)[fft2_plan(N).n2]) {
#pragma clang fp contract(off)
  constexpr F2Plan P = fft2_plan(N);
  constexpr int N1 = P.n1, N2 = P.n2, MC = N2 / N1, RP = N2 + 1;
  static_assert(N1 > 0, "no two-pass plan for this length");
  const float2* col = buf + t;
#pragma unroll
  for (int c = 0; c < MC; ++c)
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) v[c * N1 + n1] = col[N2 * n1 + N1 * c];
  fft2_sync<WAVE>();
#pragma unroll
  for (int c = 0; c < MC; ++c) reg_dft<N1, INV>(v + c * N1);
  fft2_twiddle_all<N, INV>(v, tw);
  float2* tcol = buf + t;
#pragma unroll
  for (int c = 0; c < MC; ++c)
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) tcol[RP * k1 + N1 * c] = v[c * N1 + k1];
  fft2_sync<WAVE>();
  const float2* row = buf + RP * t;
#pragma unroll
  for (int n2 = 0; n2 < N2; ++n2) v[n2] = row[n2];
  fft2_sync<WAVE>();
  reg_dft<N2, INV>(v);
}

}  // namespace simr
