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
#include "fft2.hpp"
#include "fft_api.hpp"
#include "sizes.hpp"
#include "stages.hpp"

namespace simr {

// Whether centred column xc (lateral frequency (xc - (nx/2 - 1)) dkx) lies
// beyond the filter's support for every y, so that every value of it, in
// every stage of the assembly, is zero.  The y pass writes zeros for tiles of
// such columns without reading or transforming them.
__device__ __forceinline__ bool asm_col_dead(const AsmParams& p, int xc) {
  const float kx = (float)(xc - (p.nx / 2 - 1)) * p.dkx;
  return p.rc2_zero > 0.f && kx * kx > p.rc2_zero;
}

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
  // Band value that move_kernel places at lateral (x, y) and axial z (centred
  // coordinates; x < 0 reads the Hermitian mirror, conjugated).  Every
  // address is computed first (invalid elements point at element 0 and are
  // zeroed afterwards), so all loads of a thread are in flight together.
  constexpr int B = (N * W + NT - 1) / NT;
  long long ind[B];
  unsigned okmask = 0, cjmask = 0;
  static_assert(B <= 32, "per-thread element masks hold 32 bits");
#pragma unroll
  for (int k = 0; k < B; ++k) {
    const int idx = threadIdx.x + k * NT;
    const int zout = idx / W, w = idx - (idx / W) * W;
    const int xc = xc0 + w;
    bool valid;
    int z = z_of_zout(zout, p.nz, N, valid);
    valid = valid && xc < p.nx && idx < N * W;
    int x = xc - (p.nx / 2 - 1), yy = y;
    const bool cj = x < 0;
    if (cj) {
      x = -x;
      yy = -yy;
      z = -z;
    }
    if (yy < 0) yy += p.ny;
    if (z < 0) z += p.nz;
    ind[k] = valid ? ((long long)z * p.ny + yy) * p.nxh + x : 0;
    okmask |= (valid ? 1u : 0u) << k;
    cjmask |= (cj ? 1u : 0u) << k;
  }
  float2 v[B];
  const long long dirstride = (long long)p.nbands * p.band_stride;
  if (a == 0) {  // band 0: sum over directions (move_kernel adds the order-0 bands)
#pragma unroll
    for (int k = 0; k < B; ++k) v[k] = make_float2(0.f, 0.f);
    for (int d = 0; d < p.ndirs; ++d) {
      const float2* bd = p.bands + d * dirstride;
#pragma unroll
      for (int k = 0; k < B; ++k) {
        const float2 b = bd[ind[k]];
        v[k].x += b.x;
        v[k].y += b.y;
      }
    }
#pragma unroll
    for (int k = 0; k < B; ++k)
      if ((cjmask >> k) & 1u) v[k].y = -v[k].y;
  } else {  // side band: re + i im of the order's two real-space bands
    const int no = p.norders - 1;
    const int dir = (a - 1) / no, order = 1 + (a - 1) % no;
    const float2* bre = p.bands + dir * dirstride + (2 * order - 1) * p.band_stride;
    const float2* bim = p.bands + dir * dirstride + (2 * order) * p.band_stride;
#pragma unroll
    for (int k = 0; k < B; ++k) {
      float2 re = bre[ind[k]], im = bim[ind[k]];
      if ((cjmask >> k) & 1u) {
        re.y = -re.y;
        im.y = -im.y;
      }
      v[k] = make_float2(re.x - im.y, re.y + im.x);
    }
  }
#pragma unroll
  for (int k = 0; k < B; ++k) {
    const int idx = threadIdx.x + k * NT;
    const int zout = idx / W, w = idx - (idx / W) * W;
    if (idx < N * W) lds[w * LP + lds_swz<N>(zout)] = ((okmask >> k) & 1u) ? v[k] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lds<N, true>(lds + (threadIdx.x / TPT) * LP, threadIdx.x % TPT, twr);
  for (int idx = threadIdx.x; idx < N * W; idx += NT) {
    const int zout = idx / W, w = idx - (idx / W) * W;
    const int xc = xc0 + w;
    if (xc < p.nx) p.stageZ[(((long long)a * N + zout) * p.ny + yc) * p.nx + xc] = lds[w * LP + lds_swz<N>(zout)];
  }
}

// ---- z pass, Hermitian-paired (zdim == nz, even nx and ny).
// An output column with x < 0 holds conj(B(-x, -y, -z)), and with zdim == nz
// the z map is the identity mod nz, so its inverse transform is the conjugate
// of the stored column's: one transform of stored column (x, s) gives output
// (x, s) and, conjugated, output (-x, -s).  A workgroup therefore reads the
// stored columns x in [x0, x0+W) of stored rows s and -s once (the other
// kernel reads every mirrored element a second time) and writes the four
// output columns they feed.  Side band a: value re + i im, so re and im are
// transformed separately: direct R + iI, mirror conj(R) + i conj(I).
// Band 0 (sum over directions): direct B, mirror conj(B).
// threads per workgroup of the paired z pass: W = threads / (slots * TPT)
// stored columns, so each plane's read of a stored row is W * 8 bytes.
// Side bands (four slots) take 512 threads, eight columns; band 0 (two
// slots) 256 threads, also eight columns.
template <bool B0>
constexpr int asm_zp_threads() {
  return B0 ? 256 : 512;
}
template <int N, bool B0>
constexpr int asm_zp_slots() {
  return B0 ? 2 : 4;  // transforms per stored column: rows s, -s (x re, im)
}
template <int N, bool B0>
constexpr int asm_zp_w() {
  return asm_zp_threads<B0>() / (asm_zp_slots<N, B0>() * fft_tpt(N));
}
template <int N>
constexpr bool asm_zp_ok() {
  return N > 1 && 4 * fft_tpt(N) <= 256;
}

template <int N, bool B0>
__global__ __launch_bounds__(asm_zp_threads<B0>()) void k_asm_z_pair(AsmParams p, const float2* __restrict__ tw) {
  constexpr int TPT = fft_tpt(N);
  constexpr int S = asm_zp_slots<N, B0>();
  constexpr int W = asm_zp_w<N, B0>();
  constexpr int NT = S * W * TPT;
  static_assert(NT == asm_zp_threads<B0>(), "one transform slot per TPT threads");
  constexpr int LP = lds_pitch(N);
  __shared__ float2 lds[S * W * LP];
  FftTw<N> twr;
  __shared__ float2 twstage[N > 1 ? N : 1];
  unsigned bx, by, bz;
  xcd_tile3(bx, by, bz);
  const int a = B0 ? 0 : 1 + (int)bz;
  const int s = (int)by, sp = s == 0 ? 0 : p.ny - s;  // stored rows s and -s
  const int x0 = (int)bx * W;
  const int xmax = p.nx / 2;  // stored columns 0 .. nx/2 feed the output
  const int ycs = s + p.ny / 2 - 1;                  // output row of centred y = s
  const int ycp = (s == 0 ? 0 : -s) + p.ny / 2 - 1;  // output row of centred y = -s
  const int ycp2 = s == p.ny / 2 ? ycs : ycp;        // s = ny/2 is its own mirror
  {
    // The tile's columns lie at |kx| >= x0 dkx on rows |ky| = s dky: when
    // even the nearest is beyond the filter's support, every input is zero
    // and so is every output (written without reading or transforming).
    const float kx = (float)x0 * p.dkx, ky = (float)s * p.dky;
    if (p.rc2_zero > 0.f && kx * kx + ky * ky > p.rc2_zero) {
      for (int idx = threadIdx.x; idx < W * N; idx += NT) {
        const int w = idx % W, zout = idx / W;
        const int xs = x0 + w;
        if (xs > xmax) continue;
        float2* plane = p.stageZ + ((long long)a * N + zout) * p.ny * p.nx;
        const int xd = xs + p.nx / 2 - 1;
        plane[(long long)ycs * p.nx + xd] = make_float2(0.f, 0.f);
        plane[(long long)ycp2 * p.nx + xd] = make_float2(0.f, 0.f);
        if (xs >= 1 && xs <= xmax - 1) {
          const int xm = p.nx / 2 - 1 - xs;
          plane[(long long)ycp2 * p.nx + xm] = make_float2(0.f, 0.f);
          plane[(long long)ycs * p.nx + xm] = make_float2(0.f, 0.f);
        }
      }
      return;  // uniform over the workgroup
    }
  }
  load_twiddles(twr, twstage, threadIdx.x % TPT, tw, threadIdx.x, blockDim.x);
  const long long dirstride = (long long)p.nbands * p.band_stride;
  const float2* b0;
  const float2* b1;
  if (B0) {
    b0 = p.bands;
    b1 = nullptr;
  } else {
    const int no = p.norders - 1;
    const int dir = (a - 1) / no, order = 1 + (a - 1) % no;
    b0 = p.bands + dir * dirstride + (2 * order - 1) * p.band_stride;
    b1 = b0 + p.band_stride;
  }
  // element e = (j * N + z) * W + w: slot j (B0: row s, row -s; else re s,
  // im s, re -s, im -s), plane z, column x0 + w (clamped: the columns past
  // nx/2 are transformed but never written)
  constexpr int E = S * N * W;
  constexpr int B = (E + NT - 1) / NT;
  long long off[B];
  int slot[B];
#pragma unroll
  for (int k = 0; k < B; ++k) {
    const int e = threadIdx.x + k * NT < E ? threadIdx.x + k * NT : E - 1;  // past E: a valid address, not stored
    const int w = e % W, jz = e / W;
    const int j = jz / N, z = jz - j * N;
    const int row = (B0 ? j : j >> 1) == 0 ? s : sp;
    const int x = x0 + w <= xmax ? x0 + w : xmax;
    off[k] = ((long long)z * p.ny + row) * p.nxh + x;
    slot[k] = (j * W + w) * LP + lds_swz<N>(z);
  }
  float2 v[B];
  if (B0) {
#pragma unroll
    for (int k = 0; k < B; ++k) v[k] = make_float2(0.f, 0.f);
    for (int d = 0; d < p.ndirs; ++d) {
      const float2* bd = b0 + d * dirstride;
#pragma unroll
      for (int k = 0; k < B; ++k) {
        const float2 t = bd[off[k]];
        v[k].x += t.x;
        v[k].y += t.y;
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const int e = threadIdx.x + k * NT < E ? threadIdx.x + k * NT : E - 1;
      const int j = (e / W) / N;
      v[k] = ((j & 1) ? b1 : b0)[off[k]];
    }
  }
#pragma unroll
  for (int k = 0; k < B; ++k)
    if (threadIdx.x + k * NT < E) lds[slot[k]] = v[k];
  __syncthreads();
  fft_lds<N, true>(lds + (threadIdx.x / TPT) * LP, threadIdx.x % TPT, twr);
  for (int idx = threadIdx.x; idx < W * N; idx += NT) {
    const int w = idx % W, zout = idx / W;
    const int xs = x0 + w;
    if (xs > xmax) continue;
    const int zp = lds_swz<N>(zout);
    float2 ds, dp, ms, mp;
    if (B0) {
      const float2 bs = lds[w * LP + zp], bp = lds[(W + w) * LP + zp];
      ds = bs;
      dp = bp;
      ms = make_float2(bs.x, -bs.y);
      mp = make_float2(bp.x, -bp.y);
    } else {
      const float2 rs = lds[w * LP + zp], is = lds[(W + w) * LP + zp];
      const float2 rp = lds[(2 * W + w) * LP + zp], ip = lds[(3 * W + w) * LP + zp];
      ds = make_float2(rs.x - is.y, rs.y + is.x);
      dp = make_float2(rp.x - ip.y, rp.y + ip.x);
      ms = make_float2(rs.x + is.y, is.x - rs.y);
      mp = make_float2(rp.x + ip.y, ip.x - rp.y);
    }
    float2* plane = p.stageZ + ((long long)a * N + zout) * p.ny * p.nx;
    const int xd = xs + p.nx / 2 - 1;
    plane[(long long)ycs * p.nx + xd] = ds;
    plane[(long long)ycp2 * p.nx + xd] = dp;
    if (xs >= 1 && xs <= xmax - 1) {
      const int xm = p.nx / 2 - 1 - xs;
      plane[(long long)ycp2 * p.nx + xm] = ms;  // mirror of row s lands in row -s
      plane[(long long)ycs * p.nx + xm] = mp;   // and that of row -s in row s
    }
  }
}

// ---- y pass: W adjacent xc columns of plane zout (blockIdx.y) of band a (blockIdx.z)
// Lengths with a wave-local two-pass plan use that core (natural-order LDS
// columns, register twiddles); the others the radix-stage core.
template <int N>
constexpr bool asm_y_fft2() {
  return use_fft2<N>() && fft2_tpt(N) >= 4;
}
template <int N>
constexpr int asm_y_tpt() {
  return asm_y_fft2<N>() ? fft2_tpt(N) : fft_tpt(N);
}
template <int N>
constexpr int asm_y_pitch() {
  return asm_y_fft2<N>() ? fft2_pitch<N>() : lds_pitch(N);
}
template <int N>
__device__ __forceinline__ int asm_y_pos(int Y) {
  if constexpr (asm_y_fft2<N>())
    return Y;
  else
    return lds_swz<N>(Y);
}
// Two-pass tiles hold at most 80 KB, so that two workgroups share a CU;
// their row segments (8 columns, 64 bytes) pair up in one L2 through the
// XCD-contiguous tile order.
template <int N>
constexpr int asm_y_w() {
  if (asm_y_fft2<N>()) {
    int w = 16;
    while (w > 1 && w * asm_y_pitch<N>() * 8 > 80 * 1024) w /= 2;
    while (w * asm_y_tpt<N>() < 64) w *= 2;
    return w;
  }
  constexpr int tpt = fft_tpt(N);
  int w = 1024 / tpt;
  if (w > 16) w = 16;
  while (w > 1 && w * lds_pitch(N) * 8 > 80 * 1024) w /= 2;
  return w;
}

// Column pitch of the register-in, register-out tiles (exact 2x zoom): a
// wavefront holds 8 columns x 8 threads, so the pitch is 2 mod 16 (the 16
// lanes of a 64-bit LDS access land on distinct bank pairs).
template <int N>
constexpr int asm_y_pitch_regs() {
  return asm_y_fft2<N>() ? fft2_span(N) + (18 - fft2_span(N) % 16) % 16 : 1;
}

// Exact 2x zoom (ny = N/2): slot c N1 + n1 of thread t holds row
// Y = N2 n1 + N1 c + t of the zero-padded column (source row Y + N/4 - 1 for
// Y <= N/4, Y - 3N/4 - 1 for Y >= 3N/4 + 1, zero between), read from scol
// (the column's first element, row stride nx).  The cases are compile-time
// after unrolling except for a slot that straddles an edge of the zero band.
template <int N>
__device__ __forceinline__ void asm_y_fetch_col(const float2* scol, int nx, int ny, float2 (&v)[fft2_plan(N).n2],
                                                int t) {
  constexpr int N1 = fft2_plan(N).n1, N2 = fft2_plan(N).n2, MC = N2 / N1;
  constexpr int H = N / 4, YHI = N - H + 1;
#pragma unroll
  for (int c = 0; c < MC; ++c)
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
      const int Y0 = N2 * n1 + N1 * c;
      float2& d = v[c * N1 + n1];
      if (Y0 + N1 - 1 <= H) {
        d = scol[(long long)(Y0 + t + H - 1) * nx];
      } else if (Y0 >= YHI) {
        d = scol[(long long)(Y0 + t - YHI) * nx];
      } else if (Y0 > H && Y0 + N1 - 1 < YHI) {
        d = make_float2(0.f, 0.f);
      } else {
        const int Y = Y0 + t;
        const bool lo = Y <= H, hi = Y >= YHI;
        int sr = lo ? Y + H - 1 : Y - YHI;
        sr = sr < 0 ? 0 : (sr > ny - 1 ? ny - 1 : sr);
        const float2 w = scol[(long long)sr * nx];
        d = (lo || hi) ? w : make_float2(0.f, 0.f);
      }
    }
}

template <int N>
__global__ __launch_bounds__(asm_y_w<N>() * asm_y_tpt<N>()) void k_asm_y(AsmParams p, const float2* __restrict__ tw) {
  constexpr bool F2 = asm_y_fft2<N>();
  constexpr int TPT = asm_y_tpt<N>();
  constexpr int W = asm_y_w<N>();
  constexpr int NT = W * TPT;
  constexpr int LP = asm_y_pitch<N>();
  static_assert(!F2 || NT % 64 == 0, "two-pass tiles are whole wavefronts");
  constexpr int LPR = asm_y_pitch_regs<N>();
  constexpr int LDSN = W * (LP > LPR ? LP : LPR);
  static_assert(NT <= 1024 && (LDSN + ((N > 1 && !F2) ? N : 1)) * 8 <= 160 * 1024, "the tile fits a workgroup");
  __shared__ float2 lds[LDSN];
  FftTw<F2 ? 1 : N> twr;
  __shared__ float2 twstage[(N > 1 && !F2) ? N : 1];
  Fft2Tw<F2 ? N : 64> tw2;
  const bool zoom2 = F2 && 2 * p.ny == N;
  if constexpr (F2)
    tw2.load(zoom2 ? threadIdx.x / W : threadIdx.x % TPT, tw);
  else
    load_twiddles(twr, twstage, threadIdx.x % TPT, tw, threadIdx.x, blockDim.x);
  unsigned bx, by, bz;
  xcd_tile3(bx, by, bz);
  const int a = (int)bz, zout = (int)by;
  const int xc0 = (int)bx * W;
  {
    // the tile's column nearest to the centre decides: the dead columns lie
    // at both ends of the row (uniform over the workgroup)
    const int c = p.nx / 2 - 1 < xc0 ? xc0 : (p.nx / 2 - 1 > xc0 + W - 1 ? xc0 + W - 1 : p.nx / 2 - 1);
    if (asm_col_dead(p, c)) {
      float2* dst = p.stageY + ((long long)a * p.zdim + zout) * (long long)N * p.nx;
      for (int idx = threadIdx.x; idx < N * W; idx += NT) {
        const int Y = idx / W, w = idx - (idx / W) * W;
        if (xc0 + w < p.nx) dst[(long long)Y * p.nx + xc0 + w] = make_float2(0.f, 0.f);
      }
      return;
    }
  }
  const float2* src = p.stageZ + ((long long)a * p.zdim + zout) * p.ny * p.nx;
  if constexpr (F2) {
    if (zoom2) {
      // thread (t, w) = (threadIdx.x / W, threadIdx.x % W): column xc0 + w,
      // pass-1 inputs from global memory into registers, outputs from
      // registers to global memory; each wavefront's loads and stores cover
      // whole W-column row segments of 8 consecutive rows, as the staged
      // copies do, and the padded column never passes through LDS
      constexpr int N1 = fft2_plan(N).n1, N2 = fft2_plan(N).n2;
      const int w = threadIdx.x % W, t = threadIdx.x / W;
      const int xc = xc0 + w < p.nx ? xc0 + w : p.nx - 1;
      float2 v[N2];
      asm_y_fetch_col<N>(src + xc, p.nx, p.ny, v, t);
      fft2_regs<N, true, false>(lds + w * LPR, t, tw2, v);
      if (xc0 + w < p.nx) {
        float2* dcol = p.stageY + ((long long)a * p.zdim + zout) * (long long)N * p.nx + xc0 + w;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) dcol[(long long)(t + N1 * k2) * p.nx] = v[k2];
      }
      return;
    }
  }
  // Source row sr goes to Y = sr - hy (sr >= hy) or Y = sr + N - hy (sr < hy);
  // the rows between are zero.  Loads run over source rows in batches of B
  // per thread with clamped addresses and no branches, so a batch is in
  // flight at once (one batch for a 2x zoom).
  const int hy = p.ny / 2 - 1;
  constexpr int B = (N / 2 * W + NT - 1) / NT;
  const int cnt = p.ny * W;
  for (int i0 = threadIdx.x; i0 < cnt; i0 += NT * B) {
    float2 v[B];
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const int idx = i0 + k * NT < cnt ? i0 + k * NT : cnt - 1;
      const int sr = idx / W, w = idx - (idx / W) * W;
      const int xc = xc0 + w < p.nx ? xc0 + w : p.nx - 1;
      v[k] = src[(long long)sr * p.nx + xc];
    }
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const int idx = i0 + k * NT;
      const int sr = idx / W, w = idx - (idx / W) * W;
      const int Y = sr >= hy ? sr - hy : sr + N - hy;
      if (idx < cnt) lds[w * LP + asm_y_pos<N>(Y)] = xc0 + w < p.nx ? v[k] : make_float2(0.f, 0.f);
    }
  }
  for (int idx = threadIdx.x; idx < (N - p.ny) * W; idx += NT) {
    const int Y = p.ny - p.ny / 2 + 1 + idx / W, w = idx - (idx / W) * W;
    lds[w * LP + asm_y_pos<N>(Y)] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  if constexpr (F2) {
    fft2_lds<N, true, true>(lds + (threadIdx.x / TPT) * LP, threadIdx.x % TPT, tw2);
    __syncthreads();
  } else {
    fft_lds<N, true>(lds + (threadIdx.x / TPT) * LP, threadIdx.x % TPT, twr);
  }
  float2* dst = p.stageY + ((long long)a * p.zdim + zout) * (long long)N * p.nx;
  for (int idx = threadIdx.x; idx < N * W; idx += NT) {
    const int Y = idx / W, w = idx - (idx / W) * W;
    const int xc = xc0 + w;
    if (xc < p.nx) dst[(long long)Y * p.nx + xc] = lds[w * LP + asm_y_pos<N>(Y)];
  }
}

// ---- x pass: row Y (blockIdx.x) of every output plane.
// A workgroup transforms G bands of the row at a time (TPT threads each) and
// walks all bands, so every output element is accumulated in registers and
// stored once.  The modulation phasors of the row (E per thread for bands
// 1..kAsmLutBands) are computed once per workgroup and kept in registers for
// all planes; the FFT twiddles are read from an LDS table.  Lengths with a
// two-pass plan (N = N1 * N2, fft2.hpp) use that core with G = 8 bands per
// round (one round for 3 directions x 3 orders); the others use the
// radix-stage core with 512 threads.
constexpr int kAsmXBands = 8;
constexpr int kAsmLutBands = 6;  // 1 + ndirs*(norders-1) <= 7 keeps every phasor in registers

template <int N>
constexpr bool asm_x_fft2() {
  return fft2_ok(N) && fft2_tpt(N) >= 8 && 64 % fft2_tpt(N) == 0 && (kAsmXBands * fft2_tpt(N)) % 64 == 0;
}
template <int N>
constexpr int asm_x_tpt() {
  if (asm_x_fft2<N>()) return fft2_tpt(N);
  return fft_tpt(N) >= 128 ? 128 : fft_tpt(N);
}
// The radix-stage core takes 512 threads (four rows of 128), or 256 (two
// rows) from 3072-point rows on, so that the rows and the twiddle table fit
// the 160 KB of LDS.
template <int N>
constexpr int asm_x_threads() {
  return asm_x_fft2<N>() ? kAsmXBands * fft2_tpt(N) : (N >= 3072 ? 256 : 512);
}

// modulation phasor of assembly band a (>= 1) at output column X of row Y
// (cos_sin_kernel, gpuFunctionsImpl.cu:1942-1954: float angle, then sin/cos)
template <int N>
// 2 (cos, sin) of the modulation phase (the factor 2 of write_outbuffer_kernel2,
// exact in float, folded in)
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
  return make_float2(2.f * cs, 2.f * sn);
}

// LDS row of one band: natural order for the two-pass core, swizzled for the
// radix-stage core
template <int N>
constexpr int asm_x_pitch() {
  return asm_x_fft2<N>() ? fft2_span(N) : lds_pitch(N);
}
template <int N>
__device__ __forceinline__ int asm_x_pos(int X) {
  if constexpr (asm_x_fft2<N>())
    return X;
  else
    return lds_swz<N>(X);
}

// Adds one transformed band row (band a) in LDS to the row accumulators
// (write_outbuffer_kernel1/2, gpuFunctionsImpl.cu:1956-1991: band 0 adds its
// real part, side bands add Re(2 v phasor) with a double intermediate).
// SRC >= 0: the phasor is lut[.][SRC]; SRC == -1: band 0; SRC == -2: the
// phasor is computed.  Every table index is a compile-time constant, so the
// table stays in registers.
template <int N, int NT, int E, int LB, int SRC>
__device__ __forceinline__ void asm_x_add_band(const AsmParams& p, const float2* row, int a, int Y,
                                               float (&acc)[E], const float2 (&lut)[E][LB]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int X = threadIdx.x + e * NT;
    if (X < N) {
      const float2 v = row[asm_x_pos<N>(X)];
      if constexpr (SRC == -1) {
        acc[e] += v.x;
      } else {
        float2 ph;
        if constexpr (SRC >= 0)
          ph = lut[e][SRC];
        else
          ph = asm_phasor<N>(p, a, X, Y);
        // 2 v.x cos - 2 v.y sin, ph = 2 (cos, sin): a product and a fused
        // multiply-add in float.  The reference forms this term in double and
        // rounds once when adding it to the float output; here the term itself
        // carries up to an ulp of its larger product, a few 1e-8 of the output,
        // for a third of the instructions.
        acc[e] += __builtin_fmaf(v.x, ph.x, -(v.y * ph.y));
      }
    }
  }
}

// Bands A .. AEND-1 of round R (slot A - R*G), each with its table entry.
template <int N, int NT, int G, int E, int LB, int R, int A, int AEND>
__device__ __forceinline__ void asm_x_table_bands(const AsmParams& p, const float2* lds, int Y, float (&acc)[E],
                                                  const float2 (&lut)[E][LB]) {
  if constexpr (A < AEND) {
    constexpr int LP = asm_x_pitch<N>();
    if (A < p.nasm) {
      asm_x_add_band<N, NT, E, LB, A == 0 ? -1 : A - 1>(p, lds + (A - R * G) * LP, A, Y, acc, lut);
      asm_x_table_bands<N, NT, G, E, LB, R, A + 1, AEND>(p, lds, Y, acc, lut);
    }
  }
}

// Adds the G transformed bands of round R (bands R*G + sl in LDS slot sl),
// in band order.  Band 0 and side bands 1..NTAB come with compile-time table
// entries (NTAB = LB with the phasor table, 0 without it); the rest compute
// their phasors.
template <int N, int NT, int G, int E, int LB, int NTAB, int R>
__device__ __forceinline__ void asm_x_round(const AsmParams& p, const float2* lds, int a0, int Y, float (&acc)[E],
                                            const float2 (&lut)[E][LB]) {
  constexpr int LP = asm_x_pitch<N>();
  constexpr int A0 = R * G;
  constexpr int AT = (R < 0) ? 0 : ((A0 + G < NTAB + 1) ? A0 + G : NTAB + 1);  // first band without a table entry
  int sl = 0;
  if constexpr (R >= 0 && AT > A0) {
    asm_x_table_bands<N, NT, G, E, LB, R, A0, AT>(p, lds, Y, acc, lut);
    sl = AT - A0;
  }
  for (; sl < G && a0 + sl < p.nasm; ++sl) asm_x_add_band<N, NT, E, LB, -2>(p, lds + sl * LP, a0 + sl, Y, acc, lut);
}

// Rows Y of bands a0 .. a0+R-1 of plane zout, each centred in a zero-padded
// length-N row (ifftshift layout of the zoomed grid), into LDS rows dst,
// dst + LP, ... by the 64 lanes of one wavefront.  Source column s goes to
// X = s - (half - 1) for s >= half - 1 and to X = s + N - (half - 1) below;
// the columns between are zero, and so are bands >= nasm.  The loads run
// over contiguous source columns in batches of R x B with clamped (always
// valid) addresses and no branches, so a batch is in flight at once.
template <int N, int R>
__device__ __forceinline__ void asm_x_load_rows(const AsmParams& p, float2* dst, int a0, int zout, int Y, int lane) {
  constexpr int LP = asm_x_pitch<N>();
  constexpr int B = N >= 256 ? N / 128 : 2;  // one batch for a 2x zoom
  const int half = p.nx / 2;
  const float2* src[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int a = a0 + r < p.nasm ? a0 + r : p.nasm - 1;
    src[r] = p.stageY + (((long long)a * p.zdim + zout) * p.ydim + Y) * p.nx;
  }
  for (int s0 = lane; s0 < p.nx; s0 += 64 * B) {
    float2 v[R][B];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int k = 0; k < B; ++k) {
        const int sc = s0 + 64 * k;
        v[r][k] = src[r][sc < p.nx ? sc : p.nx - 1];
      }
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int k = 0; k < B; ++k) {
        const int sc = s0 + 64 * k;
        const int X = sc >= half - 1 ? sc - (half - 1) : sc + N - (half - 1);
        if (sc < p.nx) dst[r * LP + asm_x_pos<N>(X)] = a0 + r < p.nasm ? v[r][k] : make_float2(0.f, 0.f);
      }
  }
#pragma unroll
  for (int r = 0; r < R; ++r)
    for (int X = p.nx - half + 1 + lane; X < N - (half - 1); X += 64) dst[r * LP + asm_x_pos<N>(X)] = make_float2(0.f, 0.f);
}

// The same rows when the source row fits one batch (nx <= 64 B, every 2x
// zoom): asm_x_fetch issues the global loads into registers, asm_x_put
// writes them (and the zero columns) to LDS.  The kernel fetches the next
// (plane, round) right after putting the current one, so the loads are in
// flight during the transform and accumulation instead of stalling the
// start of every plane.
template <int N>
constexpr int asm_x_batch() {
  return N >= 256 ? N / 128 : 2;
}
template <int N, int R>
struct AsmXRows {
  float2 v[R][asm_x_batch<N>()];
};
template <int N, int R>
__device__ __forceinline__ void asm_x_fetch(const AsmParams& p, AsmXRows<N, R>& b, int a0, int zout, int Y,
                                            int lane) {
  constexpr int B = asm_x_batch<N>();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int a = a0 + r < p.nasm ? a0 + r : p.nasm - 1;
    const float2* src = p.stageY + (((long long)a * p.zdim + zout) * p.ydim + Y) * p.nx;
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const int sc = lane + 64 * k;
      b.v[r][k] = src[sc < p.nx ? sc : p.nx - 1];
    }
  }
}
template <int N, int R>
__device__ __forceinline__ void asm_x_put(const AsmParams& p, float2* dst, const AsmXRows<N, R>& b, int a0,
                                          int lane) {
  constexpr int LP = asm_x_pitch<N>();
  constexpr int B = asm_x_batch<N>();
  const int half = p.nx / 2;
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int k = 0; k < B; ++k) {
      const int sc = lane + 64 * k;
      const int X = sc >= half - 1 ? sc - (half - 1) : sc + N - (half - 1);
      if (sc < p.nx) dst[r * LP + asm_x_pos<N>(X)] = a0 + r < p.nasm ? b.v[r][k] : make_float2(0.f, 0.f);
    }
#pragma unroll
  for (int r = 0; r < R; ++r)
    for (int X = p.nx - half + 1 + lane; X < N - (half - 1); X += 64) dst[r * LP + asm_x_pos<N>(X)] = make_float2(0.f, 0.f);
}

// Exact 2x zoom (nx = N/2, half = N/4): thread t's pass-1 inputs of row Y
// of band a, plane zout, loaded straight into registers, slot c N1 + n1
// holding X = N2 n1 + N1 c + t of the zero-padded row (asm_x_put's layout:
// X <= N/4 from source column X + N/4 - 1, X >= 3N/4 + 1 from X - 3N/4 - 1,
// zero between).  The slots' cases are compile-time after unrolling; only a
// slot that straddles an edge of the zero band tests per thread.
template <int N>
__device__ __forceinline__ void asm_x_fetch_cols(const AsmParams& p, float2 (&v)[fft2_plan(N).n2], int a, int zout,
                                                 int Y, int t) {
  constexpr int N1 = fft2_plan(N).n1, N2 = fft2_plan(N).n2, MC = N2 / N1;
  constexpr int H = N / 4, XHI = N - H + 1;
  const float2* src = p.stageY + (((long long)a * p.zdim + zout) * p.ydim + Y) * p.nx;
#pragma unroll
  for (int c = 0; c < MC; ++c)
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
      const int X0 = N2 * n1 + N1 * c;
      float2& d = v[c * N1 + n1];
      if (X0 + N1 - 1 <= H) {
        d = src[X0 + t + H - 1];
      } else if (X0 >= XHI) {
        d = src[X0 + t - XHI];
      } else if (X0 > H && X0 + N1 - 1 < XHI) {
        d = make_float2(0.f, 0.f);
      } else {
        const int X = X0 + t;
        const bool lo = X <= H, hi = X >= XHI;
        int sc = lo ? X + H - 1 : X - XHI;
        sc = sc < 0 ? 0 : (sc > p.nx - 1 ? p.nx - 1 : sc);
        const float2 w = src[sc];
        d = (lo || hi) ? w : make_float2(0.f, 0.f);
      }
    }
}

template <int N>
__global__ __launch_bounds__(asm_x_threads<N>(), 2) void k_asm_x(AsmParams p,
                                                                                   const float2* __restrict__ tw) {
#pragma clang fp contract(off)
  constexpr int NT = asm_x_threads<N>();
  constexpr int TPT = asm_x_tpt<N>();
  constexpr int G = NT / TPT;
  constexpr int LP = asm_x_pitch<N>();
  constexpr int E = (N + NT - 1) / NT;
  constexpr bool LUT = E <= 4;
  constexpr int LB = LUT ? kAsmLutBands : 1;
  static_assert(!asm_x_fft2<N>() || (64 % TPT == 0 && NT % 64 == 0), "two-pass transforms must not straddle waves");
  static_assert((G * LP + N) * 8 <= 160 * 1024, "band rows and twiddles fit the LDS");
  __shared__ float2 lds[G * LP];
  __shared__ float2 twt[N];
  for (int i = threadIdx.x; i < N; i += NT) twt[i] = tw[i];
  const int Y = blockIdx.x;
  const int g = threadIdx.x / TPT, t = threadIdx.x % TPT;
  const int half = p.nx / 2;
  float2 lut[E][LB];
#pragma unroll
  for (int e = 0; e < E; ++e)
#pragma unroll
    for (int b = 0; b < LB; ++b) {
      const int X = threadIdx.x + e * NT;
      lut[e][b] = (LUT && X < N && b + 1 < p.nasm) ? asm_phasor<N>(p, b + 1, X, Y) : make_float2(1.f, 0.f);
    }
  __syncthreads();
  if constexpr (asm_x_fft2<N>()) {
    if (2 * p.nx == N) {
      // exact 2x zoom: each thread's pass-1 inputs come from global memory
      // straight into registers (no LDS staging of the padded rows), and the
      // next (plane, round) is fetched while this one is transformed
      constexpr int N2 = fft2_plan(N).n2;
      float2 nxt[N2];
      asm_x_fetch_cols<N>(p, nxt, g < p.nasm ? g : p.nasm - 1, 0, Y, t);
      for (int zout = 0; zout < p.zdim; ++zout) {
        float acc[E];
#pragma unroll
        for (int e = 0; e < E; ++e) acc[e] = 0.f;
#pragma nounroll
        for (int a0 = 0; a0 < p.nasm; a0 += G) {
          float2 v[N2];
#pragma unroll
          for (int i = 0; i < N2; ++i) v[i] = nxt[i];
          int na0 = a0 + G, nz = zout;
          if (na0 >= p.nasm) {
            na0 = 0;
            ++nz;
          }
          if (nz < p.zdim) asm_x_fetch_cols<N>(p, nxt, na0 + g < p.nasm ? na0 + g : p.nasm - 1, nz, Y, t);
          fft2_lds_from_regs<N, true, true>(lds + g * LP, t, twt, v);
          __syncthreads();
          if (a0 == 0)
            asm_x_round<N, NT, G, E, LB, LUT ? LB : 0, 0>(p, lds, a0, Y, acc, lut);
          else if (a0 == G)
            asm_x_round<N, NT, G, E, LB, LUT ? LB : 0, 1>(p, lds, a0, Y, acc, lut);
          else
            asm_x_round<N, NT, G, E, LB, LUT ? LB : 0, -1>(p, lds, a0, Y, acc, lut);
          __syncthreads();
        }
        float* orow = p.out + ((long long)zout * p.ydim + Y) * N;
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int X = threadIdx.x + e * NT;
          if (X < N) orow[X] = acc[e];
        }
      }
      return;
    }
    if (p.nx <= 64 * asm_x_batch<N>()) {
      constexpr int TPW = 64 / TPT;
      const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
      AsmXRows<N, TPW> rows;
      asm_x_fetch<N, TPW>(p, rows, wave * TPW, 0, Y, lane);
      for (int zout = 0; zout < p.zdim; ++zout) {
        float acc[E];
#pragma unroll
        for (int e = 0; e < E; ++e) acc[e] = 0.f;
#pragma nounroll
        for (int a0 = 0; a0 < p.nasm; a0 += G) {
          asm_x_put<N, TPW>(p, lds + wave * TPW * LP, rows, a0 + wave * TPW, lane);
          // the next (plane, round) this wavefront loads
          int na0 = a0 + G, nz = zout;
          if (na0 >= p.nasm) {
            na0 = 0;
            ++nz;
          }
          if (nz < p.zdim) asm_x_fetch<N, TPW>(p, rows, na0 + wave * TPW, nz, Y, lane);
          wave_lds_sync();
          fft2_lds<N, true, true>(lds + g * LP, t, twt);
          __syncthreads();
          if (a0 == 0)
            asm_x_round<N, NT, G, E, LB, LUT ? LB : 0, 0>(p, lds, a0, Y, acc, lut);
          else if (a0 == G)
            asm_x_round<N, NT, G, E, LB, LUT ? LB : 0, 1>(p, lds, a0, Y, acc, lut);
          else
            asm_x_round<N, NT, G, E, LB, LUT ? LB : 0, -1>(p, lds, a0, Y, acc, lut);
          __syncthreads();
        }
        float* orow = p.out + ((long long)zout * p.ydim + Y) * N;
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int X = threadIdx.x + e * NT;
          if (X < N) orow[X] = acc[e];
        }
      }
      return;
    }
  }
  for (int zout = 0; zout < p.zdim; ++zout) {
    float acc[E];
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = 0.f;
#pragma nounroll
    for (int a0 = 0; a0 < p.nasm; a0 += G) {
      if constexpr (asm_x_fft2<N>()) {
        // each wavefront loads and transforms its own TPW bands, so only the
        // accumulation below needs the whole workgroup
        constexpr int TPW = 64 / TPT;
        const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
        asm_x_load_rows<N, TPW>(p, lds + wave * TPW * LP, a0 + wave * TPW, zout, Y, lane);
        wave_lds_sync();
        fft2_lds<N, true, true>(lds + g * LP, t, twt);
        __syncthreads();
      } else {
        // G bands of row Y of plane zout, zero-padded to N, into LDS
        for (int idx = threadIdx.x; idx < G * N; idx += NT) {
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
          lds[sl * LP + asm_x_pos<N>(X)] = v;
        }
        __syncthreads();
        fft_lds_t<N, true, TPT>(lds + g * LP, t, twt);
      }
      if (a0 == 0)
        asm_x_round<N, NT, G, E, LB, LUT ? LB : 0, 0>(p, lds, a0, Y, acc, lut);
      else if (a0 == G)
        asm_x_round<N, NT, G, E, LB, LUT ? LB : 0, 1>(p, lds, a0, Y, acc, lut);
      else
        asm_x_round<N, NT, G, E, LB, LUT ? LB : 0, -1>(p, lds, a0, Y, acc, lut);
      __syncthreads();
    }
    float* orow = p.out + ((long long)zout * p.ydim + Y) * N;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int X = threadIdx.x + e * NT;
      if (X < N) orow[X] = acc[e];
    }
  }
}

template <int N>
static bool launch_asm_z_pair(const AsmParams& p, hipStream_t s) {
  if constexpr (asm_zp_ok<N>()) {
    if (p.zdim != p.nz || p.nz == 1 || (p.nx & 1) || (p.ny & 1)) return false;
    const int nstored = p.nx / 2 + 1;
    hipLaunchKernelGGL((k_asm_z_pair<N, true>), dim3(ceil_div(nstored, asm_zp_w<N, true>()), p.ny / 2 + 1, 1),
                       dim3(asm_zp_threads<true>()), 0, s, p, twiddles(N));
    SIMR_LAUNCH_CHECK();
    if (p.nasm > 1) {
      hipLaunchKernelGGL((k_asm_z_pair<N, false>),
                         dim3(ceil_div(nstored, asm_zp_w<N, false>()), p.ny / 2 + 1, p.nasm - 1), dim3(asm_zp_threads<false>()), 0, s, p,
                         twiddles(N));
      SIMR_LAUNCH_CHECK();
    }
    return true;
  }
  return false;
}

// ---- z pass for depths without a transform kernel (any zdim): the
// move_kernel gather of k_asm_z, one element per thread, into dst
// [a][zout][y][x] (the y-pass staging buffer, idle until the y pass), then a
// direct inverse DFT along z (dft_cols) from dst into stageZ.
__global__ __launch_bounds__(256) void k_asm_z_gather(AsmParams p, int N, float2* __restrict__ dst) {
#pragma clang fp contract(off)
  const long long nxy = (long long)p.nx * p.ny;
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long long)N * nxy) return;
  const int a = blockIdx.y;
  const int zout = (int)(e / nxy);
  const long long r = e - (long long)zout * nxy;
  const int yc = (int)(r / p.nx), xc = (int)(r - (long long)yc * p.nx);
  bool valid;
  int z = z_of_zout(zout, p.nz, N, valid);
  int x = xc - (p.nx / 2 - 1), yy = yc - (p.ny / 2 - 1);
  const bool cj = x < 0;
  if (cj) {
    x = -x;
    yy = -yy;
    z = -z;
  }
  if (yy < 0) yy += p.ny;
  if (z < 0) z += p.nz;
  float2 v = make_float2(0.f, 0.f);
  if (valid) {
    const long long ind = ((long long)z * p.ny + yy) * p.nxh + x;
    const long long dirstride = (long long)p.nbands * p.band_stride;
    if (a == 0) {  // band 0: sum over directions
      for (int d = 0; d < p.ndirs; ++d) {
        const float2 b = p.bands[d * dirstride + ind];
        v.x += b.x;
        v.y += b.y;
      }
      if (cj) v.y = -v.y;
    } else {  // side band: re + i im of the order's two real-space bands
      const int no = p.norders - 1;
      const int dir = (a - 1) / no, order = 1 + (a - 1) % no;
      float2 re = p.bands[dir * dirstride + (2 * order - 1) * p.band_stride + ind];
      float2 im = p.bands[dir * dirstride + (2 * order) * p.band_stride + ind];
      if (cj) {
        re.y = -re.y;
        im.y = -im.y;
      }
      v = make_float2(re.x - im.y, re.y + im.x);
    }
  }
  dst[((long long)a * N + zout) * nxy + r] = v;
}

static void launch_asm_z_any(const AsmParams& p, hipStream_t s) {
  const long long nxy = (long long)p.nx * p.ny;
  const long long n = (long long)p.zdim * nxy;
  if (ceil_div(n, 256) > 2147483647LL || p.nasm > 65535) throw ArgError("assembly z pass: stack too large");
  hipLaunchKernelGGL(k_asm_z_gather, dim3((unsigned)ceil_div(n, 256), p.nasm), dim3(256), 0, s, p, p.zdim, p.stageY);
  SIMR_LAUNCH_CHECK();
  dft_cols(p.stageY, p.stageZ, p.zdim, nxy, nxy, n, p.nasm, true, s);
}

void launch_asm_z(const AsmParams& p, hipStream_t s) {
  switch (p.zdim) {
#define SIMR_CASE(N)                                                                              \
  case N: {                                                                                       \
    if (launch_asm_z_pair<N>(p, s)) return;                                                       \
    constexpr int W = (256 / fft_tpt(N)) > 0 ? 256 / fft_tpt(N) : 1;                              \
    hipLaunchKernelGGL(k_asm_z<N>, dim3(ceil_div(p.nx, W), p.ny, p.nasm), dim3(256), 0, s, p,     \
                       twiddles(N));                                                              \
    SIMR_LAUNCH_CHECK();                                                                          \
    return;                                                                                       \
  }
    SIMR_Z_SIZES(SIMR_CASE)
#undef SIMR_CASE
    default:
      launch_asm_z_any(p, s);
  }
}

void launch_asm_y(const AsmParams& p, hipStream_t s) {
  switch (p.ydim) {
#define SIMR_CASE(N)                                                                                \
  case N: {                                                                                         \
    constexpr int W = asm_y_w<N>();                                                                 \
    hipLaunchKernelGGL(k_asm_y<N>, dim3(ceil_div(p.nx, W), p.zdim, p.nasm), dim3(W * asm_y_tpt<N>()), \
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
    hipLaunchKernelGGL(k_asm_x<N>, dim3(p.ydim), dim3(asm_x_threads<N>()), 0, s, p,                    \
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
