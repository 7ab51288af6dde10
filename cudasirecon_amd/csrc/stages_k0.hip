// k0 search: overlap construction and inverse FFT, the z-summed overlap
// products, the cross-correlation peak, and the batched modamp reductions.
//
// Reference (src/cudaSirecon/gpuFunctionsImpl.cu): makeoverlaps 497-675 with
// makeOverlaps0Kernel 677-778 / makeOverlaps1Kernel 780-872, aTimesConjB
// 374-414, computeIntensities 416-441, findrealspacemodamp 1074-1148,
// reductionKernel 1150-1220.
//
// Differences in structure (same values up to FP reassociation):
//  * overlaps are generated inside the x-pass of the inverse FFT, only for
//    the 2*zd+1 planes the support can reach; the zero planes are never
//    written;
//  * the fit reads only the per-(x,y) z sums P = sum_z conj(X) Y, SX, SY,
//    which Parseval's identity along z forms from the stored kz planes, so
//    there is no z transform, and the 3D overlap volumes are never re-read:
//    every later line-search step is a 2D reduction of P against a phase ramp;
//  * a whole batch of line-search candidates is reduced in one launch;
//  * every launch covers the overlap pairs (or candidates) of all directions
//    being fitted, one slot per direction, so a launch fills the GPU and the
//    directions share each host read-back.
#include <type_traits>

#include "device_math.hpp"
#include "fft.hpp"
#include "fft2.hpp"
#include "fft_api.hpp"
#include "sizes.hpp"
#include "stages.hpp"

namespace simr {

// Sequential sums in block order, as the host's loops make them, of K
// interleaved components at once: the workgroup stages a chunk of the
// entries in LDS with every thread (coalesced, all loads in flight), then
// lane k adds up component k in order, eight staged values loaded ahead of
// each eight additions.  Returns lane k's sum (k < K).
constexpr int kSeqChunk = 1024;  // entries per LDS chunk

template <typename T, int K, typename Get>
__device__ __forceinline__ T seq_sum_lanes(int n, double* stage, Get get) {
#pragma clang fp contract(off)
  T acc = 0;
  const int lane = threadIdx.x;
  for (int c0 = 0; c0 < n; c0 += kSeqChunk) {
    const int m = n - c0 < kSeqChunk ? n - c0 : kSeqChunk;
    for (int i = threadIdx.x; i < m * K; i += blockDim.x) stage[i] = get(c0 + i / K, i % K);
    __syncthreads();
    if (lane < K) {
      int i = 0;
      for (; i + 8 <= m; i += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = stage[(i + u) * K + lane];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += (T)v[u];
      }
      for (; i < m; ++i) acc += (T)stage[i * K + lane];
    }
    __syncthreads();
  }
  return acc;
}

// Where one element of overlap0 (which == 0) or overlap1 (which == 1) at
// column j, row i, plane z0 comes from, before the inverse transform: its
// band index, the conjugation of the Hermitian mirror, and the OTF factor
// (makeOverlaps0Kernel / makeOverlaps1Kernel).  ok = false: the element is
// zero.  prof1/prof2: radial OTF profiles of order1/order2 at kz = z0.
//
// The reference's tests sqrtf(d) <= rdistcutoff and sqrtf(|otf|^2) * f >
// otfcutoff are made on the squares (sums of squares in fused form), the OTF
// radius index is the hardware square root of d / dkr^2 and the
// normalisation 1/sqrt(|a|^2 + |b|^2) the hardware reciprocal square root:
// within a few ulps of the reference's
// correctly rounded sqrtf and divisions (an element exactly on a cut-off
// circle may land on the other side), far inside the overlap tolerance, at
// a third of the instructions.
struct OvSrc {
  int ind;
  bool ok, conj;
  float2 fact;
};

// Bilinear profile lookup (dev_otfinterpolate's radial interpolation) at a
// radius index that is a square root, so never negative: truncation is the
// floor and the clamp at 0 cannot act; t0 + ar (t1 - t0) in fused form is
// within an ulp of (1 - ar) t0 + ar t1.
__device__ __forceinline__ float2 prof_lerp(const float2* prof, int nxotf, float krindex) {
  int irindex = (int)krindex;
  float ar = __builtin_amdgcn_fractf(krindex);
  if (irindex > nxotf - 2) {
    irindex = nxotf - 2;
    ar = 1.f;
  }
  const float2 t0 = prof[irindex], t1 = prof[irindex + 1];
  return make_float2(__builtin_fmaf(ar, t1.x - t0.x, t0.x), __builtin_fmaf(ar, t1.y - t0.y, t0.y));
}

template <bool RA>
__device__ __forceinline__ OvSrc ov_source(const OvParams& p, float kx, float ky, const float2* prof1,
                                           const float2* prof2, int which, int j, int i, int z0) {
#pragma clang fp contract(off)
  OvSrc o{0, false, false, make_float2(0.f, 0.f)};
  float x1f = (float)j, y1f = (float)i;
  if (x1f > (float)(p.nx / 2)) x1f -= (float)p.nx;
  if (y1f > (float)(p.ny / 2)) y1f -= (float)p.ny;
  x1f *= p.dkx;
  y1f *= p.dky;
  const float d1 = __builtin_fmaf(x1f, x1f, y1f * y1f);
  if (which == 0 ? !(d1 <= p.rc2) : !(d1 < p.rc2)) return o;
  const float x12 = x1f - kx, y12 = y1f - ky;
  const float d12 = __builtin_fmaf(x12, x12, y12 * y12);
  const float x21 = x1f + kx, y21 = y1f + ky;
  const float d21 = __builtin_fmaf(x21, x21, y21 * y21);
  if (!(d12 <= p.rc2 || d21 <= p.rc2)) return o;
  if (z0 == 0 && p.bNoKz0) return o;
  int iin, jin;
  if (j <= p.nx / 2) {
    iin = i;
    jin = j;
  } else {
    jin = p.nx - j;
    iin = i == 0 ? 0 : p.ny - i;
    o.conj = true;
  }
  int zin = o.conj ? -z0 : z0;
  if (zin < 0) zin += p.nz;
  float2 oa, ob;  // own-order OTF at k, other-order OTF at the shifted k
  float sa2, sb2; // their cutoff scales, squared
  if (which == 0) {
    if (!(d12 <= p.rc2)) return o;
    if constexpr (RA) {
      oa = prof_lerp(prof1, p.otf1.nxotf, __builtin_amdgcn_sqrtf(d1 * p.inv_dkr2_1));
      ob = prof_lerp(prof2, p.otf2.nxotf, __builtin_amdgcn_sqrtf(d12 * p.inv_dkr2_2));
    } else {
      oa = otf3d_interp(p.otf1, x1f, y1f, z0, p.kzscale);
      ob = otf3d_interp(p.otf2, x12, y12, z0, p.kzscale);
    }
    sa2 = 1.f;
    sb2 = p.f2;
  } else {
    if (!(d21 <= p.rc2)) return o;
    if constexpr (RA) {
      oa = prof_lerp(prof2, p.otf2.nxotf, __builtin_amdgcn_sqrtf(d1 * p.inv_dkr2_2));
      ob = prof_lerp(prof1, p.otf1.nxotf, __builtin_amdgcn_sqrtf(d21 * p.inv_dkr2_1));
    } else {
      oa = otf3d_interp(p.otf2, x1f, y1f, z0, p.kzscale);
      ob = otf3d_interp(p.otf1, x21, y21, z0, p.kzscale);
    }
    sa2 = p.f2;
    sb2 = 1.f;
  }
  const float ma = __builtin_fmaf(oa.x, oa.x, oa.y * oa.y), mb = __builtin_fmaf(ob.x, ob.x, ob.y * ob.y);
  if (!(ma * sa2 > p.cut2)) return o;
  if (!(mb * sb2 > p.cut2)) return o;
  const float rroot = __builtin_amdgcn_rsqf(ma + mb);
  o.fact = make_float2(ob.x * rroot, ob.y * rroot);
  o.ind = (zin * p.ny + iin) * p.nxh + jin;
  o.ok = true;
  return o;
}

// The overlap element from its source and the two band values read there.
__device__ __forceinline__ float2 ov_combine(const OvSrc& o, float2 vre, float2 vim, bool has_im) {
#pragma clang fp contract(off)
  if (!o.ok) return make_float2(0.f, 0.f);
  if (!has_im) vim = make_float2(0.f, 0.f);
  if (o.conj) {
    vre.y = -vre.y;
    vim.y = -vim.y;
  }
  vre = cmul_nc(vre, o.fact);
  vim = cmul_nc(vim, o.fact);
  return make_float2(vre.x - vim.y, vre.y + vim.x);
}

// Whether row i of overlap `which` can hold a nonzero element: an element at
// lateral frequency (x, y) is nonzero only if |(x, y)| and |(x, y) - s| are
// both within rdistcutoff (s = (kx, ky) for overlap0, -(kx, ky) for overlap1;
// no wrap-around), so the row meets both discs iff |sx| <= a + b with
// a^2 = rc^2 - y^2 and b^2 = rc^2 - (y - sy)^2.  The radii carry a relative
// margin of 1e-4, far beyond the rounding of the per-element tests, so a row
// this rejects is zero.
__device__ __forceinline__ bool ov_row_live(const OvParams& p, float kx, float ky, int which, int i, int z0) {
  if (i >= p.ny || (z0 == 0 && p.bNoKz0)) return false;
  float y = (float)i;
  if (y > (float)(p.ny / 2)) y -= (float)p.ny;
  y *= p.dky;
  const float sx = which == 0 ? kx : -kx, sy = which == 0 ? ky : -ky;
  const float r2 = p.rc2 * 1.0002f;
  const float a2 = r2 - y * y, b2 = r2 - (y - sy) * (y - sy);
  if (a2 < 0.f || b2 < 0.f) return false;
  return fabsf(sx) <= (sqrtf(a2) + sqrtf(b2)) * 1.0001f;
}

// x pass: rows [R*(blockIdx.x/2), +R) of overlap blockIdx.x%2 on plane zb
// (blockIdx.y) of slot blockIdx.z, R = ov_rows(N) (4, fewer when 2R
// transforms would exceed a workgroup or 96 KB of LDS); the plane's two
// radial OTF profiles are built once in LDS.
// A workgroup whose rows all lie outside the overlap's support (about a third
// of them at 3D-SIM geometries) writes zeros and returns.
constexpr int kOvStaticLds = 96 * 1024;
constexpr int ov_rows(int N) {
  int r = 4;
  while (r > 1 && (2 * r * fft_tpt(N) > 1024 || (2 * r * lds_pitch(N) + N) * 8 > kOvStaticLds)) r /= 2;
  return r;
}

template <int N, bool RA>
__global__ __launch_bounds__(ov_rows(N) * fft_tpt(N)) void k_overlaps_x(OvBatch b, const float2* __restrict__ tw) {
  constexpr int kOvRows = ov_rows(N);
  static_assert((kOvRows * lds_pitch(N) + N) * 8 <= kOvStaticLds, "rows and twiddles leave room for the profiles");
  const OvParams& p = b.p[blockIdx.z];
  const int zb = blockIdx.y;
  if (zb > 2 * p.zd) return;  // a slot with a narrower z support (whole block)
  float2* __restrict__ ovx = p.ovx;
  const int nzs = b.nzs_stride;
  constexpr int TPT = fft_tpt(N);
  constexpr int RW = kOvRows;  // rows per workgroup
  constexpr int NT = RW * TPT;
  constexpr int P = lds_pitch(N);
  constexpr int EPT = (RW * N + NT - 1) / NT;
  const int which = blockIdx.x & 1;
  const int i0 = (blockIdx.x >> 1) * RW;
  const int z0 = zb - p.zd;
  float kx = p.kx, ky = p.ky;
  if (p.k0d != nullptr) {
#pragma clang fp contract(off)
    const float2 k = *p.k0d;
    kx = k.x * p.kmul;
    ky = k.y * p.kmul;
  }
  float2* __restrict__ orow = ovx + (((long long)which * nzs + zb) * p.ny + i0) * p.nx;
  {
    bool live = false;
#pragma unroll
    for (int r = 0; r < RW; ++r) live = live || ov_row_live(p, kx, ky, which, i0 + r, z0);
    if (!live) {
#pragma unroll
      for (int q = 0; q < EPT; ++q) {
        const int idx = threadIdx.x + q * NT;
        if (idx < RW * N && i0 + idx / N < p.ny) orow[idx] = make_float2(0.f, 0.f);
      }
      return;  // uniform over the workgroup
    }
  }
  __shared__ float2 lds[RW * P];
  extern __shared__ float2 prof[];  // [2][nxotf]
  FftTw<N> twr;
  __shared__ float2 twstage[N > 1 ? N : 1];
  load_twiddles(twr, twstage, threadIdx.x % TPT, tw, threadIdx.x, blockDim.x);
  const int nxotf = p.otf1.nxotf;
  if constexpr (RA)
    for (int k = threadIdx.x; k < 2 * nxotf; k += NT)
      prof[k] = otf_profile_entry(k < nxotf ? p.otf1 : p.otf2, k < nxotf ? k : k - nxotf, z0, p.kzscale);
  __syncthreads();
  // In chunks of CH elements per thread: the sources of a chunk first, then
  // all its band loads (clamped to a valid index for zero elements) at
  // once, then the products.  Chunks keep the registers (and so the
  // workgroups per CU) within bounds.
  constexpr int CH = EPT < 4 ? EPT : 4;
#pragma unroll
  for (int q0 = 0; q0 < EPT; q0 += CH) {
    OvSrc src[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int idx = threadIdx.x + (q0 + c) * NT;
      const int t = idx / N, j = idx - (idx / N) * N;  // t = row in the workgroup
      const int i = i0 + t;
      src[c] = (q0 + c < EPT && idx < RW * N && i < p.ny)
                   ? ov_source<RA>(p, kx, ky, prof, prof + nxotf, which, j, i, z0)
                   : OvSrc{0, false, false, make_float2(0.f, 0.f)};
    }
    const float2* bre = which == 0 ? p.b1re : p.b2re;
    const float2* bim = which == 0 ? (p.order1 > 0 ? p.b1im : p.b1re) : p.b2im;
    float2 vre[CH], vim[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      vre[c] = bre[src[c].ind];
      vim[c] = bim[src[c].ind];
    }
    const bool has_im = which == 1 || p.order1 > 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int idx = threadIdx.x + (q0 + c) * NT;
      if (q0 + c < EPT && idx < RW * N) {
        const int t = idx / N, j = idx - (idx / N) * N;
        lds[t * P + lds_swz<N>(j)] = ov_combine(src[c], vre[c], vim[c], has_im);
      }
    }
  }
  __syncthreads();
  fft_lds<N, true>(lds + (threadIdx.x / TPT) * P, threadIdx.x % TPT, twr);
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int idx = threadIdx.x + q * NT;
    if (idx < RW * N) {
      const int t = idx / N, j = idx - (idx / N) * N;
      if (i0 + t < p.ny) orow[idx] = lds[t * P + lds_swz<N>(j)];
    }
  }
}

template <int N>
static void launch_ov_x(const OvBatch& b, int nzs_max, hipStream_t s) {
  const OvParams& p = b.p[0];
  const dim3 grid(2 * ceil_div(p.ny, ov_rows(N)), nzs_max, b.n), block(ov_rows(N) * fft_tpt(N));
  if (p.otf1.ra) {
    const size_t lds = (size_t)2 * p.otf1.nxotf * sizeof(float2);
    if (lds > 160 * 1024 - kOvStaticLds) throw ArgError("OTF table too long for the overlap pass's LDS profiles");
    hipLaunchKernelGGL((k_overlaps_x<N, true>), grid, block, lds, s, b, twiddles(N));
  } else {
    hipLaunchKernelGGL((k_overlaps_x<N, false>), grid, block, 0, s, b, twiddles(N));
  }
  SIMR_LAUNCH_CHECK();
}

void launch_overlaps_x(const OvBatch& b, hipStream_t s) {
  if (b.n < 1 || b.n > kMaxDirs) throw ArgError("overlap batch: 1..8 slots");
  int nzs_max = 0;
  for (int k = 0; k < b.n; ++k) {
    const OvParams& p = b.p[k];
    if (p.nx != b.p[0].nx || p.ny != b.p[0].ny || p.otf1.nxotf != b.p[0].otf1.nxotf ||
        p.otf2.nxotf != b.p[0].otf1.nxotf || p.otf1.ra != b.p[0].otf1.ra || p.otf2.ra != b.p[0].otf1.ra)
      throw ArgError("overlap batch: slots of different geometry");
    if (p.zd < 0 || 2 * p.zd + 1 > b.nzs_stride) throw ArgError("overlap z support wider than its planes");
    nzs_max = std::max(nzs_max, 2 * p.zd + 1);
  }
  switch (b.p[0].nx) {
#define SIMR_CASE(N)                   \
  case N:                              \
    launch_ov_x<N>(b, nzs_max, s);     \
    break;
    SIMR_XY_SIZES(SIMR_CASE)
#undef SIMR_CASE
    default:
      throw ArgError("image width " + std::to_string(b.p[0].nx) + " is not a supported size");
  }
}

__global__ __launch_bounds__(256) void k_zsum_totals(ZsumBatch b, int zblocks);

// ---------------------------------------------------------------------------
// The overlaps' inverse transforms with the z sums fused, for the FP64 fit:
// the formation runs with the y transform, by columns (k_ovf_cols), and the
// x transform with the z sums, by rows (k_ovf_rows): a workgroup takes row y
// of every stored plane of both overlaps, so P(x, y) and the SX / SY totals
// accumulate in registers plane after plane.  The transformed planes are
// written once and read once (the three-pass form writes them twice and
// reads them twice).  The order of the two 1D transforms is the other one
// than the three-pass form's, so P agrees with it to rounding, not bit for
// bit.

// Whether column j of overlap `which` can hold a nonzero element (the row
// test of ov_row_live with x and y exchanged).
__device__ __forceinline__ bool ov_col_live(const OvParams& p, float kx, float ky, int which, int j, int z0) {
  if (j >= p.nx || (z0 == 0 && p.bNoKz0)) return false;
  float x = (float)j;
  if (x > (float)(p.nx / 2)) x -= (float)p.nx;
  x *= p.dkx;
  const float sx = which == 0 ? kx : -kx, sy = which == 0 ? ky : -ky;
  const float r2 = p.rc2 * 1.0002f;
  const float a2 = r2 - x * x, b2 = r2 - (x - sx) * (x - sx);
  if (a2 < 0.f || b2 < 0.f) return false;
  return fabsf(sy) <= (sqrtf(a2) + sqrtf(b2)) * 1.0001f;
}

// columns per workgroup of the column pass (radix-stage core: its rows are
// a quarter of the two-pass core's LDS per thread, which keeps the occupancy
// the formation's gathers need)
template <int NY>
constexpr int ovf_cols_w() {
  int w = 8;
  while (w > 1 && w * fft_tpt(NY) > 512) w /= 2;
  return w;
}
template <int NY>
constexpr bool ovf_cols_ok() {
  return fft_supported(NY) && NY > 1 && (ovf_cols_w<NY>() * lds_pitch(NY) + NY) * 8 <= 64 * 1024 &&
         (NY * ovf_cols_w<NY>()) % (ovf_cols_w<NY>() * fft_tpt(NY)) == 0;
}

// Columns [W*(blockIdx.x/2), +W) of overlap blockIdx.x%2 on plane blockIdx.y
// of slot blockIdx.z: formation (as k_overlaps_x) and the inverse y transform.
template <int NY, bool RA>
__global__ __launch_bounds__(ovf_cols_w<NY>() * fft_tpt(NY)) void k_ovf_cols(OvBatch b,
                                                                          const float2* __restrict__ tw) {
  constexpr int W = ovf_cols_w<NY>();
  constexpr int TPT = fft_tpt(NY);
  constexpr int NT = W * TPT;
  constexpr int P = lds_pitch(NY);
  constexpr int EPT = NY * W / NT;
  constexpr int CH = EPT < 4 ? EPT : 4;
  static_assert(EPT % CH == 0, "chunks cover the elements");
  // XCD-contiguous tile order: neighbouring column tiles, whose 64-byte
  // segments of the 257-element band rows share cache lines, run in one L2
  unsigned bx, by, bz;
  xcd_tile3(bx, by, bz);
  const OvParams& p = b.p[bz];
  const int zb = by;
  if (zb > 2 * p.zd) return;  // a slot with a narrower z support (whole block)
  const int which = bx & 1;
  const int c0 = (bx >> 1) * W;
  const int z0 = zb - p.zd;
  float kx = p.kx, ky = p.ky;
  if (p.k0d != nullptr) {
#pragma clang fp contract(off)
    const float2 k = *p.k0d;
    kx = k.x * p.kmul;
    ky = k.y * p.kmul;
  }
  float2* __restrict__ plane = p.ovx + ((long long)which * b.nzs_stride + zb) * NY * p.nx;
  {
    bool live = false;
    for (int w = 0; w < W; ++w) live = live || ov_col_live(p, kx, ky, which, c0 + w, z0);
    if (!live) {
      for (int idx = threadIdx.x; idx < NY * W; idx += NT) {
        const int i = idx / W, w = idx % W;
        if (c0 + w < p.nx) plane[(long long)i * p.nx + c0 + w] = make_float2(0.f, 0.f);
      }
      return;  // uniform over the workgroup
    }
  }
  __shared__ float2 lds[W * P];
  extern __shared__ float2 prof[];  // [2][nxotf]
  const int tid = threadIdx.x;
  FftTw<NY> twr;
  __shared__ float2 twstage[NY];
  load_twiddles(twr, twstage, tid % TPT, tw, tid, NT);
  const int nxotf = p.otf1.nxotf;
  if constexpr (RA)
    for (int k = tid; k < 2 * nxotf; k += NT)
      prof[k] = otf_profile_entry(k < nxotf ? p.otf1 : p.otf2, k < nxotf ? k : k - nxotf, z0, p.kzscale);
  __syncthreads();
  const float2* bre = which == 0 ? p.b1re : p.b2re;
  const float2* bim = which == 0 ? (p.order1 > 0 ? p.b1im : p.b1re) : p.b2im;
  const bool has_im = which == 1 || p.order1 > 0;
#pragma unroll
  for (int q0 = 0; q0 < EPT; q0 += CH) {
    OvSrc src[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int idx = tid + (q0 + c) * NT;
      const int i = idx / W, j = c0 + idx % W;
      src[c] = j < p.nx ? ov_source<RA>(p, kx, ky, prof, prof + nxotf, which, j, i, z0)
                        : OvSrc{0, false, false, make_float2(0.f, 0.f)};
    }
    float2 vre[CH], vim[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      vre[c] = bre[src[c].ind];
      vim[c] = bim[src[c].ind];
    }
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int idx = tid + (q0 + c) * NT;
      lds[(idx % W) * P + lds_swz<NY>(idx / W)] = ov_combine(src[c], vre[c], vim[c], has_im);
    }
  }
  __syncthreads();
  fft_lds<NY, true>(lds + (tid / TPT) * P, tid % TPT, twr);
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int idx = tid + k * NT;
    const int i = idx / W, w = idx % W;
    if (c0 + w < p.nx) plane[(long long)i * p.nx + c0 + w] = lds[w * P + lds_swz<NY>(i)];
  }
}

// planes per round of the row pass: G of each overlap, 2G transforms on the
// radix-stage core
constexpr int kOvfG = 4;
template <int NX>
constexpr bool ovf_rows_ok() {
  return fft_supported(NX) && NX > 1 && 2 * kOvfG * fft_tpt(NX) <= 1024 && 2 * kOvfG * fft_tpt(NX) >= 64 &&
         (2 * kOvfG * lds_pitch(NX) + NX) * 8 <= 96 * 1024 && (2 * kOvfG * NX) % (2 * kOvfG * fft_tpt(NX)) == 0;
}

// Row y = blockIdx.x of every stored plane of both overlaps of slot
// blockIdx.y: the inverse x transforms, kOvfG planes at a time, then
// P(x, y) = nz sum_kz conj(X) Y in plane order (as k_overlaps_zsum), and the
// block's share of the SX and SY totals.
template <int NX>
__global__ __launch_bounds__(2 * kOvfG * fft_tpt(NX)) void k_ovf_rows(ZsumBatch b, const float2* __restrict__ tw) {
#pragma clang fp contract(off)
  constexpr int TPT = fft_tpt(NX);
  constexpr int NT = 2 * kOvfG * TPT;
  constexpr int P = lds_pitch(NX);
  constexpr int EPT = (NX + NT - 1) / NT;  // output elements per thread
  constexpr int LPT = 2 * kOvfG * NX / NT;  // loads per thread per round
  __shared__ float2 lds[2 * kOvfG * P];
  __shared__ double red[NT];
  const int tid = threadIdx.x;
  const int y = blockIdx.x, slot = blockIdx.y;
  const int nzs = b.nzs[slot];
  const long long nxy = b.nxy;
  const float2* X = b.ovx[slot] + (long long)y * NX;
  const float2* Y = X + (long long)b.nzs_stride * nxy;
  FftTw<NX> twr;
  __shared__ float2 twstage[NX];
  load_twiddles(twr, twstage, tid % TPT, tw, tid, NT);
  double sxr[EPT], sxi[EPT], sx[EPT], sy[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) sxr[e] = sxi[e] = sx[e] = sy[e] = 0.0;
  const double n = (double)b.nz;
  for (int g0 = 0; g0 < nzs; g0 += kOvfG) {
    const int cnt = nzs - g0 < kOvfG ? nzs - g0 : kOvfG;  // uniform
    // transform t < kOvfG: X of plane g0 + t; t >= kOvfG: Y of plane g0 + t - kOvfG
    float2 v[LPT];
#pragma unroll
    for (int k = 0; k < LPT; ++k) {
      const int idx = tid + k * NT;
      const int t = idx / NX, x = idx % NX;
      const int pl = t < kOvfG ? t : t - kOvfG;
      const float2* src = t < kOvfG ? X : Y;
      v[k] = src[(long long)(g0 + (pl < cnt ? pl : 0)) * nxy + x];
    }
    __syncthreads();  // the previous round's reads of lds are done
#pragma unroll
    for (int k = 0; k < LPT; ++k) {
      const int idx = tid + k * NT;
      lds[(idx / NX) * P + lds_swz<NX>(idx % NX)] = v[k];
    }
    __syncthreads();
    fft_lds<NX, true>(lds + (tid / TPT) * P, tid % TPT, twr);
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int x = tid + e * NT;
      if (x < NX) {
        const int xs = lds_swz<NX>(x);
        for (int k = 0; k < cnt; ++k) {
          const float2 xa = lds[k * P + xs], yb = lds[(kOvfG + k) * P + xs];
          const double ar = xa.x, ai = xa.y, br = yb.x, bi = yb.y;
          // float x float products are exact in double: a product plus a
          // fused multiply-add is the correctly rounded a*b + c*d
          sxr[e] += __builtin_fma(ar, br, ai * bi);
          sxi[e] += __builtin_fma(ar, bi, -(ai * br));
          sx[e] += __builtin_fma(ar, ar, ai * ai);
          sy[e] += __builtin_fma(br, br, bi * bi);
        }
      }
    }
  }
  double2* Pout = b.P[slot] + (long long)y * NX;
  double tsx = 0.0, tsy = 0.0;
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int x = tid + e * NT;
    if (x < NX) {
      Pout[x] = make_double2(sxr[e] * n, sxi[e] * n);
      tsx += sx[e] * n;
      tsy += sy[e] * n;
    }
  }
  const double a = block_sum_d<NT>(tsx, red);
  const double c = block_sum_d<NT>(tsy, red);
  if (tid == 0) {
    b.totals[slot][y] = a;
    b.totals[slot][gridDim.x + y] = c;
  }
}

template <int NY>
static bool ovf_cols_launch(const OvBatch& ob, int nzs_max, bool run, hipStream_t s) {
  if constexpr (ovf_cols_ok<NY>()) {
    const OvParams& p = ob.p[0];
    const size_t lds = p.otf1.ra ? (size_t)2 * p.otf1.nxotf * sizeof(float2) : 0;
    if (lds > 32 * 1024) return false;
    if (!run) return true;
    const dim3 g1(2 * ceil_div(p.nx, ovf_cols_w<NY>()), nzs_max, ob.n), b1(ovf_cols_w<NY>() * fft_tpt(NY));
    if (p.otf1.ra)
      hipLaunchKernelGGL((k_ovf_cols<NY, true>), g1, b1, lds, s, ob, twiddles(NY));
    else
      hipLaunchKernelGGL((k_ovf_cols<NY, false>), g1, b1, 0, s, ob, twiddles(NY));
    SIMR_LAUNCH_CHECK();
    return true;
  } else {
    return false;
  }
}

template <int NX>
static bool ovf_rows_launch(const ZsumBatch& zb, int ny, bool run, hipStream_t s) {
  if constexpr (ovf_rows_ok<NX>()) {
    if (!run) return true;
    hipLaunchKernelGGL(k_ovf_rows<NX>, dim3(ny, zb.n), dim3(2 * kOvfG * fft_tpt(NX)), 0, s, zb, twiddles(NX));
    SIMR_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_zsum_totals, dim3(zb.n), dim3(256), 0, s, zb, ny);
    SIMR_LAUNCH_CHECK();
    return true;
  } else {
    return false;
  }
}

static int ovf_nzs_max(const OvBatch& ob) {
  int m = 0;
  for (int k = 0; k < ob.n; ++k) m = std::max(m, 2 * ob.p[k].zd + 1);
  return m;
}

static bool ovf_cols_dispatch(const OvBatch& ob, bool run, hipStream_t s) {
  switch (ob.p[0].ny) {
    case 256: return ovf_cols_launch<256>(ob, ovf_nzs_max(ob), run, s);
    case 512: return ovf_cols_launch<512>(ob, ovf_nzs_max(ob), run, s);
    case 1024: return ovf_cols_launch<1024>(ob, ovf_nzs_max(ob), run, s);
    default: return false;
  }
}

static bool ovf_rows_dispatch(const ZsumBatch& zb, int nx, int ny, bool run, hipStream_t s) {
  switch (nx) {
    case 256: return ovf_rows_launch<256>(zb, ny, run, s);
    case 512: return ovf_rows_launch<512>(zb, ny, run, s);
    case 1024: return ovf_rows_launch<1024>(zb, ny, run, s);
    default: return false;
  }
}

bool overlaps_fused_ok(const OvBatch& ob, const ZsumBatch& zb) {
  if (ob.n < 1 || ob.n > kMaxDirs || zb.n != ob.n) return false;
  const OvParams& p0 = ob.p[0];
  if (p0.ny > kZsumMaxBlocks) return false;
  for (int k = 0; k < ob.n; ++k) {
    const OvParams& p = ob.p[k];
    if (p.nx != p0.nx || p.ny != p0.ny || p.otf1.nxotf != p0.otf1.nxotf || p.otf2.nxotf != p0.otf1.nxotf ||
        p.otf1.ra != p0.otf1.ra || p.otf2.ra != p0.otf1.ra)
      return false;
    if (p.zd < 0 || 2 * p.zd + 1 > ob.nzs_stride || zb.nzs[k] != 2 * p.zd + 1 || zb.ovx[k] != p.ovx) return false;
  }
  return ovf_cols_dispatch(ob, false, nullptr) && ovf_rows_dispatch(zb, p0.nx, p0.ny, false, nullptr);
}

void launch_overlaps_fused_cols(const OvBatch& ob, hipStream_t s) {
  if (!ovf_cols_dispatch(ob, true, s)) throw ArgError("two-pass overlaps: unsupported geometry");
}

void launch_overlaps_fused_rows(const ZsumBatch& zb, int nx, int ny, hipStream_t s) {
  if (!ovf_rows_dispatch(zb, nx, ny, true, s)) throw ArgError("two-pass overlaps: unsupported geometry");
}

// Every reserved plane is transformed (planes beyond a slot's own 2zd+1 hold
// stale values that nothing reads), so the slots form one batch of columns.
void launch_overlaps_y(int nx, int ny, int nzs_stride, float2* ovx, int nslots, hipStream_t s) {
  const long long nxy = (long long)nx * ny;
  fft_cols(ny, ovx, nx, nx, 2LL * nzs_stride * nslots, nxy, true, s);
}

// z pass, reduced to what the fit reads.  The modamp and cross-correlation
// kernels use only the per-(x,y) sums over z of conj(X) Y, |X|^2 and |Y|^2 of
// the inverse-transformed overlaps, and the phase factor they apply depends on
// (x, y) alone (reductionKernel, gpuFunctionsImpl.cu:1150-1190; aTimesConjB
// 391-413).  For the unnormalised inverse DFT along z, Parseval gives
// sum_z conj(x_z) y_z = nz sum_kz conj(X_kz) Y_kz, and only the 2zd+1 stored
// kz planes are nonzero, so the sums are formed from those planes directly
// and the z transform is not needed.  Loads of eight planes of both overlaps
// are in flight at once.
// FP32: float sums, stored per (x, y) in P, SX, SY (k_modamp_fp32 reduces
// them in the reference's blocks).  FP64: double sums; P per (x, y), and of
// SX and SY only what the fit reads, their totals, as one partial per block
// (totals[block], totals[nblocks + block]); ELEM also stores them per (x, y).
int zsum_blocks(long long nxy) { return (int)std::min<long long>(ceil_div(nxy, 256), kZsumMaxBlocks); }

template <bool FP32, bool ELEM>
__global__ __launch_bounds__(256) void k_overlaps_zsum(ZsumBatch b) {
#pragma clang fp contract(off)
  using A = typename std::conditional<FP32, float, double>::type;
  constexpr int KB = 8;
  const int slot = blockIdx.y;
  const float2* __restrict__ ovx = b.ovx[slot];
  double2* __restrict__ P = b.P[slot];
  double* __restrict__ SX = b.SX[slot];
  double* __restrict__ SY = b.SY[slot];
  double* __restrict__ totals = b.totals[slot];
  const int nzs = b.nzs[slot];
  const long long nxy = b.nxy;
  const A n = (A)b.nz;
  double tsx = 0.0, tsy = 0.0;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < nxy; e += (long long)gridDim.x * 256) {
    A sxr = 0, sxi = 0, sx = 0, sy = 0;
    const float2* X = ovx + e;
    const float2* Y = ovx + (long long)b.nzs_stride * nxy + e;
    for (int s = 0; s < nzs; s += KB) {
      const int cnt = nzs - s < KB ? nzs - s : KB;  // uniform
      float2 xa[KB], yb[KB];
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        const long long o = (long long)(k < cnt ? s + k : s) * nxy;
        xa[k] = X[o];
        yb[k] = Y[o];
      }
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        if (k >= cnt) break;
        const A ar = xa[k].x, ai = xa[k].y, br = yb[k].x, bi = yb[k].y;
        if constexpr (FP32) {
          sxr += ar * br + ai * bi;
          sxi += ar * bi - ai * br;
          sx += ar * ar + ai * ai;
          sy += br * br + bi * bi;
        } else {
          // float x float products are exact in double, so a product plus a
          // fused multiply-add gives the correctly rounded a*b + c*d of the
          // plain form
          sxr += __builtin_fma(ar, br, ai * bi);
          sxi += __builtin_fma(ar, bi, -(ai * br));
          sx += __builtin_fma(ar, ar, ai * ai);
          sy += __builtin_fma(br, br, bi * bi);
        }
      }
    }
    P[e] = make_double2((double)(sxr * n), (double)(sxi * n));
    if constexpr (FP32 || ELEM) {
      SX[e] = (double)(sx * n);
      SY[e] = (double)(sy * n);
    }
    if constexpr (!FP32) {
      tsx += sx * n;
      tsy += sy * n;
    }
  }
  if constexpr (!FP32) {
    __shared__ double red[256];
    const double a = block_sum_d<256>(tsx, red);
    const double b = block_sum_d<256>(tsy, red);
    if (threadIdx.x == 0) {
      totals[blockIdx.x] = a;
      totals[gridDim.x + blockIdx.x] = b;
    }
  }
}

// The two z-sum totals of each slot (blockIdx.x), summed in block order.
__global__ __launch_bounds__(256) void k_zsum_totals(ZsumBatch b, int zblocks) {
  __shared__ double stage[2 * kSeqChunk];
  double* t = b.totals[blockIdx.x];
  const double v = seq_sum_lanes<double, 2>(zblocks, stage, [&](int i, int k) { return t[k * zblocks + i]; });
  if (threadIdx.x < 2) t[kZsumFinal + threadIdx.x] = v;
}

void launch_overlaps_z_reduce(const ZsumBatch& b, bool fp32, bool per_elem, hipStream_t s) {
  if (b.n < 1 || b.n > kMaxDirs) throw ArgError("z-sum batch: 1..8 slots");
  for (int k = 0; k < b.n; ++k)
    if (b.nzs[k] < 1 || b.nzs[k] > b.nz || b.nzs[k] > b.nzs_stride)
      throw ArgError("overlap z support wider than the stack");
  const dim3 grid(zsum_blocks(b.nxy), b.n);
  if (fp32)
    hipLaunchKernelGGL((k_overlaps_zsum<true, true>), grid, dim3(256), 0, s, b);
  else if (per_elem)
    hipLaunchKernelGGL((k_overlaps_zsum<false, true>), grid, dim3(256), 0, s, b);
  else
    hipLaunchKernelGGL((k_overlaps_zsum<false, false>), grid, dim3(256), 0, s, b);
  SIMR_LAUNCH_CHECK();
  if (!fp32) {
    hipLaunchKernelGGL(k_zsum_totals, dim3(b.n), dim3(256), 0, s, b, (int)grid.x);
    SIMR_LAUNCH_CHECK();
  }
}

// ----------------------------------------------------------------- findk0 peak
__global__ void k_cc_prep(PeakBatch b, float2* __restrict__ cc, long long n) {
  const double2* __restrict__ P = b.P[blockIdx.y];
  cc += (long long)blockIdx.y * n;
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) {
    const double2 v = P[e];
    cc[e] = make_float2((float)v.x, -(float)v.y);  // aTimesConjB: X conj(Y) = conj(conj(X) Y)
  }
}

__device__ __forceinline__ void better(float& v, int& i, float v2, int i2) {
  if (v2 > v || (v2 == v && i2 < i)) {
    v = v2;
    i = i2;
  }
}

__device__ __forceinline__ float intensity(float2 a) {
#pragma clang fp contract(off)
  return a.x * a.x + a.y * a.y;
}

__global__ __launch_bounds__(256) void k_argmax_partial(const float2* __restrict__ cc, long long n,
                                                        float* __restrict__ pv, int* __restrict__ pi) {
  cc += (long long)blockIdx.y * n;
  pv += blockIdx.y * gridDim.x;
  pi += blockIdx.y * gridDim.x;
  __shared__ float sv[256];
  __shared__ int si[256];
  const long long chunk = (n + gridDim.x - 1) / gridDim.x;
  const long long b0 = blockIdx.x * chunk;
  const long long b1 = min(n, b0 + chunk);
  float v = -1e30f;
  int idx = 0x7fffffff;
  for (long long e = b0 + threadIdx.x; e < b1; e += 256) better(v, idx, intensity(cc[e]), (int)e);
  sv[threadIdx.x] = v;
  si[threadIdx.x] = idx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      float a = sv[threadIdx.x];
      int ai = si[threadIdx.x];
      better(a, ai, sv[threadIdx.x + s], si[threadIdx.x + s]);
      sv[threadIdx.x] = a;
      si[threadIdx.x] = ai;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = sv[0];
    pi[blockIdx.x] = si[0];
  }
}

__global__ __launch_bounds__(256) void k_argmax_final(const float2* __restrict__ cc, const float* __restrict__ pv,
                                                      const int* __restrict__ pi, int nparts, int nx, int ny,
                                                      float* __restrict__ vals, int* __restrict__ idx_out) {
  cc += (long long)blockIdx.x * nx * ny;
  pv += blockIdx.x * nparts;
  pi += blockIdx.x * nparts;
  vals += blockIdx.x * 8;
  idx_out += blockIdx.x;
  __shared__ float sv[256];
  __shared__ int si[256];
  float v = -1e30f;
  int idx = 0x7fffffff;
  for (int b = threadIdx.x; b < nparts; b += 256) better(v, idx, pv[b], pi[b]);
  sv[threadIdx.x] = v;
  si[threadIdx.x] = idx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      float a = sv[threadIdx.x];
      int ai = si[threadIdx.x];
      better(a, ai, sv[threadIdx.x + s], si[threadIdx.x + s]);
      sv[threadIdx.x] = a;
      si[threadIdx.x] = ai;
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  idx = si[0];
  // No element won (every intensity NaN, e.g. from one NaN input pixel spread
  // by the transforms): findpeak keeps its initial centre (0, 0)
  // (gpuFunctionsImpl.cu:443-455).
  if (idx < 0 || idx >= nx * ny) idx = 0;
  const int xc = idx % nx, yc = idx / nx;
  auto I = [&](int y, int x) { return intensity(cc[(long long)y * nx + x]); };
  vals[0] = I(yc, xc);
  vals[1] = I(yc, xc == 0 ? nx - 1 : xc - 1);
  vals[2] = I(yc, xc == nx - 1 ? 0 : xc + 1);
  vals[3] = I(yc == 0 ? ny - 1 : yc - 1, xc);
  vals[4] = I(yc == ny - 1 ? 0 : yc + 1, xc);
  idx_out[0] = idx;
  __threadfence_system();  // the outputs may be host memory
}

void launch_findpeak(const PeakBatch& b, float2* cc, float* partial_val, int* partial_idx, float* peak_vals,
                     int* peak_idx, int nx, int ny, hipStream_t s) {
  if (b.n < 1 || b.n > kMaxDirs) throw ArgError("peak batch: 1..8 slots");
  const long long n = (long long)nx * ny;
  hipLaunchKernelGGL(k_cc_prep, dim3(ceil_div(n, 256), b.n), dim3(256), 0, s, b, cc, n);
  SIMR_LAUNCH_CHECK();
  fft_rows(nx, cc, nx, (long long)ny * b.n, false, s);
  fft_cols(ny, cc, nx, nx, b.n, n, false, s);
  hipLaunchKernelGGL(k_argmax_partial, dim3(kArgBlocks, b.n), dim3(256), 0, s, cc, n, partial_val, partial_idx);
  SIMR_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_argmax_final, dim3(b.n), dim3(256), 0, s, cc, partial_val, partial_idx, kArgBlocks, nx, ny,
                     peak_vals, peak_idx);
  SIMR_LAUNCH_CHECK();
}

// ------------------------------------------------------------ modamp, FP64
// e^{i phi} = e^{i a(x)} e^{i b(y)} with a = 2 pi dxy (x - nx/2) kx and
// b = 2 pi dxy (y - ny/2) ky: one table of nx phasors per candidate.
constexpr int kModRows = 4;

int modamp_partials_fp64(int nx, int ny) { return (ny + kModRows - 1) / kModRows; }

__device__ __forceinline__ float2 cand_k(const CandSet& s, int c) { return s.kc ? s.kc[c] : s.k[c]; }
__device__ __forceinline__ int cand_slot(const CandSet& s, int c) { return s.ks ? s.ks[c] : s.slot[c]; }

__global__ __launch_bounds__(256) void k_modamp_fp64(ModampSlots m, CandSet cs, int nx, int ny, float dxy,
                                                     double2* __restrict__ partials) {
  extern __shared__ double2 alpha[];  // nx phasors
  __shared__ double2 beta[kModRows];
  __shared__ double rx[256], ry[256];
  const int c = blockIdx.y;
  if (cs.nd != nullptr && c >= *cs.nd) return;  // whole block
  const double2* __restrict__ P = m.P[cand_slot(cs, c)];
  const float2 kc = cand_k(cs, c);
  const double kx = (double)kc.x, ky = (double)kc.y;
  const double twopi_dxy = 2.0 * 3.14159265358979323846 * (double)dxy;
  for (int i = threadIdx.x; i < nx; i += 256) {
    double sn, cs;
    sincos(twopi_dxy * (((double)i - 0.5 * nx) * kx), &sn, &cs);
    alpha[i] = make_double2(cs, sn);
  }
  const int j0 = blockIdx.x * kModRows;
  if (threadIdx.x < kModRows) {
    double sn, cs;
    sincos(twopi_dxy * (((double)(j0 + threadIdx.x) - 0.5 * ny) * ky), &sn, &cs);
    beta[threadIdx.x] = make_double2(cs, sn);
  }
  __syncthreads();
  double ar = 0.0, ai = 0.0;
  const int nrows = min(kModRows, ny - j0);
  for (int idx = threadIdx.x; idx < nrows * nx; idx += 256) {
    const int jj = idx / nx, i = idx - jj * nx;
    const double2 pv = P[(long long)(j0 + jj) * nx + i];
    const double2 a = alpha[i], b = beta[jj];
    const double er = a.x * b.x - a.y * b.y, ei = a.x * b.y + a.y * b.x;
    ar += pv.x * er - pv.y * ei;
    ai += pv.x * ei + pv.y * er;
  }
  rx[threadIdx.x] = ar;
  ry[threadIdx.x] = ai;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      rx[threadIdx.x] += rx[threadIdx.x + s];
      ry[threadIdx.x] += ry[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[(long long)c * gridDim.x + blockIdx.x] = make_double2(rx[0], ry[0]);
}

static void check_cands(const CandSet& c, int ncand) {
  if (ncand < 1 || (c.kc == nullptr && ncand > kArgCand) || ncand > 65535)
    throw ArgError("modamp: 1..128 candidates per launch (by value), 1..65535 from device arrays");
  if ((c.kc == nullptr) != (c.ks == nullptr)) throw ArgError("modamp: candidates and slots from the same source");
}

void launch_modamp_fp64(const ModampSlots& m, const CandSet& c, int ncand, int nx, int ny, float dxy,
                        double2* partials, hipStream_t s) {
  check_cands(c, ncand);
  dim3 grid(modamp_partials_fp64(nx, ny), ncand);
  hipLaunchKernelGGL(k_modamp_fp64, grid, dim3(256), sizeof(double2) * nx, s, m, c, nx, ny, dxy, partials);
  SIMR_LAUNCH_CHECK();
}

// The result arithmetic of one modamp candidate, in the host's order and
// precision (one thread: the sums are sequential in block order).
__global__ __launch_bounds__(256) void k_modamp_finish(int fp32, ModampSlots m, const int* __restrict__ kslot,
                                                       const double2* __restrict__ part64,
                                                       const float4* __restrict__ part32, int nparts, int zblocks,
                                                       float2* __restrict__ amp, float* __restrict__ corr, int stride,
                                                       int order) {
#pragma clang fp contract(off)
  __shared__ double stage[4 * kSeqChunk];
  __shared__ double sums[4];
  const int c = blockIdx.x, slot = kslot[c];
  const double* __restrict__ totals = m.totals[slot];
  part64 += (long long)c * nparts;
  part32 += (long long)c * nparts;
  amp += slot * stride + order;
  corr += slot * stride + order;
  if (fp32) {
    const float v = seq_sum_lanes<float, 4>(nparts, stage, [&](int i, int k) {
      const float4 q = part32[i];
      return (double)(k == 0 ? q.x : k == 1 ? q.y : k == 2 ? q.z : q.w);
    });
    if (threadIdx.x < 4) sums[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x != 0) return;
    const float xr = (float)sums[0], xi = (float)sums[1], sx = (float)sums[2], sy = (float)sums[3];
    const float mx = xr / sx, my = xi / sx;
    *amp = make_float2(mx, my);
    *corr = sqrtf((xr * xr + xi * xi) / (sx * sy));
  } else {
    const double v = seq_sum_lanes<double, 2>(nparts, stage, [&](int i, int k) {
      return k == 0 ? part64[i].x : part64[i].y;
    });
    if (threadIdx.x < 2) sums[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double xr = sums[0], xi = sums[1], sx = totals[kZsumFinal], sy = totals[kZsumFinal + 1];
    *amp = make_float2((float)(xr / sx), (float)(xi / sx));
    *corr = (float)sqrt((xr * xr + xi * xi) / (sx * sy));
  }
}

void launch_modamp_finish(bool fp32, const ModampSlots& m, const int* kslot, int ncand, const double2* part64,
                          const float4* part32, int nparts, int zblocks, float2* amp, float* corr, int stride,
                          int order, hipStream_t s) {
  if (ncand < 1) return;
  hipLaunchKernelGGL(k_modamp_finish, dim3(ncand), dim3(256), 0, s, fp32 ? 1 : 0, m, kslot, part64, part32, nparts,
                     zblocks, amp, corr, stride, order);
  SIMR_LAUNCH_CHECK();
}

__global__ __launch_bounds__(256) void k_modamp_evals(int fp32, ModampSlots m, CandSet cs,
                                                      const double2* __restrict__ part64,
                                                      const float4* __restrict__ part32, int nparts, int zblocks,
                                                      double* __restrict__ out, int host_out) {
#pragma clang fp contract(off)
  __shared__ double stage[4 * kSeqChunk];
  const int c = blockIdx.x;
  if (cs.nd != nullptr && c >= *cs.nd) return;  // whole block
  out += 4LL * c;
  if (fp32) {
    const float4* pp = part32 + (long long)c * nparts;
    const float v = seq_sum_lanes<float, 4>(nparts, stage, [&](int i, int k) {
      const float4 q = pp[i];
      return (double)(k == 0 ? q.x : k == 1 ? q.y : k == 2 ? q.z : q.w);
    });
    if (threadIdx.x < 4) out[threadIdx.x] = v;
    if (host_out) __threadfence_system();
    return;
  }
  const double2* pp = part64 + (long long)c * nparts;
  const double v = seq_sum_lanes<double, 2>(nparts, stage, [&](int i, int k) { return k == 0 ? pp[i].x : pp[i].y; });
  if (threadIdx.x < 2) out[threadIdx.x] = v;
  if (threadIdx.x == 2 || threadIdx.x == 3)
    out[threadIdx.x] = m.totals[cand_slot(cs, c)][kZsumFinal + threadIdx.x - 2];
  if (host_out) __threadfence_system();
}

void launch_modamp_evals(bool fp32, const ModampSlots& m, const CandSet& c, int ncand, const double2* part64,
                         const float4* part32, int nparts, int zblocks, double* out, bool host_out, hipStream_t s) {
  check_cands(c, ncand);
  hipLaunchKernelGGL(k_modamp_evals, dim3(ncand), dim3(256), 0, s, fp32 ? 1 : 0, m, c, part64, part32, nparts,
                     zblocks, out, host_out ? 1 : 0);
  SIMR_LAUNCH_CHECK();
}

__global__ void k_fit_signal(unsigned* flag, unsigned seq) {
  __threadfence_system();
  __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void launch_fit_signal(unsigned* flag, unsigned seq, hipStream_t s) {
  hipLaunchKernelGGL(k_fit_signal, dim3(1), dim3(1), 0, s, flag, seq);
  SIMR_LAUNCH_CHECK();
}

__device__ __forceinline__ float2 cmul_dev(float2 a, float2 b) {
#pragma clang fp contract(off)
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__global__ void k_amp_policy(AmpPolicy a, float2* __restrict__ amps) {
#pragma clang fp contract(off)
  const int dir = threadIdx.x;
  if (dir >= a.ndirs) return;
  float2* am = amps + dir * a.norders;
  if (a.chain)
    for (int order = 2; order < a.norders; ++order) am[order] = cmul_dev(am[order], am[order - 1]);
  if (a.nforceamp > 0 && a.forceamp[0] > 0.0f) {
    for (int order = 1; order < a.norders; ++order) {
      const float m = sqrtf(am[order].x * am[order].x + am[order].y * am[order].y);
      if (order - 1 < a.nforceamp && m < a.forceamp[order - 1]) {
        const float f = a.forceamp[order - 1] / m;
        am[order].x *= f;
        am[order].y *= f;
      }
    }
  }
  if (!a.fit_all_phases) {
    const float base_phase = atan2f(am[1].y, am[1].x);
    for (int order = 2; order < a.norders; ++order) {
      const float2 amplitude = make_float2(sqrtf(am[order].x * am[order].x + am[order].y * am[order].y), 0.f);
      const float phi = (float)order * base_phase;
      am[order] = cmul_dev(amplitude, make_float2((float)cos((double)phi), (float)sin((double)phi)));
    }
  }
}

void launch_amp_policy(const AmpPolicy& a, float2* amps, hipStream_t s) {
  if (a.ndirs > 64) throw ArgError("amp policy: too many directions");
  hipLaunchKernelGGL(k_amp_policy, dim3(1), dim3(64), 0, s, a, amps);
  SIMR_LAUNCH_CHECK();
}

// ------------------------------------------------------------ modamp, FP32
// reductionKernel on 64x4 blocks, float throughout, the thread z-sums being
// the float P/SX/SY of k_overlaps_zsum; block partials are summed on the host in
// block-index order as cpuReduce does.
int modamp_partials_fp32(int nx, int ny) { return ((nx + 63) / 64) * ((ny + 3) / 4); }

__global__ __launch_bounds__(256) void k_modamp_fp32(ModampSlots m, CandSet cand, int nx, int ny, float dxy,
                                                     float4* __restrict__ partials) {
#pragma clang fp contract(off)
  __shared__ float bx[256], by[256], bs[256], bt[256];
  const int c = blockIdx.z;
  if (cand.nd != nullptr && c >= *cand.nd) return;  // whole block
  const int slot = cand_slot(cand, c);
  const double2* __restrict__ P = m.P[slot];
  const double* __restrict__ SX = m.SX[slot];
  const double* __restrict__ SY = m.SY[slot];
  const float2 kc = cand_k(cand, c);
  const float k0x = kc.x, k0y = kc.y;
  const int i = threadIdx.x + blockIdx.x * 64;
  const int j = threadIdx.y + blockIdx.y * 4;
  const float a = ((float)i - 0.5f * (float)nx) * k0x;
  const float b = ((float)j - 0.5f * (float)ny) * k0y;
  const float angle = (float)(2.0f * 3.14159265358979323846 * (double)(a + b) * (double)dxy);
  float sn, cs;
  sincosf(angle, &sn, &cs);
  float xr = 0.f, xi = 0.f, sx = 0.f, sy = 0.f;
  if (i < nx && j < ny) {
    const long long e = (long long)j * nx + i;
    xr = (float)P[e].x;
    xi = (float)P[e].y;
    sx = (float)SX[e];
    sy = (float)SY[e];
  }
  const float tr = xr * cs - xi * sn;
  const float ti = xr * sn + xi * cs;
  const int lt = threadIdx.y * 64 + threadIdx.x;
  bx[lt] = tr;
  by[lt] = ti;
  bs[lt] = sx;
  bt[lt] = sy;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (lt < s) {
      bx[lt] += bx[lt + s];
      by[lt] += by[lt + s];
      bs[lt] += bs[lt + s];
      bt[lt] += bt[lt + s];
    }
    __syncthreads();
  }
  if (lt == 0) {
    const long long blk = (long long)blockIdx.y * gridDim.x + blockIdx.x;
    partials[(long long)c * gridDim.x * gridDim.y + blk] = make_float4(bx[0], by[0], bs[0], bt[0]);
  }
}

void launch_modamp_fp32(const ModampSlots& m, const CandSet& c, int ncand, int nx, int ny, float dxy,
                        float4* partials, hipStream_t s) {
  check_cands(c, ncand);
  dim3 grid((nx + 63) / 64, (ny + 3) / 4, ncand);
  hipLaunchKernelGGL(k_modamp_fp32, grid, dim3(64, 4), 0, s, m, c, nx, ny, dxy, partials);
  SIMR_LAUNCH_CHECK();
}

// Real-space overlap volumes (the debug dump of the overlap pair,
// cudaSirecon.cpp:330-352,433-445): the inverse DFT along z, unnormalised
// and exp(+i) like the reference's inverse 3D transform, of the 2zd+1 stored
// kz planes (plane j holds kz = j - zd; every other kz is zero) of both
// overlaps: vol[which][z][y][x].  A direct sum over the stored planes; this
// runs once per dumped direction, not in the reconstruction.
__global__ __launch_bounds__(256) void k_overlap_volumes(const float2* __restrict__ ovx, int zd, int nzs_stride, int nz,
                                                         long long nxy, float2* __restrict__ vol) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= nxy) return;
  const int z = blockIdx.y, which = blockIdx.z;
  const int nzs = 2 * zd + 1;
  const float2* src = ovx + (long long)which * nzs_stride * nxy + e;
  float re = 0.f, im = 0.f;
  for (int j = 0; j < nzs; ++j) {
    int m = ((j - zd) * z) % nz;  // exp(2 pi i kz z / nz) depends on kz*z mod nz only
    if (m < 0) m += nz;
    double sd, cd;
    sincospi(2.0 * (double)m / (double)nz, &sd, &cd);
    const float c = (float)cd, sn = (float)sd;
    const float2 v = src[(long long)j * nxy];
    re += v.x * c - v.y * sn;
    im += v.x * sn + v.y * c;
  }
  vol[((long long)which * nz + z) * nxy + e] = make_float2(re, im);
}

void launch_overlap_volumes(const float2* ovx, int zd, int nzs_stride, int nz, long long nxy, float2* vol,
                            hipStream_t s) {
  if (zd < 0 || 2 * zd + 1 > nz || 2 * zd + 1 > nzs_stride) throw ArgError("overlap z support wider than the stack");
  if (nz > 65535) throw ArgError("overlap volumes: too many z planes");
  hipLaunchKernelGGL(k_overlap_volumes, dim3((unsigned)ceil_div(nxy, 256), nz, 2), dim3(256), 0, s, ovx, zd,
                     nzs_stride, nz, nxy, vol);
  SIMR_LAUNCH_CHECK();
}

}  // namespace simr
