// Kernel parameter blocks and host launchers of the reconstruction stages
// (stages.hip).  Everything is passed by value; no __constant__ state, so any
// number of reconstruction contexts can run concurrently on separate streams.
#pragma once
#include "common.hpp"

namespace simr {

constexpr int kMaxDirs = 8;
constexpr int kMaxOrders = 3;
constexpr int kMaxBands = 5;

// OTF table of one (direction, order).
//   ra = 1: rotationally averaged, [kr][kz], kz fastest (nzotf entries per
//           kr, nxotf radial samples; nyotf = 1).
//   ra = 0: the full OTF over kx >= 0, [kz][ky][kx] (nxotf x nyotf per kz
//           plane, nzotf planes, ky and kz in FFT order), read by trilinear
//           interpolation (the reference's GPU path loads such a table but
//           evaluates it as zero, gpuFunctionsImpl.cu:1680-1714).
struct OtfTab {
  const float2* t;
  int nzotf, nxotf;
  float dkrotf;
  int nyotf, ra;
};

// ---- forward path -------------------------------------------------------
struct FwdParams {
  int nx, ny, nz, nxh;
  int ndirs, nphases, nbands;
  int fastSI;
  int napod;
  float inscale;
  float bg0;                         // uniform background when no flat-field maps are set
  float sep[kMaxBands * kMaxBands];  // separation matrix [band][phase], row stride kMaxBands
};

// Flat-fielded section sums: partials[sec * kSumSplit + part] (double).
constexpr int kSumSplit = 8;
// bg/slope == nullptr selects the uniform flat field (bg0, slope 1).
void launch_section_sums(const void* raw, bool u16, const float* bg, const float* slope, float bg0,
                         float inscale, int nx, int ny, int nimg, double* partials, hipStream_t s);
// Bleach-correction ratios per section (reference rescale semantics).
void launch_ratios(const double* partials, double* sdp, float* ratio, int ndirs, int nz, int nphases, int fastSI,
                   int it, int equalizez, int equalizet, int do_rescale, hipStream_t s);
// Fused flatfield + rescale + apodize + phase separation + R2C along x.
// bands: [dir][band][z][y][nxh] complex.
// deskewOneSection (cudaSirecon.cpp:1733-1760) of the flat-fielded sections
// (load_and_flatfield, 1715-1731): raw [nimg][ny][nx_raw] -> out [nimg][ny][nx].
void launch_resample(const void* raw, bool u16, const float* bg, const float* slope, float bg0, float inscale,
                     int nx_raw, int nx, int ny, int nz, int nimg, int ndirs, int nphases, int fastSI, float factor,
                     int shift, float* out, hipStream_t s);
void launch_rows_fwd(const FwdParams& p, const void* raw, bool u16, const float* bg, const float* slope,
                     const float* ratio, const float* apofact, float2* bands, hipStream_t s);

// ---- k0 search --------------------------------------------------------------
// The directions' overlap pairs, cross-correlations and modamp candidates are
// processed together: every launch below covers up to kMaxDirs slots (one per
// direction being fitted), each with its own parameters and buffers.
struct OvParams {
  int nx, ny, nz, nxh;
  int zd;  // z support half-width in planes
  int order1, order2;
  float kx, ky;  // (order2-order1) * k0, 1/um
  float rdistcutoff, otfcutoff, order0_2_factor;
  float dkx, dky, kzscale;
  int bNoKz0;
  OtfTab otf1, otf2;
  const float2 *b1re, *b1im, *b2re, *b2im;  // b1im null for order1 == 0
  // host-derived: rdistcutoff^2, otfcutoff^2, order0_2_factor^2, 1/dkrotf^2 of otf1 and otf2
  float rc2, cut2, f2, inv_dkr2_1, inv_dkr2_2;
  float2* ovx;  // [2][nzs_stride][ny][nx]: overlap0 planes, then overlap1 planes
  // non-null: (kx, ky) = *k0d * kmul, read on the device (a k0 fitted there)
  const float2* k0d;
  float kmul;
};
struct OvBatch {
  OvParams p[kMaxDirs];
  int n;           // slots
  int nzs_stride;  // planes reserved per overlap (2 zd_max + 1 >= 2 zd + 1)
};
// Builds both overlaps of every slot and inverse-transforms them along x.
void launch_overlaps_x(const OvBatch& b, hipStream_t s);
// The y inverse transforms of nslots ovx sets laid out one after another from
// ovx (each [2][nzs_stride][ny][nx]).
void launch_overlaps_y(int nx, int ny, int nzs_stride, float2* ovx, int nslots, hipStream_t s);
// The per-(x,y) z sums of each slot's inverse-transformed overlap pair,
// P = sum_z conj(X) Y, SX = sum |X|^2, SY = sum |Y|^2 (double storage).
// fp32: accumulated in float as the reference's reductionKernel threads, P,
// SX, SY stored per (x, y).  Otherwise in double, P per (x, y) and the totals
// of SX and SY as zsum_blocks(nx*ny) partials each: totals[b], totals[nb + b]
// (per_elem: SX and SY per (x, y) as well), and their sums in block order at
// totals[kZsumFinal], totals[kZsumFinal + 1].
constexpr int kZsumMaxBlocks = 2048;
// totals buffer of one slot: [2][kZsumMaxBlocks] block partials, then the two
// totals summed in block order (fp64, written by the z-sum pass)
constexpr int kZsumFinal = 2 * kZsumMaxBlocks;
constexpr int kZsumTotals = kZsumFinal + 2;
int zsum_blocks(long long nxy);
struct ZsumBatch {
  const float2* ovx[kMaxDirs];
  double2* P[kMaxDirs];
  double* SX[kMaxDirs];
  double* SY[kMaxDirs];
  double* totals[kMaxDirs];
  int nzs[kMaxDirs];
  int n, nzs_stride, nz;
  long long nxy;
};
void launch_overlaps_z_reduce(const ZsumBatch& b, bool fp32, bool per_elem, hipStream_t s);
// The FP64 fit's overlap construction in two passes instead of three: the
// formation with the inverse y transform (by columns, ..._cols), then the
// inverse x transform with the z sums (by rows, ..._rows); P and the totals
// as launch_overlaps_z_reduce(fp32 = false, per_elem = false) leaves them.
// overlaps_fused_ok: whether the geometry is covered.
bool overlaps_fused_ok(const OvBatch& ob, const ZsumBatch& zb);
void launch_overlaps_fused_cols(const OvBatch& ob, hipStream_t s);
void launch_overlaps_fused_rows(const ZsumBatch& zb, int nx, int ny, hipStream_t s);
// Inverse z DFT of the stored kz planes of both overlaps into real-space
// volumes vol[2][nz][nx*ny] (debug dump only).
void launch_overlap_volumes(const float2* ovx, int zd, int nzs_stride, int nz, long long nxy, float2* vol,
                            hipStream_t s);
// crosscorr = conj(P) as float, then its forward 2D FFT and |.|^2 argmax, for
// n slots: cc [n][nx*ny] (contiguous), partials [n][kArgBlocks],
// peak_vals [n][8] (the 5 intensities centre, x-1, x+1, y-1, y+1),
// peak_idx [n] (index of the first maximum).
constexpr int kArgBlocks = 256;
struct PeakBatch {
  const double2* P[kMaxDirs];
  int n;
};
void launch_findpeak(const PeakBatch& b, float2* cc, float* partial_val, int* partial_idx, float* peak_vals,
                     int* peak_idx, int nx, int ny, hipStream_t s);
// Modamp reductions for a batch of candidate vectors kcand[c] on slot kslot[c].
//   fp64: out[c] = sum P e^{i phi_c} over the image in double; partials [ncand][nparts] double2.
//   fp32: reference reductionKernel on 64x4 blocks; partials [ncand][nblocks] of
//         (xsy.x, xsy.y, sx, sy) float4, summed on the host in block order.
struct ModampSlots {
  const double2* P[kMaxDirs];
  const double* SX[kMaxDirs];
  const double* SY[kMaxDirs];
  const double* totals[kMaxDirs];  // the z-sum totals (fp64 finish)
};
// The candidates of one batch: device arrays (kc, ks; the graph's persistent
// ones), or, with kc == nullptr, the values themselves in the kernel
// arguments, so that a line-search round needs no copy.
constexpr int kArgCand = 128;
struct CandSet {
  const float2* kc;
  const int* ks;
  float2 k[kArgCand];
  int slot[kArgCand];
  // non-null: the number of candidates, read on the device (blocks of
  // candidates at or past it do nothing); the launch covers the most there
  // can be
  const int* nd;
};
int modamp_partials_fp64(int nx, int ny);
int modamp_partials_fp32(int nx, int ny);
void launch_modamp_fp64(const ModampSlots& m, const CandSet& c, int ncand, int nx, int ny, float dxy,
                        double2* partials, hipStream_t s);
void launch_modamp_fp32(const ModampSlots& m, const CandSet& c, int ncand, int nx, int ny, float dxy,
                        float4* partials, hipStream_t s);
// The sums the host makes of the modamp partials (cpuReduce,
// gpuFunctionsImpl_hh.cu:162-180), on the device: for candidate c, its
// partials summed in block order and, fp64, its slot's z-sum totals;
// out[4c..4c+3] = (xr, xi, sx, sy) (fp32: the float sums).  host_out: out is
// host memory, and the writes are made visible system-wide before the kernel
// ends.
void launch_modamp_evals(bool fp32, const ModampSlots& m, const CandSet& c, int ncand, const double2* part64,
                         const float4* part32, int nparts, int zblocks, double* out, bool host_out, hipStream_t s);
// Writes seq to *flag (host memory) once the work queued before it is done,
// with system-scope ordering: the host polls the flag instead of an event.
void launch_fit_signal(unsigned* flag, unsigned seq, hipStream_t s);

// ------------------------------------------------------------ device-driven fit
// findk0's peak refinement and fitk0andmodamps' two line searches
// (gpuFunctionsImpl.cu:874-1020) run on the device, every direction at once:
// a walker kernel replays each direction's walk over its cache of evaluated
// (angle, mag) candidates, exactly as the host does, and writes the next
// candidates for the modamp kernels; the host queues a fixed number of
// (evaluate, walk) rounds and reads the results once.
constexpr int kFitLook = 6;                  // steps evaluated ahead in each direction
constexpr int kFitReq = 2 * kFitLook + 1;    // candidates a direction requests in one round
constexpr int kFitMaxRounds = 8;
constexpr int kFitCache = kFitReq * (kFitMaxRounds + 1);  // evaluated candidates kept per direction
enum { kFitAngle = 0, kFitMag = 1, kFitFinal = 2, kFitDone = 3 };
struct FitDir {
  float2 k0;             // findk0's result
  float2 kov;            // vector of the next overlap construction: the fit's start, then the fitted k0
  float k0mag, k0angle;  // the walks' start
  float angle, mag;      // fitted
  int phase, started;    // walk phase; whether its first batch was requested
  int nreq, cand0;       // candidates requested in the last round, their first index
  float req_a[kFitReq], req_m[kFitReq];
  int n_angle_evals, n_mag_evals, n_reductions;
  int ncache;
  float ca[kFitCache], cm[kFitCache];
  double cev[kFitCache][4];
};
struct FitState {
  FitDir d[kMaxDirs];
  float2 kc[kArgCand];  // the round's candidate vectors and their directions
  int ks[kArgCand];
  int ncand;
  int ndone, rounds, overflow;
  float peak_vals[kMaxDirs * 8];  // findk0's cross-correlation peaks
  int peak_idx[kMaxDirs];
  double ev[kArgCand][4];  // the round's sums (xr, xi, sx, sy) per candidate
};
struct FitArgs {
  int ndirs, norders, fo2, fp32, nx, ny;
  float dkx, dky, deltaangle, deltamag;
  float2 guess[kMaxDirs];  // the k0 findk0's overlaps were built at
};
// Refines every direction's cross-correlation peak into k0 (findk0,
// gpuFunctionsImpl.cu:316-370) and sets up the walks.
void launch_fit_start(FitState* fs, const FitArgs& a, hipStream_t s);
// One walker step: merges the last round's sums (fs->ev) into the caches,
// advances every direction's walks, and writes the next round's candidates.
// A direction that finishes gets its fitted k0 (kov) and its order-fo2
// modamp and correlation (amps / corr [dir][norders]).
void launch_fit_walk(FitState* fs, const FitArgs& a, float2* amps, float* corr, hipStream_t s);
// kc[d] = kov[d] * mult for every direction
void launch_fit_kc(const FitState* fs, int ndirs, float mult, float2* kc, hipStream_t s);
// The results for the host, into (host-mapped) out: per direction k0 (the
// search's), kov (the fitted k0) and the counters, then amps and corr
// [dir][norders], then ndone, rounds and overflow.
struct FitOut {
  float2 k0[kMaxDirs], kov[kMaxDirs];
  int n_angle_evals[kMaxDirs], n_mag_evals[kMaxDirs], n_reductions[kMaxDirs];
  float2 amps[kMaxDirs * 5];
  float corr[kMaxDirs * 5];
  int ndone, rounds, overflow;
};
// ... then writes seq into *flag (host memory) once every result is visible
// to the host, as launch_fit_signal does.
void launch_fit_publish(const FitState* fs, const FitArgs& a, const float2* amps, const float* corr, FitOut* out,
                        unsigned* flag, unsigned seq, hipStream_t s);
// findrealspacemodamp's result (gpuFunctionsImpl.cu:1128-1147) on the device,
// for each of ncand candidates: the partials summed in block order as the
// host does (cpuReduce, gpuFunctionsImpl_hh.cu:162-180), then
// modamp = XStarY / sumX and corr = |XStarY| / sqrt(sumX sumY), written to
// amp[kslot[c] * stride + order] and corr[same].
// fp64: part64 [ncand][nparts] and the slot's z-sum totals [2][zblocks];
// fp32: part32 [ncand][nparts].
void launch_modamp_finish(bool fp32, const ModampSlots& m, const int* kslot, int ncand, const double2* part64,
                          const float4* part32, int nparts, int zblocks, float2* amp, float* corr, int stride,
                          int order, hipStream_t s);
// The modamp policy after every order is fitted at a known k0
// (cudaSirecon.cpp:450-491): 2D chaining amp[o] *= amp[o-1], forcemodamp
// (raise magnitudes below the forced ones), and with fit_all_phases == 0 the
// phases of orders >= 2 tied to order 1's.  amps: [ndirs][norders], in place.
struct AmpPolicy {
  int ndirs, norders, chain, nforceamp, fit_all_phases;
  float forceamp[kMaxOrders];
};
void launch_amp_policy(const AmpPolicy& a, float2* amps, hipStream_t s);

// ---- Wiener filter ------------------------------------------------------------
struct FilterParams {
  int nx, ny, nz, nxh;
  // dir, order, conjamp, bp, bp2 describe one band: the host builds the
  // launch's FilterBand entries from them; the kernel reads the entry
  int dir, ndirs, order, norders;
  int zd[kMaxOrders];
  float rdistcutoff, apocutoff, zapocutoff, kzscale, wiener;
  float dkx, dky, min_dkr, suppRadius;
  int bSuppress, bDampenOrder0, bNoKz0, bFilteroverlaps, apodizeoutput;
  int bessel;  // apodization radius x/rdistcutoff, y/apocutoff, z/zapocutoff (gpuFunctionsImpl.cu:1570-1577)
  float2 k0[kMaxDirs];
  float ampmag2_all[kMaxDirs * kMaxOrders];
  float noiseVar[kMaxDirs * kMaxOrders];
  float2 conjamp;  // conj(amp[dir][order])
  // non-null: the modamps [ndirs][norders] on the device (fitted there);
  // ampmag2_all and conjamp are then formed from them in the kernel
  const float2* amps;
  OtfTab otf[kMaxOrders];
  float2 *bp, *bp2;
  // host-derived constants (same float/double expressions as the per-thread ones)
  float inv_nv[kMaxDirs * kMaxOrders];              // (float)(1.0 / (double)noiseVar)
  int unit_nv;                                      // every noiseVar == 1: the divisions are exact no-ops
  float cut2;       // rdistcutoff^2
  float inv_dkr2;   // 1 / dkrotf^2
  float inv_rcut2;  // 1 / rdistcutoff^2
  float supp2;      // suppRadius^2
  float inv_min_dkr2, inv_zd0, inv_rcut, inv_apo, inv_apo2, inv_zapo;  // reciprocals of the divisors
  float inv_nxh;    // 1 / nxh
};
// The bands one filter launch covers: every field of FilterParams that
// differs between (dir, order) comes from the entry (blockIdx.z), the rest
// from the shared FilterParams; the bands are independent, so one launch
// filters a whole stack.
struct FilterBand {
  int dir, order;
  float2 conjamp;
  float2 *bp, *bp2;
  long long otf_off;  // elements added to every otf[o].t (per-direction OTFs)
};
struct FilterBatch {
  int n;
  FilterBand b[kMaxDirs * kMaxOrders];
};
void launch_filter_batch(const FilterParams& p, const FilterBatch& b, hipStream_t s);
// Fill n float2 with zero.
void launch_zero(float2* p, long long n, hipStream_t s);

// ---- assembly ----------------------------------------------------------------
struct AsmParams {
  int nx, ny, nz, nxh;
  int xdim, ydim, zdim;
  int nasm;                 // assembly bands: 0 = order-0 sum over dirs, then (dir, order>=1)
  int ndirs, norders;
  const float2* bands;      // [dir][band][z][y][nxh]
  long long band_stride;    // elements between consecutive bands (H)
  int nbands;
  float2* stageZ;           // [asm][zdim][ny][nx]
  float2* stageY;           // [asm][zdim][ydim][nx]
  float* out;               // [zdim][ydim][xdim]
  float fact, dxyo;
  float2 k0[kMaxDirs];
  // the filter's lateral support: it leaves zeros wherever
  // (x dkx)^2 + (y dky)^2 exceeds rdistcutoff^2; rc2_zero is that square with
  // a margin (0: unknown, nothing is skipped)
  float dkx, dky, rc2_zero;
};
void launch_asm_z(const AsmParams& p, hipStream_t s);
void launch_asm_y(const AsmParams& p, hipStream_t s);
// all assembly bands of stageY -> out (overwritten)
void launch_asm_x(const AsmParams& p, hipStream_t s);

}  // namespace simr
