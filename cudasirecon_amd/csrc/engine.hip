// Host orchestration of one reconstruction context.  See engine.hpp.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>

#include "dtrig.hpp"
#include "engine.hpp"
#include "fft_api.hpp"
#include "sizes.hpp"

namespace simr {

namespace {
constexpr double kPI = 3.14159265358979323846;
constexpr float K0_WARNING_THRESH = 0.05f;  // cudaSireconImpl.h:54

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr int kMaxCand = kArgCand;  // modamp candidates (of every direction) reduced per launch
// the fit's zero-copy host region: the modamp sums (4 doubles each) of two
// queued batches of kMaxCand, the peak values (8 floats per direction) and
// peak indices, and the completion flag
constexpr int kZcCand = 2 * kMaxCand;
constexpr size_t kZcEval = 0;
constexpr size_t kZcPeakV = kZcEval + (size_t)kZcCand * 32;
constexpr size_t kZcPeakI = kZcPeakV + (size_t)kMaxDirs * 32;
constexpr size_t kZcFlag = kZcPeakI + 256;
constexpr size_t kZcFit = kZcFlag + 256;  // FitOut of the device-driven fit
constexpr size_t kZcBytes = kZcFit + ((sizeof(FitOut) + 255) & ~(size_t)255);

int norders_of(const simr_params& p) { return p.norders_output != 0 ? p.norders_output : p.nphases / 2 + 1; }

// makeoverlaps' axial support (gpuFunctionsImpl.cu:508-548), float semantics.
// Without --bessel it does not depend on k0.  With --bessel it is
// ceil((kzExMax sin(halfangle) + (1 - cos(alpha)) / lambdaem) / dkz) with
// cos(halfangle) = |k0| order2 / (norders - 1) / kzExMax; k0mag < 0 gives the
// bound over every k0 (halfangle = pi/2), which sizes the overlap stage.
int overlap_zd(const simr_params& p, float k0mag = -1.f, int order2 = 1, int norders = 2) {
  const int nz = p.nz;
  const float dz = p.dz;
  const float dkz = dz > 0 ? 1.0f / ((float)nz * dz) : p.dkzotf;
  const float lambdaem = (float)((double)((float)p.wave / p.nimm) / 1.0e3);
  const float alpha = asinf(p.na / p.nimm);
  float zd;
  if (!p.bBessel) {
    zd = (float)(int)std::ceil(((1.0 - (double)cosf(alpha)) / (double)lambdaem) / (double)dkz);
  } else {
    const float kzExMax = (float)(2.0 * (double)p.BesselNA / (double)p.BesselLambdaEx);
    float s = 1.0f;
    if (k0mag >= 0.f) {
      const float c = k0mag * (float)order2 / (float)(norders - 1) / kzExMax;
      SIMR_REQUIRE(c <= 1.0f, "Bessel excitation kz extent 2*besselNA/besselExWave is smaller than the "
                              "pattern's lateral frequency |k0|*order/(norders-1)");
      s = sinf(acosf(c));
    }
    const double q = ((double)(kzExMax * s) + (1.0 - (double)cosf(alpha)) / (double)lambdaem) / (double)dkz;
    zd = (float)std::ceil(q);
    // the bound: a margin for the float rounding of sin(acos(c)) near 1, and
    // no clamp, since any value up to it can come out unclamped
    if (k0mag < 0.f) return (int)zd + 1;
  }
  if (zd > (float)(nz / 2)) zd = (float)(((nz / 2 - 1) > 0) ? (nz / 2 - 1) : 0);
  return (int)zd;
}

#define SIMR_CASE(N) case N:
bool in_list_xy(int n) {
  switch (n) {
    SIMR_XY_SIZES(SIMR_CASE)
    return true;
    default:
      return false;
  }
}
bool in_list_zoom(int n) {
  switch (n) {
    SIMR_ZOOM_SIZES(SIMR_CASE)
    return true;
    default:
      return false;
  }
}
bool in_list_z(int n) {
  switch (n) {
    SIMR_Z_SIZES(SIMR_CASE)
    return true;
    default:
      return false;
  }
}
#undef SIMR_CASE
}  // namespace

// ------------------------------------------------------------------- arena
void Arena::reserve(size_t bytes) {
  release();
  if (bytes == 0) return;
  hipError_t e = hipMalloc(&base_, bytes);
  if (e != hipSuccess) {
    base_ = nullptr;
    (void)hipGetLastError();
    throw NoMemError("device arena of " + std::to_string(bytes >> 20) + " MiB could not be allocated: " +
                     hipGetErrorString(e));
  }
  cap_ = bytes;
  off_ = 0;
}

void* Arena::take(size_t bytes) {
  bytes = align256(bytes);
  if (off_ + bytes > cap_) throw NoMemError("device arena exhausted");
  void* p = base_ + off_;
  off_ += bytes;
  return p;
}

void Arena::release() {
  if (base_) (void)hipFree(base_);
  base_ = nullptr;
  cap_ = off_ = 0;
}

// ------------------------------------------------------------------- layout
namespace {
// The fit's device scratch (aliasing the assembly staging region), in the
// order the constructor carves it: every direction's overlap stage, one after
// another; every direction's cross-correlation, one after another; per
// direction P, SX, SY and the z-sum partials; the batch buffers.
size_t ovx_elems(const Layout& L) { return (size_t)2 * (2 * L.zd_ov_max + 1) * L.nxy; }
size_t fit_block_bytes(const Layout& L) {
  size_t b = align256((size_t)L.ndirs * ovx_elems(L) * 8);
  b += align256((size_t)L.ndirs * L.nxy * 8);
  b += (size_t)L.ndirs * (align256((size_t)L.nxy * 16) + 2 * align256((size_t)L.nxy * 8) +
                          align256((size_t)kZsumTotals * 8));
  b += align256((size_t)kMaxCand * modamp_partials_fp64(L.nx, L.ny) * 16);
  b += align256((size_t)kMaxCand * modamp_partials_fp32(L.nx, L.ny) * 16);
  b += 2 * align256((size_t)kMaxDirs * kArgBlocks * 4);
  b += align256(sizeof(FitState));  // device-driven fit
  return b;
}
// The overlap dumps' scratch, in its own region: one overlap stage, P, SX,
// SY and the z-sum partials.
size_t dbg_set_bytes(const Layout& L) {
  return align256(ovx_elems(L) * 8) + align256((size_t)L.nxy * 16) + 2 * align256((size_t)L.nxy * 8) +
         align256((size_t)kZsumTotals * 8);
}
}  // namespace

Layout Layout::of(const simr_params& p) {
  Layout L{};
  SIMR_REQUIRE(p.nx > 0 && p.ny > 0 && p.nz > 0, "image dimensions must be positive");
  SIMR_REQUIRE(p.ndirs >= 1 && p.ndirs <= SIMR_MAX_DIRS, "ndirs must be in 1..8");
  L.norders = norders_of(p);
  SIMR_REQUIRE(L.norders >= 1 && L.norders <= SIMR_MAX_ORDERS, "norders must be in 1..3");
  SIMR_REQUIRE(2 * L.norders - 1 == p.nphases, "nphases must equal 2*norders-1");
  if (p.bRadAvgOTF)
    SIMR_REQUIRE(p.nxotf >= 2 && p.nzotf >= 1 && p.nyotf == 1, "invalid radially averaged OTF geometry");
  else
    SIMR_REQUIRE(p.nxotf >= 2 && p.nyotf >= 1 && p.nzotf >= 1, "invalid OTF geometry");
  SIMR_REQUIRE(p.dkrotf > 0 && p.dkzotf > 0, "OTF pixel sizes must be positive");
  SIMR_REQUIRE(p.dxy > 0, "dxy must be positive");
  SIMR_REQUIRE(p.zoomfact >= 1 && p.z_zoom >= 1, "zoom factors must be >= 1");
  L.nx = p.nx;
  L.ny = p.ny;
  L.nz = p.nz;
  L.nxh = p.nx / 2 + 1;
  L.ndirs = p.ndirs;
  L.nphases = p.nphases;
  L.nbands = 2 * L.norders - 1;
  L.nimg = p.nz * p.ndirs * p.nphases;
  L.nDirsOTF = p.bOneOTFperAngle ? p.ndirs : 1;
  L.H = (long long)L.nxh * p.ny * p.nz;
  L.nxy = (long long)p.nx * p.ny;
  SIMR_REQUIRE(p.nx_raw >= 0, "nx_raw must not be negative");
  L.nx_raw = p.nx_raw > 0 ? p.nx_raw : p.nx;
  L.nxy_raw = (long long)L.nx_raw * p.ny;
  L.resample = L.nx_raw != p.nx || p.deskew_factor != 0.f;
  L.xdim = (int)std::rint((double)p.zoomfact * p.nx);
  L.ydim = (int)std::rint((double)p.zoomfact * p.ny);
  L.zdim = p.z_zoom * p.nz;
  L.nasm = 1 + p.ndirs * (L.norders - 1);
  L.zd_ov_max = overlap_zd(p);
  // Bessel: every per-call support up to the bound, within 2 zd + 1 <= nz
  if (p.bBessel) L.zd_ov_max = std::min(L.zd_ov_max, (p.nz - 1) / 2);
  L.otfsize = (size_t)p.nzotf * p.nxotf * p.nyotf;
  SIMR_REQUIRE(in_list_xy(p.nx), "image width " + std::to_string(p.nx) + " is not a supported size");
  SIMR_REQUIRE(fft_size_supported(p.ny) && p.ny <= 4096, "image height " + std::to_string(p.ny) + " is not a supported size");
  // depths outside the transform-kernel list take the direct-DFT z passes
  SIMR_REQUIRE(p.nz >= 1 && p.nz <= 512, "stack depth " + std::to_string(p.nz) + " is not a supported size (1 to 512)");
  SIMR_REQUIRE(in_list_zoom(L.xdim) && in_list_zoom(L.ydim), "zoomed lateral size is not a supported size");
  SIMR_REQUIRE(L.zdim >= 1 && L.zdim <= 1024, "zoomed stack depth is not a supported size (1 to 1024)");
  SIMR_REQUIRE(2 * L.zd_ov_max + 1 <= p.nz, "overlap z support wider than the stack");
  size_t b = 0;
  b += align256((size_t)L.nimg * L.nxy_raw * 4);        // raw (float or uint16)
  b += 2 * align256((size_t)L.nxy_raw * 4);              // bg, slope
  b += align256((size_t)L.nimg * kSumSplit * 8) + align256((size_t)L.nimg * 4) + align256((size_t)p.nz * 8) + align256(256);
  b += align256((size_t)L.ndirs * L.nbands * L.H * 8);   // bands
  b += align256((size_t)L.nDirsOTF * L.norders * L.otfsize * 8);
  b += 3 * align256((size_t)L.ndirs * SIMR_MAX_ORDERS * 8);  // device modamps of graphed time points
  b += align256((size_t)kMaxDirs * 4);                       // their candidates' direction slots
  b += dbg_set_bytes(L);  // the overlap dumps' scratch
  // assembly staging; the fit scratch aliases it
  const size_t stage = align256((size_t)L.nasm * L.zdim * L.nxy * 8) +            // stageZ
                       align256((size_t)L.nasm * L.zdim * (size_t)L.ydim * p.nx * 8);  // stageY
  // (and the resampled float planes, which live there while the bands are formed)
  b += std::max(std::max(stage, fit_block_bytes(L)), L.resample ? (size_t)L.nimg * (size_t)L.nxy * 4 : (size_t)0);
  b += align256((size_t)L.zdim * L.ydim * (size_t)L.xdim * 4);              // output
  L.bytes_total = b;
  return L;
}

// ------------------------------------------------------------------- engine
Engine::Engine(const simr_params& p, const float* otf, int device) : P(p), L(Layout::of(p)), device_(device) {
  SIMR_REQUIRE(otf != nullptr, "OTF data is required");
  SIMR_HIP(hipSetDevice(device_));
  SIMR_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  fit_fp32_ = P.fit_fp64 == 0;
  // SIMR_FIT_ON_HOST=1: the line searches replay on the host with a read-back
  // per round (the device-driven fit's reference in tests)
  if (const char* e = std::getenv("SIMR_FIT_ON_HOST")) device_fit_ = !(e[0] == '1' && e[1] == 0);
  // SIMR_FIT_ROUNDS=n: the device-driven fit's first rounds (tests of the
  // fall-back to the host walk when they are too few)
  if (const char* e = std::getenv("SIMR_FIT_ROUNDS")) fit_rounds_ = std::min(kFitMaxRounds, std::max(1, std::atoi(e)));
  // SIMR_OVERLAPS_3PASS=1: the three-pass overlap construction everywhere
  if (const char* e = std::getenv("SIMR_OVERLAPS_3PASS")) fused_ov_ = !(e[0] == '1' && e[1] == 0);
  // every twiddle table this context's transforms use (rows, columns, z, the
  // zoomed assembly lengths), so that no launch path allocates one
  for (int n : {L.nx, L.ny, L.nz, L.xdim, L.ydim, L.zdim})
    if (n > 1) (void)twiddles(n);
  arena_.reserve(L.bytes_total);
  raw_ = arena_.take((size_t)L.nimg * L.nxy_raw * 4);
  bg_ = (float*)arena_.take((size_t)L.nxy_raw * 4);
  slope_ = (float*)arena_.take((size_t)L.nxy_raw * 4);
  sums_ = (double*)arena_.take((size_t)L.nimg * kSumSplit * 8);
  ratio_ = (float*)arena_.take((size_t)L.nimg * 4);
  sdp_ = (double*)arena_.take((size_t)L.nz * 8);
  apofact_ = (float*)arena_.take(256);
  bands_ = (float2*)arena_.take((size_t)L.ndirs * L.nbands * L.H * 8);
  otf_ = (float2*)arena_.take((size_t)L.nDirsOTF * L.norders * L.otfsize * 8);
  kc_dev_ = (float2*)arena_.take((size_t)L.ndirs * SIMR_MAX_ORDERS * 8);
  amps_dev_ = (float2*)arena_.take((size_t)L.ndirs * SIMR_MAX_ORDERS * 8);
  corr_dev_ = (float*)arena_.take((size_t)L.ndirs * SIMR_MAX_ORDERS * 8);
  slot_ids_ = (int*)arena_.take((size_t)kMaxDirs * 4);
  char* dbg = (char*)arena_.take(dbg_set_bytes(L));
  const size_t zb = align256((size_t)L.nasm * L.zdim * L.nxy * 8);
  const size_t yb = align256((size_t)L.nasm * L.zdim * (size_t)L.ydim * L.nx * 8);
  stage_bytes_ = std::max(std::max(zb + yb, fit_block_bytes(L)),
                          L.resample ? (size_t)L.nimg * (size_t)L.nxy * 4 : (size_t)0);
  char* stage = (char*)arena_.take(stage_bytes_);
  stageZ_ = (float2*)stage;
  resampled_ = (float*)stage;
  stageY_ = (float2*)(stage + zb);
  out_ = (float*)arena_.take((size_t)L.zdim * L.ydim * (size_t)L.xdim * 4);

  // fit scratch: the layout fit_block_bytes counts, at the start of the staging region
  nzs_stride_ = 2 * L.zd_ov_max + 1;
  {
    char* at = stage;
    auto take = [&](size_t n) {
      char* q = at;
      at += align256(n);
      return q;
    };
    fw_.resize(L.ndirs);
    float2* ovx = (float2*)take((size_t)L.ndirs * ovx_elems(L) * 8);
    float2* cc = (float2*)take((size_t)L.ndirs * L.nxy * 8);
    for (int d = 0; d < L.ndirs; ++d) {
      FitWork& w = fw_[d];
      w.ovx = ovx + (size_t)d * ovx_elems(L);
      w.cc = cc + (size_t)d * L.nxy;
      w.P = (double2*)take((size_t)L.nxy * 16);
      w.SX = (double*)take((size_t)L.nxy * 8);
      w.SY = (double*)take((size_t)L.nxy * 8);
      w.sumparts = (double*)take((size_t)kZsumTotals * 8);
    }
    part64_ = (double2*)take((size_t)kMaxCand * modamp_partials_fp64(L.nx, L.ny) * 16);
    part32_ = (float4*)take((size_t)kMaxCand * modamp_partials_fp32(L.nx, L.ny) * 16);
    pv_ = (float*)take((size_t)kMaxDirs * kArgBlocks * 4);
    pi_ = (int*)take((size_t)kMaxDirs * kArgBlocks * 4);
    fitst_ = (FitState*)take(sizeof(FitState));
    SIMR_REQUIRE((size_t)(at - stage) <= stage_bytes_, "fit scratch larger than the staging region");
  }
  {
    char* at = dbg;
    auto take = [&](size_t n) {
      char* q = at;
      at += align256(n);
      return q;
    };
    dbg_.ovx = (float2*)take(ovx_elems(L) * 8);
    dbg_.P = (double2*)take((size_t)L.nxy * 16);
    dbg_.SX = (double*)take((size_t)L.nxy * 8);
    dbg_.SY = (double*)take((size_t)L.nxy * 8);
    dbg_.sumparts = (double*)take((size_t)kZsumTotals * 8);
  }
  {
    void* h = nullptr;
    void* d = nullptr;
    SIMR_HIP(hipHostMalloc(&h, kZcBytes, hipHostMallocMapped | hipHostMallocCoherent));
    hzc_ = (char*)h;
    SIMR_HIP(hipHostGetDevicePointer(&d, h, 0));
    dzc_ = (char*)d;
    std::memset(hzc_, 0, kZcBytes);
  }
  {
    int ids[kMaxDirs];
    for (int k = 0; k < kMaxDirs; ++k) ids[k] = k;
    SIMR_HIP(hipMemcpyAsync(slot_ids_, ids, sizeof(ids), hipMemcpyHostToDevice, stream_));
    SIMR_HIP(hipStreamSynchronize(stream_));
  }

  SIMR_HIP(hipHostMalloc(&fit_host_, (size_t)L.ndirs * L.norders * 12, hipHostMallocDefault));
  {
    // order 0's modamp is 1 (setup_common, cudaSirecon.cpp:1257-1260); the
    // graphed fit writes the others
    std::vector<float2> one((size_t)L.ndirs * L.norders, make_float2(1.f, 0.f));
    SIMR_HIP(hipMemcpyAsync(amps_dev_, one.data(), one.size() * 8, hipMemcpyHostToDevice, stream_));
    SIMR_HIP(hipStreamSynchronize(stream_));
  }

  // OTF tables
  const size_t notf = (size_t)L.nDirsOTF * L.norders * L.otfsize;
  SIMR_HIP(hipMemcpyAsync(otf_, otf, notf * 8, hipMemcpyHostToDevice, stream_));
  set_flatfield(nullptr, nullptr);
  std::vector<double> zero(L.nz, 0.0);
  SIMR_HIP(hipMemcpyAsync(sdp_, zero.data(), L.nz * 8, hipMemcpyHostToDevice, stream_));
  k0_.assign(L.ndirs, make_float2(0, 0));
  k0_time0_ = k0_;
  derive();
  amp_.assign(L.ndirs, std::vector<float2>(L.norders, make_float2(0, 0)));
  for (int i = 0; i < L.ndirs; ++i) amp_[i][0] = make_float2(1, 0);
  corr_.assign(L.ndirs, std::vector<float>(SIMR_MAX_ORDERS, 0.f));
  stats_.assign(L.ndirs, simr_fit{});
}

// Host-side state derived from the parameters: separation matrix, inscale,
// edge-blend factors and k0 guesses.
void Engine::derive() {
  // separation matrix, equally spaced phases (makematrix, cudaSirecon.cpp:591-617)
  sep_.assign((size_t)L.nbands * L.nphases, 0.f);
  const float phi = (float)(2.0 * kPI / L.nphases);
  for (int j = 0; j < L.nphases; ++j) {
    sep_[0 * L.nphases + j] = 1.0f;
    for (int order = 1; order < L.norders; ++order) {
      const float arg = (float)(j * order) * phi;
      sep_[(2 * order - 1) * L.nphases + j] = (float)std::cos((double)arg);
      sep_[2 * order * L.nphases + j] = (float)std::sin((double)arg);
    }
  }
  // inscale (cudaSirecon.cpp:1231-1232)
  {
    const int ip = L.nx * L.ny * L.nz;
    float f = (float)ip * P.zoomfact;
    f = f * P.zoomfact;
    f = f * (float)P.z_zoom;
    f = f * (float)L.ndirs;
    inscale_ = (float)(1.0 / (double)f);
  }
  // edge-blend factors of apodize_x/y_kernel (gpuFunctionsImpl.cu:61,77)
  SIMR_REQUIRE(P.napodize <= 64, "napodize must be <= 64");
  apof_.assign(64, 0.f);
  for (int l = 0; l < P.napodize; ++l) {
    const double a = ((double)l + 0.5) / (double)(float)P.napodize * kPI * 0.5;
    apof_[l] = (float)(1.0 - (double)sinf((float)a));
  }
  SIMR_HIP(hipMemcpyAsync(apofact_, apof_.data(), 64 * 4, hipMemcpyHostToDevice, stream_));
  SIMR_HIP(hipStreamSynchronize(stream_));

  // k0 guesses (setup_common, cudaSirecon.cpp:1233-1252)
  k0guess_.assign(L.ndirs, make_float2(0, 0));
  const float delta_angle = (float)(kPI / L.ndirs);
  float k0magguess = (float)(1.0 / (double)P.linespacing);
  if (L.nz > 1) k0magguess = k0magguess / (float)(L.nphases / 2 + 1 - 1);
  for (int i = 0; i < L.ndirs; ++i) {
    const float ang = P.nk0angles < L.ndirs ? P.k0startangle + (float)i * delta_angle : P.k0angles[i];
    k0guess_[i] = make_float2((float)((double)k0magguess * std::cos((double)ang)),
                              (float)((double)k0magguess * std::sin((double)ang)));
  }
  SIMR_HIP(hipStreamSynchronize(stream_));
}

// Replace the parameters of a live context.  Everything that sizes device
// buffers or the OTF tables must stay the same; fit state (the time-0 k0 of
// a series) is kept.
void Engine::update_params(const simr_params& p) {
  const Layout N = Layout::of(p);
  SIMR_REQUIRE(N.nx == L.nx && N.ny == L.ny && N.nz == L.nz && N.ndirs == L.ndirs && N.nphases == L.nphases &&
                   N.nx_raw == L.nx_raw && N.resample == L.resample &&
                   N.norders == L.norders && N.xdim == L.xdim && N.ydim == L.ydim && N.zdim == L.zdim &&
                   N.nDirsOTF == L.nDirsOTF && N.otfsize == L.otfsize && N.zd_ov_max <= L.zd_ov_max,
               "parameter change alters the context's buffer layout; create a new context");
  sync();
  refresh_fit();
  P = p;
  fit_fp32_ = P.fit_fp64 == 0;
  derive();
  ++param_version_;  // a captured graph holds the old parameters
}

Engine::~Engine() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  if (fit_host_) (void)hipHostFree(fit_host_);
  if (hzc_) (void)hipHostFree(hzc_);
  for (auto& e : events_) {
    (void)hipEventDestroy(e.second.first);
    (void)hipEventDestroy(e.second.second);
  }
  for (hipEvent_t e : ev_pool_) (void)hipEventDestroy(e);
  arena_.release();
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::sync() { SIMR_HIP(hipStreamSynchronize(stream_)); }

// Waits for the work queued on the context stream.  The fit's read-backs are
// small and on its critical path: a one-thread kernel queued behind the work
// writes a sequence number into the zero-copy region, and the host polls that
// word, which ends the wait within a few microseconds of the device's last
// write (an event's completion signal takes tens of them to be seen).  After
// 2 ms of polling (a busy device, e.g. other contexts' work) it falls back to
// the blocking synchronize, which also reports a failed kernel.
void Engine::wait_fit() {
  const unsigned seq = ++fit_seq_;
  launch_fit_signal((unsigned*)(dzc_ + kZcFlag), seq, stream_);
  wait_flag(seq);
}

// Waits until the flag word holds seq (written by the last kernel queued).
void Engine::wait_flag(unsigned seq) {
  const volatile unsigned* flag = (const volatile unsigned*)(hzc_ + kZcFlag);
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0;; ++i) {
    if (*flag == seq) {
      std::atomic_thread_fence(std::memory_order_acquire);
      return;
    }
    if ((i & 63) == 63 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) break;
    __builtin_ia32_pause();
  }
  SIMR_HIP(hipStreamSynchronize(stream_));
  SIMR_REQUIRE(*flag == seq, "fit completion flag not written");
  std::atomic_thread_fence(std::memory_order_acquire);
}

void Engine::enable_timing(bool on) {
  timing_ = on;
  if (!on) return;
  // the timed region takes its events from here instead of creating them
  std::lock_guard<std::mutex> lk(ev_mu_);
  while (ev_pool_.size() < 2048) {
    hipEvent_t e;
    SIMR_HIP(hipEventCreate(&e));
    ev_pool_.push_back(e);
  }
}

// Stage timing: a pair of events around the stage's launches on its stream.
int Engine::stage_begin(const char* name, hipStream_t s) {
  if (!timing_ || (!timing_only_.empty() && timing_only_ != name)) return -1;
  hipEvent_t a, b;
  int id;
  {
    std::lock_guard<std::mutex> lk(ev_mu_);
    for (hipEvent_t* e : {&a, &b}) {
      if (ev_pool_.empty()) {
        SIMR_HIP(hipEventCreate(e));
      } else {
        *e = ev_pool_.back();
        ev_pool_.pop_back();
      }
    }
    events_.push_back({name, {a, b}});
    id = (int)events_.size() - 1;
  }
  SIMR_HIP(hipEventRecord(a, s));
  return id;
}
void Engine::stage_end(int id, hipStream_t s) {
  if (id < 0) return;
  hipEvent_t b;
  {
    std::lock_guard<std::mutex> lk(ev_mu_);
    b = events_[id].second.second;
  }
  SIMR_HIP(hipEventRecord(b, s));
}

int Engine::timings(const char** names, float* ms, int cap) {
  sync();
  for (auto& e : events_) {
    float t = 0;
    SIMR_HIP(hipEventElapsedTime(&t, e.second.first, e.second.second));
    last_timings_.push_back({e.first, t});
    ev_pool_.push_back(e.second.first);
    ev_pool_.push_back(e.second.second);
  }
  events_.clear();
  static thread_local std::vector<std::string> keep;
  keep.clear();
  int n = 0;
  for (auto& t : last_timings_) {
    if (n < cap) {
      keep.push_back(t.first);
      ms[n] = t.second;
    }
    ++n;
  }
  for (int i = 0; i < (int)keep.size() && i < cap; ++i) names[i] = keep[i].c_str();
  last_timings_.clear();
  return std::min(n, cap);
}

void Engine::set_raw(const void* src, bool u16, bool on_device) {
  SIMR_REQUIRE(src != nullptr, "raw data pointer is null");
  const size_t bytes = (size_t)L.nimg * L.nxy_raw * (u16 ? 2 : 4);
  SIMR_HIP(hipMemcpyAsync(raw_, src, bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream_));
  raw_u16_ = u16;
  raw_set_ = true;
  if (!on_device) sync();
}

void Engine::set_flatfield(const float* bg, const float* slope) {
  // bgAndSlope (cudaSirecon.cpp:139-161): without maps the background is the
  // constant `background` option and the slope 1; that case needs no arrays.
  flat_uniform_ = bg == nullptr && slope == nullptr;
  if (flat_uniform_) return;
  std::vector<float> b((size_t)L.nxy_raw), s((size_t)L.nxy_raw);
  for (long long e = 0; e < L.nxy_raw; ++e) {
    b[e] = bg ? bg[e] : P.constbkgd;
    s[e] = slope ? slope[e] : 1.0f;
  }
  SIMR_HIP(hipMemcpyAsync(bg_, b.data(), L.nxy_raw * 4, hipMemcpyHostToDevice, stream_));
  SIMR_HIP(hipMemcpyAsync(slope_, s.data(), L.nxy_raw * 4, hipMemcpyHostToDevice, stream_));
  sync();
}

// ------------------------------------------------------------ forward path
void Engine::load_and_rescale(int it) {
  if (!raw_set_) throw StateError("raw data has not been set");
  const float* bgp = flat_uniform_ ? nullptr : bg_;
  const float* slp = flat_uniform_ ? nullptr : slope_;
  const void* src = raw_;
  bool src_u16 = raw_u16_;
  float bg0 = P.constbkgd, insc = inscale_;
  if (L.resample) {
    // flat-field and resample into float planes; the stages below then see
    // them with the identity flat-field ((v - 0 - 0) * 1 * 1 = v exactly)
    cur_stage_ = stage_begin("resample", stream_);
    launch_resample(raw_, raw_u16_, bgp, slp, P.constbkgd, inscale_, L.nx_raw, L.nx, L.ny, L.nz, L.nimg, L.ndirs,
                    L.nphases, P.bFastSIM, P.deskew_factor, P.deskew_shift, resampled_, stream_);
    stage_end(cur_stage_, stream_);
    src = resampled_;
    src_u16 = false;
    bgp = slp = nullptr;
    bg0 = 0.f;
    insc = 1.f;
  }
  cur_stage_ = stage_begin("section_sums", stream_);
  launch_section_sums(src, src_u16, bgp, slp, bg0, insc, L.nx, L.ny, L.nimg, sums_, stream_);
  launch_ratios(sums_, sdp_, ratio_, L.ndirs, L.nz, L.nphases, P.bFastSIM, it, P.equalizez, P.equalizet,
                P.do_rescale, stream_);
  stage_end(cur_stage_, stream_);
  FwdParams fp{};
  fp.nx = L.nx;
  fp.ny = L.ny;
  fp.nz = L.nz;
  fp.nxh = L.nxh;
  fp.ndirs = L.ndirs;
  fp.nphases = L.nphases;
  fp.nbands = L.nbands;
  fp.fastSI = P.bFastSIM;
  fp.napod = P.napodize > 0 ? P.napodize : 0;
  fp.inscale = insc;
  fp.bg0 = bg0;
  for (int b = 0; b < L.nbands; ++b)
    for (int ph = 0; ph < L.nphases; ++ph) fp.sep[b * kMaxBands + ph] = sep_[b * L.nphases + ph];
  cur_stage_ = stage_begin("rows_fwd", stream_);
  launch_rows_fwd(fp, src, src_u16, bgp, slp, ratio_, apofact_, bands_, stream_);
  stage_end(cur_stage_, stream_);
  cur_stage_ = stage_begin("cols_fwd_y", stream_);
  fft_cols(L.ny, bands_, L.nxh, L.nxh, (long long)L.ndirs * L.nbands * L.nz, (long long)L.nxh * L.ny, false,
           stream_);
  stage_end(cur_stage_, stream_);
  if (L.nz > 1) {
    cur_stage_ = stage_begin("cols_fwd_z", stream_);
    const long long plane = (long long)L.nxh * L.ny;
    if (in_list_z(L.nz)) {
      fft_cols(L.nz, bands_, plane, plane, (long long)L.ndirs * L.nbands, L.H, false, stream_);
    } else {
      // depths without a transform kernel: direct DFT along z (dft_cols),
      // (direction, band) volumes in chunks through the
      // assembly staging region, which is idle until the fit and the assembly
      const long long vols = (long long)L.ndirs * L.nbands;
      const long long chunk =
          std::max<long long>(1, std::min<long long>(vols, (long long)(stage_bytes_ / ((size_t)L.H * 8))));
      SIMR_REQUIRE((size_t)L.H * 8 <= stage_bytes_, "staging region smaller than one band volume");
      for (long long v0 = 0; v0 < vols; v0 += chunk) {
        const long long nv = std::min(chunk, vols - v0);
        float2* src = bands_ + v0 * L.H;
        dft_cols(src, stageZ_, L.nz, plane, plane, L.H, nv, false, stream_);
        SIMR_HIP(hipMemcpyAsync(src, stageZ_, (size_t)nv * L.H * 8, hipMemcpyDeviceToDevice, stream_));
      }
    }
    stage_end(cur_stage_, stream_);
  }
  preprocessed_ = true;
  fitted_ = processed_ = false;
}

// ------------------------------------------------------------ geometry
Engine::Geo Engine::geometry() const {
  Geo g;
  g.dkx = 1.0f / ((float)L.nx * P.dxy);
  g.dky = 1.0f / ((float)L.ny * P.dxy);
  g.dkz = P.dz > 0 ? 1.0f / ((float)L.nz * P.dz) : P.dkzotf;
  g.kzscale = g.dkz / P.dkzotf;
  g.rdistcutoff = (float)((double)P.na * 2.0 / ((double)P.wave * 0.001));
  if ((double)g.rdistcutoff > 1. / (2. * (double)P.dxy)) g.rdistcutoff = (float)(1. / (2. * (double)P.dxy));
  g.lambdaem = (float)((double)((float)P.wave / P.nimm) / 1.0e3);
  g.lambdaexc = (float)(0.88 * (double)g.lambdaem);
  g.alpha = asinf(P.na / P.nimm);
  return g;
}

int Engine::zd_overlaps() const { return L.zd_ov_max; }

// fitorder2 of findk0 / fitk0andmodamps (gpuFunctionsImpl.cu:316-323, 880-887)
int Engine::fit_order2() const { return L.nz > 1 && !P.bBessel ? 2 : 1; }

OtfTab Engine::otf_table(int dir, int order) const {
  const int d = P.bOneOTFperAngle ? dir : 0;
  OtfTab o;
  o.t = otf_ + ((size_t)d * L.norders + order) * L.otfsize;
  o.nzotf = P.nzotf;
  o.nxotf = P.nxotf;
  o.dkrotf = P.dkrotf;
  o.nyotf = P.bRadAvgOTF ? 1 : P.nyotf;
  o.ra = P.bRadAvgOTF ? 1 : 0;
  return o;
}

// ------------------------------------------------------------ overlaps / fit
OvParams Engine::ov_params(int dir, int o1, int o2, float k0x, float k0y) const {
  const Geo g = geometry();
  OvParams p{};
  p.nx = L.nx;
  p.ny = L.ny;
  p.nz = L.nz;
  p.nxh = L.nxh;
  if (P.bBessel) {
    // the reference's k0mag = sqrt(k0x*k0x + k0y*k0y) of the float operands
    const float k0mag = (float)std::sqrt((double)(k0x * k0x + k0y * k0y));
    p.zd = overlap_zd(P, k0mag, o2, L.norders);
    SIMR_REQUIRE(p.zd <= L.zd_ov_max, "overlap z support wider than the stack");
  } else {
    p.zd = zd_overlaps();
  }
  p.order1 = o1;
  p.order2 = o2;
  p.kx = k0x * (float)(o2 - o1);
  p.ky = k0y * (float)(o2 - o1);
  p.rdistcutoff = g.rdistcutoff;
  p.otfcutoff = P.bBessel ? 0.01f : P.otfcutoff;
  p.order0_2_factor = L.nz > 1 ? (P.bBessel ? 4.0f : 5.0f) : 1.0f;
  p.dkx = g.dkx;
  p.dky = g.dky;
  p.kzscale = g.kzscale;
  p.bNoKz0 = P.bNoKz0;
  p.otf1 = otf_table(dir, o1);
  p.otf2 = otf_table(dir, o2);
  const float2* db = bands_ + (size_t)dir * L.nbands * L.H;
  p.b1re = o1 == 0 ? db : db + (size_t)(2 * o1 - 1) * L.H;
  p.b1im = o1 == 0 ? nullptr : db + (size_t)(2 * o1) * L.H;
  p.b2re = db + (size_t)(2 * o2 - 1) * L.H;
  p.b2im = db + (size_t)(2 * o2) * L.H;
  p.rc2 = p.rdistcutoff * p.rdistcutoff;
  p.cut2 = p.otfcutoff * p.otfcutoff;
  p.f2 = p.order0_2_factor * p.order0_2_factor;
  p.inv_dkr2_1 = (float)(1.0 / ((double)p.otf1.dkrotf * (double)p.otf1.dkrotf));
  p.inv_dkr2_2 = (float)(1.0 / ((double)p.otf2.dkrotf * (double)p.otf2.dkrotf));
  return p;
}

// makeoverlaps (gpuFunctionsImpl.cu:497-675) for every request, at most one
// per direction, each in its direction's slot (or all in `only`, the dumps'
// scratch, for a single request): one launch per pass covers every request.
void Engine::make_overlaps(const std::vector<OvReq>& reqs, bool per_elem, FitWork* only) {
  SIMR_REQUIRE(!reqs.empty() && (int)reqs.size() <= kMaxDirs && (!only || reqs.size() == 1),
               "overlap batch: 1..8 requests");
  OvBatch b{};
  ZsumBatch z{};
  b.n = z.n = (int)reqs.size();
  b.nzs_stride = z.nzs_stride = nzs_stride_;
  z.nz = L.nz;
  z.nxy = L.nxy;
  int dmin = L.ndirs, dmax = -1;
  for (int k = 0; k < b.n; ++k) {
    const OvReq& r = reqs[k];
    SIMR_REQUIRE(r.dir >= 0 && r.dir < L.ndirs, "direction out of range");
    FitWork& w = only ? *only : fw_[r.dir];
    for (int j = 0; j < k; ++j) SIMR_REQUIRE(reqs[j].dir != r.dir, "overlap batch: one request per direction");
    SIMR_REQUIRE(r.k0d == nullptr || !P.bBessel, "device-side overlap shifts need a fixed z support");
    b.p[k] = ov_params(r.dir, r.o1, r.o2, r.k0x, r.k0y);
    b.p[k].ovx = w.ovx;
    b.p[k].k0d = r.k0d;
    b.p[k].kmul = r.kmul;
    z.ovx[k] = w.ovx;
    z.P[k] = w.P;
    z.SX[k] = w.SX;
    z.SY[k] = w.SY;
    z.totals[k] = w.sumparts;
    z.nzs[k] = 2 * b.p[k].zd + 1;
    w.zd = b.p[k].zd;
    w.cache.clear();
    dmin = std::min(dmin, r.dir);
    dmax = std::max(dmax, r.dir);
  }
  // the FP64 fit's overlaps in two passes where the sizes allow
  if (!only && !per_elem && !fit_fp32_ && fused_ov_ && overlaps_fused_ok(b, z)) {
    int st = stage_begin("ov_cols", stream_);
    launch_overlaps_fused_cols(b, stream_);
    stage_end(st, stream_);
    st = stage_begin("ov_rows", stream_);
    launch_overlaps_fused_rows(z, L.nx, L.ny, stream_);
    stage_end(st, stream_);
    return;
  }
  int st = stage_begin("ov_x", stream_);
  launch_overlaps_x(b, stream_);
  stage_end(st, stream_);
  st = stage_begin("ov_y", stream_);
  // the directions' stages lie one after another: one batch from the first
  // requested to the last (a direction in between is transformed unused)
  if (only)
    launch_overlaps_y(L.nx, L.ny, nzs_stride_, only->ovx, 1, stream_);
  else
    launch_overlaps_y(L.nx, L.ny, nzs_stride_, fw_[dmin].ovx, dmax - dmin + 1, stream_);
  stage_end(st, stream_);
  st = stage_begin("ov_zsum", stream_);
  launch_overlaps_z_reduce(z, fit_fp32_, per_elem, stream_);
  stage_end(st, stream_);
}

ModampSlots Engine::modamp_slots() const {
  ModampSlots m{};
  for (int d = 0; d < L.ndirs; ++d) {
    m.P[d] = fw_[d].P;
    m.SX[d] = fw_[d].SX;
    m.SY[d] = fw_[d].SY;
    m.totals[d] = fw_[d].sumparts;
  }
  return m;
}

// Queues the modamp reductions of n <= kMaxCand candidates (direction,
// vector) at entries [at, at + n) of the zero-copy region: the candidates
// travel in the kernel arguments, and the partials are summed per candidate in
// block order on the device as the host does (cpuReduce,
// gpuFunctionsImpl_hh.cu:162-180), straight into the region.  The sums are
// read (eval_read) after the next wait_fit().
void Engine::eval_enqueue(const Cand* c, int n, int at) {
  SIMR_REQUIRE(n >= 1 && n <= kMaxCand && at >= 0 && at + n <= kZcCand, "modamp batch out of range");
  CandSet cs{};  // every pointer null: candidates by value, sums by k_modamp_evals
  for (int i = 0; i < n; ++i) {
    SIMR_REQUIRE(c[i].first >= 0 && c[i].first < L.ndirs, "direction out of range");
    cs.k[i] = c[i].second;
    cs.slot[i] = c[i].first;
  }
  for (int i = n; i < kArgCand; ++i) {
    cs.k[i] = make_float2(0.f, 0.f);
    cs.slot[i] = 0;
  }
  const ModampSlots m = modamp_slots();
  const int nb = fit_fp32_ ? modamp_partials_fp32(L.nx, L.ny) : modamp_partials_fp64(L.nx, L.ny);
  double* dev = (double*)(dzc_ + kZcEval) + 4 * (size_t)at;
  const int st = stage_begin("modamp", stream_);
  if (fit_fp32_)
    launch_modamp_fp32(m, cs, n, L.nx, L.ny, P.dxy, part32_, stream_);
  else
    launch_modamp_fp64(m, cs, n, L.nx, L.ny, P.dxy, part64_, stream_);
  launch_modamp_evals(fit_fp32_, m, cs, n, part64_, part32_, nb, zsum_blocks(L.nxy), dev, true, stream_);
  stage_end(st, stream_);
}

void Engine::eval_read(int n, int at, Eval* out) const {
  const double* ev = (const double*)(hzc_ + kZcEval) + 4 * (size_t)at;
  for (int i = 0; i < n; ++i) out[i] = {ev[4 * i], ev[4 * i + 1], ev[4 * i + 2], ev[4 * i + 3]};
}

// The modamp reductions of every direction's candidates (req: direction and
// its candidate vectors), in rounds of up to kMaxCand candidates: one copy,
// two kernels and one wait each.
std::vector<std::vector<Engine::Eval>> Engine::eval_multi(
    const std::vector<std::pair<int, std::vector<float2>>>& req) {
  std::vector<Cand> flat;
  for (auto& r : req)
    for (auto& k : r.second) flat.push_back({r.first, k});
  std::vector<Eval> ev(flat.size());
  for (size_t c0 = 0; c0 < flat.size(); c0 += kMaxCand) {
    const int nc = (int)std::min<size_t>(kMaxCand, flat.size() - c0);
    eval_enqueue(flat.data() + c0, nc, 0);
    wait_fit();
    eval_read(nc, 0, ev.data() + c0);
  }
  std::vector<std::vector<Eval>> res(req.size());
  size_t i = 0;
  for (size_t r = 0; r < req.size(); ++r)
    for (size_t c = 0; c < req[r].second.size(); ++c) res[r].push_back(ev[i++]);
  return res;
}

// findrealspacemodamp's result arithmetic (gpuFunctionsImpl.cu:1128-1147)
void Engine::modamp_of(const Eval& e, float2* modamp, double* amp2, float* corr) const {
  if (fit_fp32_) {
    const float xr = (float)e.xr, xi = (float)e.xi, sx = (float)e.sx, sy = (float)e.sy;
    const float mx = xr / sx, my = xi / sx;
    *modamp = make_float2(mx, my);
    const float a2 = mx * mx + my * my;
    *amp2 = (double)a2;
    const float cc = (xr * xr + xi * xi) / (sx * sy);
    if (corr) *corr = sqrtf(cc);
  } else {
    const double mx = e.xr / e.sx, my = e.xi / e.sx;
    *modamp = make_float2((float)mx, (float)my);
    *amp2 = mx * mx + my * my;
    if (corr) *corr = (float)std::sqrt((e.xr * e.xr + e.xi * e.xi) / (e.sx * e.sy));
  }
}

static std::pair<uint32_t, uint32_t> key_of(float a, float b) {
  uint32_t x, y;
  std::memcpy(&x, &a, 4);
  std::memcpy(&y, &b, 4);
  return {x, y};
}

static float2 kvec_of(float angle, float mag, int o1, int o2) {
  // getmodamp (gpuFunctionsImpl.cu:1058-1059) then findrealspacemodamp (1093-1094)
  const float kx = mag * dtrig::fcos(angle), ky = mag * dtrig::fsin(angle);
  return make_float2(kx * (float)(o2 - o1), ky * (float)(o2 - o1));
}

// Evaluates the (angle, mag) candidates of each direction that its cache
// does not hold yet, all in one batch.
void Engine::prefetch_multi(const std::vector<std::pair<int, Walk>>& req, int o1, int o2) {
  std::vector<std::pair<int, std::vector<float2>>> kv;
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> keys;
  for (auto& r : req) {
    FitWork& w = fw_[r.first];
    std::vector<float2> v;
    std::vector<std::pair<uint32_t, uint32_t>> ks;
    for (auto& x : r.second) {
      auto k = key_of(x.first, x.second);
      if (w.cache.count(k) || std::find(ks.begin(), ks.end(), k) != ks.end()) continue;
      ks.push_back(k);
      v.push_back(kvec_of(x.first, x.second, o1, o2));
    }
    if (v.empty()) continue;
    kv.push_back({r.first, v});
    keys.push_back(ks);
  }
  if (kv.empty()) return;
  auto ev = eval_multi(kv);
  for (size_t r = 0; r < kv.size(); ++r)
    for (size_t i = 0; i < keys[r].size(); ++i) fw_[kv[r].first].cache[keys[r][i]] = ev[r][i];
}

// findrealspacemodamp at each direction's k0_ with overlaps of orders
// (o1, o2) built there (redoarrays): amp_[dir][o2] and corr_[dir][o2].
// npend evaluations already queued at entries [0, npend) come back in *pend,
// read after the same wait.
void Engine::realspacemodamp_multi(const std::vector<int>& dirs, int o1, int o2, int npend,
                                   std::vector<Eval>* pend) {
  if (dirs.empty() && npend == 0) return;
  std::vector<Cand> c;
  if (!dirs.empty()) {
    std::vector<OvReq> reqs;
    for (int d : dirs) {
      reqs.push_back({d, o1, o2, k0_[d].x, k0_[d].y});
      c.push_back({d, make_float2(k0_[d].x * (float)(o2 - o1), k0_[d].y * (float)(o2 - o1))});
    }
    make_overlaps(reqs);
    eval_enqueue(c.data(), (int)c.size(), npend);
  }
  wait_fit();
  if (npend > 0) {
    pend->resize(npend);
    eval_read(npend, 0, pend->data());
  }
  std::vector<Eval> ev(c.size());
  if (!c.empty()) eval_read((int)c.size(), npend, ev.data());
  for (size_t i = 0; i < dirs.size(); ++i) {
    const int d = dirs[i];
    stats_[d].n_reductions++;
    double a2;
    float2 m;
    float cc;
    modamp_of(ev[i], &m, &a2, &cc);
    amp_[d][o2] = m;
    corr_[d][o2] = cc;
  }
}

// findk0 (gpuFunctionsImpl.cu:306-372) with findpeak/fitparabola (443-495)
static float fitparabola(float a1, float a2, float a3) {
  const float slope = (float)(0.5 * (double)(a3 - a1));
  const float curve = (a3 + a1) - 2 * a2;
  if (curve == 0) return 0.0f;
  const float peak = -slope / curve;
  if (peak > 1.5f || peak < -1.5f) return 0.0f;
  return peak;
}

void Engine::findk0_all(const std::vector<int>& dirs) {
  if (dirs.empty()) return;
  const int fo2 = fit_order2();
  std::vector<OvReq> reqs;
  int dmin = L.ndirs, dmax = -1;
  for (int d : dirs) {
    reqs.push_back({d, 0, fo2, k0_[d].x, k0_[d].y});
    dmin = std::min(dmin, d);
    dmax = std::max(dmax, d);
  }
  make_overlaps(reqs);
  // the cross-correlations of directions dmin..dmax, one after another
  PeakBatch pb{};
  pb.n = dmax - dmin + 1;
  for (int k = 0; k < pb.n; ++k) pb.P[k] = fw_[dmin + k].P;
  const int st = stage_begin("findk0_peak", stream_);
  // the peak values and indices land in the zero-copy region
  launch_findpeak(pb, fw_[dmin].cc, pv_, pi_, (float*)(dzc_ + kZcPeakV), (int*)(dzc_ + kZcPeakI), L.nx, L.ny,
                  stream_);
  stage_end(st, stream_);
  wait_fit();
  const Geo g = geometry();
  for (int d : dirs) {
    const float* v = (const float*)(hzc_ + kZcPeakV) + (size_t)(d - dmin) * 8;
    const int idx = ((const int*)(hzc_ + kZcPeakI))[d - dmin];
    const int xc = idx % L.nx, yc = idx / L.nx;
    float2 pk;
    pk.x = fitparabola(v[1], v[0], v[2]) + (float)xc;
    pk.y = fitparabola(v[3], v[0], v[4]) + (float)yc;
    const float2 old = k0_[d];
    if (old.x / g.dkx < pk.x - (float)(L.nx / 2)) pk.x -= (float)L.nx;
    if (old.x / g.dkx > pk.x + (float)(L.nx / 2)) pk.x += (float)L.nx;
    if (old.y / g.dky < pk.y - (float)(L.ny / 2)) pk.y -= (float)L.ny;
    if (old.y / g.dky > pk.y + (float)(L.ny / 2)) pk.y += (float)L.ny;
    pk.x *= g.dkx / (float)fo2;
    pk.y *= g.dky / (float)fo2;
    k0_[d] = pk;
  }
}

template <typename U>
static U fitxyparabola(U x1, U y1, U x2, U y2, U x3, U y3) {
  if (x1 == x2 || x2 == x3 || x3 == x1) return 0;
  const U xbar1 = (U)(0.5 * (double)(x1 + x2));
  const U xbar2 = (U)(0.5 * (double)(x2 + x3));
  const U slope1 = (y2 - y1) / (x2 - x1);
  const U slope2 = (y3 - y2) / (x3 - x2);
  const U curve = (slope2 - slope1) / (xbar2 - xbar1);
  if (curve == 0) return 0;
  return xbar2 - slope2 / curve;
}

namespace {
constexpr int kLook = 6;  // line-search steps evaluated ahead in each direction

// The next kLook points of a walk from `start` (angle walks vary the angle at
// a fixed magnitude, magnitude walks the magnitude at a fixed angle), by the
// reference's float accumulation of the step.
std::vector<std::pair<float, float>> walk_from(float start, float other, bool is_angle, float delta, int sign) {
  std::vector<std::pair<float, float>> v;
  float a = start;
  for (int m = 0; m < kLook; ++m) {
    v.push_back(is_angle ? std::make_pair(a, other) : std::make_pair(other, a));
    a = sign > 0 ? a + delta : a - delta;
  }
  return v;
}
}  // namespace

// One line-search walk of fitk0andmodamps (gpuFunctionsImpl.cu:912-983):
// from `start` step by `delta` while the modamp magnitude grows, then the
// parabola through the last three points.  The steps the reference visits
// are deterministic, so the walk is replayed over the direction's cache of
// evaluated candidates: true (and the vertex, and the number of steps) when
// every value it visits is cached; otherwise false, with `need` the next
// kLook steps from the first value missing.
bool Engine::replay_walk(int dir, bool is_angle, float start, float other, float delta, int o1, int o2, float* out,
                         int* nevals, Walk* need) {
  FitWork& w = fw_[dir];
  int n = 0;
  int sign = +1;
  auto f = [&](float a, double* amp) {
    auto k = is_angle ? key_of(a, other) : key_of(other, a);
    auto it = w.cache.find(k);
    if (it == w.cache.end()) {
      *need = walk_from(a, other, is_angle, delta, sign);
      return false;
    }
    float2 m;
    modamp_of(it->second, &m, amp, nullptr);
    ++n;
    return true;
  };
  float x2 = start;
  double amp2, amp3, amp1 = 0;
  if (!f(start, &amp2)) return false;
  float a = start + delta;
  float x3 = a;
  if (!f(a, &amp3)) return false;
  float x1 = 0;
  if (amp3 > amp2) {
    while (amp3 > amp2) {
      amp1 = amp2;
      x1 = x2;
      amp2 = amp3;
      x2 = x3;
      a += delta;
      x3 = a;
      if (!f(a, &amp3)) return false;
    }
  } else {
    a = start;
    double t = amp3;
    amp3 = amp2;
    amp2 = t;
    float tx = x3;
    x3 = x2;
    x2 = tx;
    sign = -1;
    while (amp3 > amp2) {
      amp1 = amp2;
      x1 = x2;
      amp2 = amp3;
      x2 = x3;
      a -= delta;
      x3 = a;
      if (!f(a, &amp3)) return false;
    }
  }
  if (!fit_fp32_)
    *out = (float)fitxyparabola<double>(x1, amp1, x2, amp2, x3, amp3);
  else
    *out = fitxyparabola<float>(x1, (float)amp1, x2, (float)amp2, x3, (float)amp3);
  *nevals = n;
  return true;
}

// fitk0andmodamps (gpuFunctionsImpl.cu:874-1020) for every direction in
// `dirs` at once.  The reference evaluates the line-search steps one at a
// time with a synchronous reduction each; here each walk's next kLook steps
// in both directions are reduced together, every direction's in the same
// launch, and each direction's walk is replayed on the results; a round
// fetches what the walks that are not done yet need next.
void Engine::fitk0_all(const std::vector<int>& dirs) {
  if (dirs.empty()) return;
  const int fo1 = 0, fo2 = fit_order2();
  const float deltaangle = 0.001f;
  const float deltamag = (float)(0.1 / (double)((float)std::max(L.nx, L.ny) * P.dxy));
  std::vector<float> k0mag(L.ndirs), k0angle(L.ndirs), angle(L.ndirs), mag(L.ndirs);
  // first evaluation recomputes the overlaps (redoarrays = recalcarrays >= 1)
  {
    std::vector<OvReq> reqs;
    for (int d : dirs) {
      const float2 kv = k0_[d];
      k0mag[d] = sqrtf(kv.x * kv.x + kv.y * kv.y);
      k0angle[d] = dtrig::fatan2(kv.y, kv.x);
      reqs.push_back({d, fo1, fo2, k0mag[d] * dtrig::fcos(k0angle[d]), k0mag[d] * dtrig::fsin(k0angle[d])});
    }
    make_overlaps(reqs);
  }
  auto run_walks = [&](bool is_angle) {
    std::vector<std::pair<int, Walk>> first;
    for (int d : dirs) {
      const float s0 = is_angle ? k0angle[d] : k0mag[d];
      const float other = is_angle ? k0mag[d] : angle[d];
      const float delta = is_angle ? deltaangle : deltamag;
      // the start and the first kLook steps both ways in one batch
      Walk all{is_angle ? std::make_pair(s0, other) : std::make_pair(other, s0)};
      auto up = walk_from(s0 + delta, other, is_angle, delta, +1);
      auto dn = walk_from(s0 - delta, other, is_angle, delta, -1);
      all.insert(all.end(), up.begin(), up.end());
      all.insert(all.end(), dn.begin(), dn.end());
      first.push_back({d, all});
    }
    prefetch_multi(first, fo1, fo2);
    std::vector<bool> done(L.ndirs, false);
    for (;;) {
      std::vector<std::pair<int, Walk>> more;
      for (int d : dirs) {
        if (done[d]) continue;
        const float s0 = is_angle ? k0angle[d] : k0mag[d];
        const float other = is_angle ? k0mag[d] : angle[d];
        float r;
        int ne = 0;
        Walk need;
        if (replay_walk(d, is_angle, s0, other, is_angle ? deltaangle : deltamag, fo1, fo2, &r, &ne, &need)) {
          (is_angle ? angle : mag)[d] = r;
          (is_angle ? stats_[d].n_angle_evals : stats_[d].n_mag_evals) += ne;
          stats_[d].n_reductions += ne;
          done[d] = true;
        } else {
          more.push_back({d, need});
        }
      }
      if (more.empty()) break;
      prefetch_multi(more, fo1, fo2);
    }
  };
  run_walks(true);   // angle walk at the start magnitude
  run_walks(false);  // magnitude walk at the fitted angle
  // The last getmodamp at the fitted (angle, mag) is queued behind the walks
  // and read, with the other orders' modamps at the fitted k0, after one wait.
  std::vector<Cand> fin;
  std::vector<int> fin_dir;
  std::vector<Eval> final_ev(L.ndirs);  // the next overlaps clear the caches
  for (int d : dirs) {
    k0_[d] = make_float2(mag[d] * dtrig::fcos(angle[d]), mag[d] * dtrig::fsin(angle[d]));
    auto it = fw_[d].cache.find(key_of(angle[d], mag[d]));
    if (it != fw_[d].cache.end()) {
      final_ev[d] = it->second;
    } else {
      fin.push_back({d, kvec_of(angle[d], mag[d], fo1, fo2)});
      fin_dir.push_back(d);
    }
  }
  if (!fin.empty()) eval_enqueue(fin.data(), (int)fin.size(), 0);
  std::vector<std::pair<int, int>> others;  // (order1, order2) at the fitted k0
  if (L.nz == 1) {
    for (int order = 2; order < L.norders; ++order) others.push_back({order - 1, order});
  } else {
    for (int order = 1; order < L.norders; ++order)
      if (order != fo2) others.push_back({0, order});
  }
  std::vector<Eval> fin_ev;
  // the first other order's overlaps reuse the fit's scratch after the queued
  // evaluation has read it (stream order)
  if (others.empty())
    realspacemodamp_multi({}, 0, 0, (int)fin.size(), &fin_ev);
  else
    realspacemodamp_multi(dirs, others[0].first, others[0].second, (int)fin.size(), &fin_ev);
  for (size_t i = 0; i < fin.size(); ++i) final_ev[fin_dir[i]] = fin_ev[i];
  for (int d : dirs) {
    float2 modamp;
    double a2;
    float cc = 0;
    modamp_of(final_ev[d], &modamp, &a2, &cc);
    stats_[d].n_reductions++;
    corr_[d][fo2] = cc;
    amp_[d][fo2] = modamp;
  }
  for (size_t k = 1; k < others.size(); ++k) realspacemodamp_multi(dirs, others[k].first, others[k].second);
}

// findk0_all + fitk0_all + the other orders' modamps with the decisions on
// the device: the cross-correlation peaks are refined there, the fit's
// overlaps read their shift from device memory, fit_rounds_ (evaluate, walk)
// rounds run the line searches, and the other orders' overlaps and modamps
// follow at the fitted k0; the host waits once, for the results.  false (the
// host path then runs instead) when the walks needed more rounds than were
// queued.
bool Engine::fit_on_device() {
  const int fo2 = fit_order2();
  const Geo g = geometry();
  FitArgs a{};
  a.ndirs = L.ndirs;
  a.norders = L.norders;
  a.fo2 = fo2;
  a.fp32 = fit_fp32_ ? 1 : 0;
  a.nx = L.nx;
  a.ny = L.ny;
  a.dkx = g.dkx;
  a.dky = g.dky;
  a.deltaangle = 0.001f;
  a.deltamag = (float)(0.1 / (double)((float)std::max(L.nx, L.ny) * P.dxy));
  for (int d = 0; d < L.ndirs; ++d) a.guess[d] = k0_[d];
  // findk0's overlaps at the guesses and their cross-correlation peaks
  std::vector<OvReq> reqs;
  for (int d = 0; d < L.ndirs; ++d) reqs.push_back({d, 0, fo2, k0_[d].x, k0_[d].y});
  make_overlaps(reqs);
  PeakBatch pb{};
  pb.n = L.ndirs;
  for (int k = 0; k < pb.n; ++k) pb.P[k] = fw_[k].P;
  int st = stage_begin("findk0_peak", stream_);
  launch_findpeak(pb, fw_[0].cc, pv_, pi_, fitst_->peak_vals, fitst_->peak_idx, L.nx, L.ny, stream_);
  launch_fit_start(fitst_, a, stream_);
  stage_end(st, stream_);
  // the fit's overlaps at every direction's start vector, then the rounds
  reqs.clear();
  for (int d = 0; d < L.ndirs; ++d) reqs.push_back({d, 0, fo2, 0.f, 0.f, &fitst_->d[d].kov, (float)fo2});
  make_overlaps(reqs);
  const ModampSlots ms = modamp_slots();
  launch_fit_walk(fitst_, a, amps_dev_, corr_dev_, stream_);
  CandSet cs{};
  cs.kc = fitst_->kc;
  cs.ks = fitst_->ks;
  cs.nd = &fitst_->ncand;
  const int np = fit_fp32_ ? modamp_partials_fp32(L.nx, L.ny) : modamp_partials_fp64(L.nx, L.ny);
  const int rounds = fit_rounds_;
  for (int r = 0; r < rounds; ++r) {
    st = stage_begin("modamp", stream_);
    if (fit_fp32_)
      launch_modamp_fp32(ms, cs, kFitReq * L.ndirs, L.nx, L.ny, P.dxy, part32_, stream_);
    else
      launch_modamp_fp64(ms, cs, kFitReq * L.ndirs, L.nx, L.ny, P.dxy, part64_, stream_);
    launch_modamp_evals(fit_fp32_, ms, cs, kFitReq * L.ndirs, part64_, part32_, np, zsum_blocks(L.nxy),
                        &fitst_->ev[0][0], false, stream_);
    launch_fit_walk(fitst_, a, amps_dev_, corr_dev_, stream_);
    stage_end(st, stream_);
  }
  // the other orders' modamps at the fitted k0 (as fitk0_all)
  std::vector<std::pair<int, int>> others;
  if (L.nz == 1) {
    for (int order = 2; order < L.norders; ++order) others.push_back({order - 1, order});
  } else {
    for (int order = 1; order < L.norders; ++order)
      if (order != fo2) others.push_back({0, order});
  }
  const int nparts = fit_fp32_ ? modamp_partials_fp32(L.nx, L.ny) : modamp_partials_fp64(L.nx, L.ny);
  for (auto& oo : others) {
    const float mult = (float)(oo.second - oo.first);
    reqs.clear();
    for (int d = 0; d < L.ndirs; ++d) reqs.push_back({d, oo.first, oo.second, 0.f, 0.f, &fitst_->d[d].kov, mult});
    make_overlaps(reqs);
    float2* kc = kc_dev_ + (size_t)oo.second * L.ndirs;
    launch_fit_kc(fitst_, L.ndirs, mult, kc, stream_);
    CandSet c2{};
    c2.kc = kc;
    c2.ks = slot_ids_;
    st = stage_begin("modamp", stream_);
    const ModampSlots m = modamp_slots();
    if (fit_fp32_)
      launch_modamp_fp32(m, c2, L.ndirs, L.nx, L.ny, P.dxy, part32_, stream_);
    else
      launch_modamp_fp64(m, c2, L.ndirs, L.nx, L.ny, P.dxy, part64_, stream_);
    launch_modamp_finish(fit_fp32_, m, slot_ids_, L.ndirs, part64_, part32_, nparts, zsum_blocks(L.nxy), amps_dev_,
                         corr_dev_, L.norders, oo.second, stream_);
    stage_end(st, stream_);
  }
  FitOut* out = (FitOut*)(dzc_ + kZcFit);
  const unsigned seq = ++fit_seq_;
  launch_fit_publish(fitst_, a, amps_dev_, corr_dev_, out, (unsigned*)(dzc_ + kZcFlag), seq, stream_);
  wait_flag(seq);
  const FitOut& o = *(const FitOut*)(hzc_ + kZcFit);
  if (o.overflow || o.ndone < L.ndirs) {
    fit_rounds_ = std::min(kFitMaxRounds, rounds + 2);
    return false;
  }
  // queue what this fit needed, and one round more
  fit_rounds_ = std::min(kFitMaxRounds, std::max(2, o.rounds + 1));
  for (int d = 0; d < L.ndirs; ++d) {
    k0_[d] = o.k0[d];
    require_finite_k0(d, "the k0 search");
    k0_[d] = o.kov[d];
    stats_[d].n_angle_evals = o.n_angle_evals[d];
    stats_[d].n_mag_evals = o.n_mag_evals[d];
    stats_[d].n_reductions = o.n_reductions[d] + (int)others.size();
    for (int order = 1; order < L.norders; ++order) {
      amp_[d][order] = o.amps[d * L.norders + order];
      corr_[d][order] = o.corr[d * L.norders + order];
    }
  }
  return true;
}

// A non-finite k0 (the reference carries it on into NaN output) only comes
// from non-finite or empty input: a NaN/Inf pixel spreads through the
// transforms to every overlap product, and a stack with no modulated signal
// makes the modamp ratios 0/0.  It is reported here, before any kernel takes
// it as a frequency shift.
void Engine::require_finite_k0(int dir, const char* where) const {
  if (std::isfinite(k0_[dir].x) && std::isfinite(k0_[dir].y)) return;
  throw ArgError(std::string(where) + " of direction " + std::to_string(dir) +
                 " gave a non-finite modulation vector: the stack holds NaN/Inf pixels or no modulated signal");
}

static float2 cmulh(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
static float cmagh(float2 a) { return sqrtf(a.x * a.x + a.y * a.y); }

// The k0 policy of findModulationVectorsAndPhasesForAllDirections
// (cudaSirecon.cpp:296-492).  The reference fits the directions one after
// another; they are independent, so here they advance together and share
// every launch and read-back.
void Engine::find_modulation() {
  if (!preprocessed_) throw StateError("find_modulation before load_and_rescale");
  refresh_fit();
  const bool reuse = P.ntimes > 1 && cur_t_ > 0 && P.bUseTime0k0;
  std::vector<int> all(L.ndirs);
  for (int d = 0; d < L.ndirs; ++d) {
    all[d] = d;
    stats_[d] = simr_fit{};
    if (!reuse) k0_[d] = k0guess_[d];
  }
  if (P.bSearchforvector && !reuse) {
    // the device-driven fit covers every case but a later time point's search
    // (whose fallback to time 0's k0 needs the host) and Bessel SIM (whose
    // overlap z support follows |k0|)
    const bool on_device = device_fit_ && !P.bBessel && !(P.ntimes > 1 && cur_t_ > 0) && fit_on_device();
    if (!on_device) {
      for (int d : all) {
        stats_[d] = simr_fit{};
        k0_[d] = k0guess_[d];
      }
      findk0_all(all);
      for (int d : all) require_finite_k0(d, "the k0 search");
      fitk0_all(all);
    }
    std::vector<int> fallback;
    for (int d : all) {
      require_finite_k0(d, "the k0 fit");
      if (cur_t_ == 0) k0_time0_[d] = k0_[d];
      const float dx = k0_[d].x - k0guess_[d].x, dy = k0_[d].y - k0guess_[d].y;
      const float dist = sqrtf(dx * dx + dy * dy);
      if (P.ntimes > 1 && cur_t_ > 0 && dist > 2 * K0_WARNING_THRESH) {
        k0_[d] = k0_time0_[d];
        fallback.push_back(d);
      }
    }
    for (int order = 1; order < L.norders; ++order)
      realspacemodamp_multi(fallback, L.nz > 1 ? 0 : order - 1, order);
  } else {
    for (int order = 1; order < L.norders; ++order) realspacemodamp_multi(all, 0, order);
  }
  for (int dir = 0; dir < L.ndirs; ++dir) {
    if (L.nz == 1)
      for (int order = 2; order < L.norders; ++order) amp_[dir][order] = cmulh(amp_[dir][order], amp_[dir][order - 1]);
    if (P.nforceamp > 0 && P.forceamp[0] > 0.0f) {
      for (int order = 1; order < L.norders; ++order) {
        const float a = cmagh(amp_[dir][order]);
        if (order - 1 < P.nforceamp && a < P.forceamp[order - 1]) {
          const float f = P.forceamp[order - 1] / a;
          amp_[dir][order].x *= f;
          amp_[dir][order].y *= f;
        }
      }
    }
    if (!P.bFitallphases) {
      const float base_phase = atan2f(amp_[dir][1].y, amp_[dir][1].x);
      for (int order = 2; order < L.norders; ++order) {
        const float2 amplitude = make_float2(cmagh(amp_[dir][order]), 0);
        const float phi = (float)order * base_phase;
        amp_[dir][order] = cmulh(amplitude, make_float2((float)std::cos((double)phi), (float)std::sin((double)phi)));
      }
    }
  }
  fitted_ = true;
}

// ------------------------------------------------------------ filter
FilterParams Engine::filter_params(int dir, int order) const {
  FilterParams f{};
  const Geo g = geometry();
  f.nx = L.nx;
  f.ny = L.ny;
  f.nz = L.nz;
  f.nxh = L.nxh;
  f.dir = dir;
  f.ndirs = L.ndirs;
  f.order = order;
  f.norders = L.norders;
  f.wiener = P.wiener * P.wiener;
  f.kzscale = g.kzscale;
  const float k0mag = sqrtf(k0_[0].x * k0_[0].x + k0_[0].y * k0_[0].y);
  const float lambdaem = (float)((double)((float)P.wave / P.nimm) / 1000.0);
  const float alpha = asinf(P.na / P.nimm);
  const float lambdaexc = (float)(0.88 * (double)lambdaem);
  float rdistcutoff = (float)((double)(P.na * 2.0f) / ((double)P.wave / 1000.0));
  if ((double)rdistcutoff > 1. / (2. * (double)P.dxy)) rdistcutoff = (float)(1. / (2. * (double)P.dxy));
  if (!P.bBessel) {
    float zq = (1.0f - cosf(alpha)) / lambdaem;
    zq = zq / g.dkz;
    const int zd0 = (int)std::ceil(zq);
    f.zd[0] = zd0;
    f.zd[L.norders - 1] = (int)(1.3 * (double)zd0);
    if (L.norders >= 3)
      for (int o = 1; o < L.norders - 1; o++) f.zd[o] = (int)((1.0f + lambdaem / lambdaexc) * (float)zd0);
  } else {
    // gpuFunctionsImpl.cu:1260-1271: order 0 spans the excitation's kz extent
    // plus the detection OTF's; order o the part of the excitation left at
    // lateral frequency |k0| o / (norders - 1)
    const float kzExMax = (2.0f * P.BesselNA) / P.BesselLambdaEx;
    f.zd[0] = (int)std::rint((kzExMax + (1.0f - cosf(alpha)) / lambdaem) / g.dkz);
    for (int o = 1; o < L.norders; o++) {
      const float c = k0mag * (float)o / (float)(L.norders - 1) / kzExMax;
      SIMR_REQUIRE(c <= 1.0f, "Bessel excitation kz extent 2*besselNA/besselExWave is smaller than the "
                              "pattern's lateral frequency |k0|*order/(norders-1)");
      const double q = ((double)(kzExMax * sinf(acosf(c))) + (1.0 - (double)cosf(alpha)) / (double)lambdaem) /
                       (double)g.dkz;
      f.zd[o] = (int)std::ceil(q);
    }
  }
  for (int o = 0; o < L.norders; o++)
    if (f.zd[o] >= L.nz / 2) f.zd[o] = ((L.nz / 2 - 1) > 0 ? (L.nz / 2 - 1) : 0);
  f.rdistcutoff = rdistcutoff;
  f.apocutoff = rdistcutoff + k0mag * (float)(L.norders - 1);
  f.zapocutoff = (L.norders > 1 && !P.bBessel) ? (float)f.zd[1] : (float)f.zd[0];
  f.bessel = P.bBessel ? 1 : 0;
  f.dkx = (float)(1. / (double)((float)L.nx * P.dxy));
  f.dky = (float)(1. / (double)((float)L.ny * P.dxy));
  f.min_dkr = f.dkx < f.dky ? f.dkx : f.dky;
  f.suppRadius = (float)P.suppression_radius * f.min_dkr;
  f.bSuppress = P.bSuppress_singularities;
  f.bDampenOrder0 = P.bDampenOrder0;
  f.bNoKz0 = P.bNoKz0;
  f.bFilteroverlaps = P.bFilteroverlaps;
  f.apodizeoutput = P.apodizeoutput;
  for (int d = 0; d < L.ndirs; ++d) {
    f.k0[d] = k0_[d];
    for (int o = 0; o < L.norders; ++o) {
      f.ampmag2_all[d * L.norders + o] = amp_[d][o].x * amp_[d][o].x + amp_[d][o].y * amp_[d][o].y;
      f.noiseVar[d * L.norders + o] = 1.0f;  // allocSepMatrixAndNoiseVarFactors (cudaSirecon.cpp:728)
    }
  }
  f.conjamp = make_float2(amp_[dir][order].x, -amp_[dir][order].y);
  for (int o = 0; o < L.norders; ++o) f.otf[o] = otf_table(dir, o);
  f.unit_nv = 1;
  for (int d = 0; d < L.ndirs; ++d) {
    for (int o = 0; o < L.norders; ++o) {
      const float nv = f.noiseVar[d * L.norders + o];
      f.inv_nv[d * L.norders + o] = (float)(1. / (double)nv);
      if (nv != 1.0f) f.unit_nv = 0;
    }
  }
  f.cut2 = (float)((double)rdistcutoff * (double)rdistcutoff);
  f.inv_dkr2 = (float)(1.0 / ((double)f.otf[0].dkrotf * (double)f.otf[0].dkrotf));
  f.inv_rcut2 = (float)(1.0 / ((double)rdistcutoff * (double)rdistcutoff));
  f.supp2 = (float)((double)f.suppRadius * (double)f.suppRadius);
  f.inv_min_dkr2 = (float)(1.0 / ((double)f.min_dkr * (double)f.min_dkr));
  f.inv_zd0 = 1.0f / (float)f.zd[0];  // inf for zd[0] == 0: 0 * inf is the reference's 0 / 0
  f.inv_rcut = 1.0f / f.rdistcutoff;
  f.inv_apo = 1.0f / f.apocutoff;
  f.inv_apo2 = (float)(1.0 / ((double)f.apocutoff * (double)f.apocutoff));
  f.inv_zapo = f.zapocutoff > 0 ? 1.0f / f.zapocutoff : 0.f;
  f.inv_nxh = 1.0f / (float)L.nxh;
  SIMR_REQUIRE((long long)L.nxh * L.ny < (1LL << 24), "filter: plane too large");
  float2* db = bands_ + (size_t)dir * L.nbands * L.H;
  f.bp = order == 0 ? db : db + (size_t)(2 * order - 1) * L.H;
  f.bp2 = order == 0 ? nullptr : db + (size_t)(2 * order) * L.H;
  return f;
}

// Every order of directions d0 .. d1-1 in one launch; amps: the modamps on
// the device (graphed time points) or null (the host's).
void Engine::filter_dirs(int d0, int d1, const float2* amps) {
  if (amps == nullptr) refresh_fit();  // (a graph capture passes the device modamps)
  FilterParams f = filter_params(d0, 0);
  f.amps = amps;
  const float2* otf0 = f.otf[0].t;
  FilterBatch b{};
  for (int d = d0; d < d1; ++d)
    for (int o = 0; o < L.norders; ++o) {
      const FilterParams g = filter_params(d, o);
      b.b[b.n++] = FilterBand{d, o, g.conjamp, g.bp, g.bp2, (long long)(g.otf[0].t - otf0)};
    }
  const int st = stage_begin("filter", stream_);
  launch_filter_batch(f, b, stream_);
  stage_end(st, stream_);
}

void Engine::filter_dir(int dir) { filter_dirs(dir, dir + 1, nullptr); }

void Engine::assemble() {
  AsmParams a{};
  a.nx = L.nx;
  a.ny = L.ny;
  a.nz = L.nz;
  a.nxh = L.nxh;
  a.xdim = L.xdim;
  a.ydim = L.ydim;
  a.zdim = L.zdim;
  a.nasm = L.nasm;
  a.ndirs = L.ndirs;
  a.norders = L.norders;
  a.bands = bands_;
  a.band_stride = L.H;
  a.nbands = L.nbands;
  a.stageZ = stageZ_;
  a.stageY = stageY_;
  a.out = out_;
  a.fact = (float)((double)P.explodefact / 0.5);
  a.dxyo = P.dxy / P.zoomfact;
  for (int d = 0; d < L.ndirs; ++d) a.k0[d] = k0_[d];
  {
    // the filter zeroes every band element beyond rdistcutoff (filter_params)
    const Geo g = geometry();
    a.dkx = g.dkx;
    a.dky = g.dky;
    a.rc2_zero = (float)((double)g.rdistcutoff * (double)g.rdistcutoff * 1.0002);
  }
  cur_stage_ = stage_begin("asm_z", stream_);
  launch_asm_z(a, stream_);
  stage_end(cur_stage_, stream_);
  cur_stage_ = stage_begin("asm_y", stream_);
  launch_asm_y(a, stream_);
  stage_end(cur_stage_, stream_);
  cur_stage_ = stage_begin("asm_x", stream_);
  launch_asm_x(a, stream_);
  stage_end(cur_stage_, stream_);
}

void Engine::filter_assemble() {
  if (!fitted_) throw StateError("filter_assemble before find_modulation");
  filter_dirs(0, L.ndirs, nullptr);
  assemble();
  processed_ = true;
}

void Engine::process_volume() {
  if (graph_eligible()) {
    run_graphed();
    return;
  }
  find_modulation();
  filter_assemble();
}

// ------------------------------------------------------------ graphed time points
// A time point that takes time 0's k0 (cudaSirecon.cpp:296-331) needs no
// host decision: its modamps are findrealspacemodamp at the known k0 for
// every order (realspacemodamp with order1 = 0, then the policy of 450-491),
// so the fit, the filter and the assembly go on the device as one graph:
// per order, every direction's overlaps in one launch per pass and their
// modamp reductions in one launch, the modamp result arithmetic and the
// policy in small kernels, and the filter reading the modamps from device
// memory.  The
// graph is captured once per context and replayed while k0 and the
// parameters stay the same.
bool Engine::graph_eligible() const {
  return P.use_graphs != 0 && !timing_ && preprocessed_ && P.ntimes > 1 && cur_t_ > 0 && P.bUseTime0k0 != 0;
}

void Engine::build_graph() {
  if (graph_exec_) SIMR_HIP(hipGraphExecDestroy(graph_exec_));
  if (graph_) SIMR_HIP(hipGraphDestroy(graph_));
  graph_exec_ = nullptr;
  graph_ = nullptr;
  // candidate vectors of findrealspacemodamp (gpuFunctionsImpl.cu:1093-1094),
  // order1 = 0: [order][dir] = k0[dir] * order
  std::vector<float2> kc((size_t)L.norders * L.ndirs, make_float2(0.f, 0.f));
  for (int o = 1; o < L.norders; ++o)
    for (int d = 0; d < L.ndirs; ++d)
      kc[(size_t)o * L.ndirs + d] = make_float2(k0_[d].x * (float)o, k0_[d].y * (float)o);
  SIMR_HIP(hipMemcpyAsync(kc_dev_, kc.data(), kc.size() * 8, hipMemcpyHostToDevice, stream_));
  sync();
  SIMR_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  try {
    const int nparts = fit_fp32_ ? modamp_partials_fp32(L.nx, L.ny) : modamp_partials_fp64(L.nx, L.ny);
    const ModampSlots m = modamp_slots();
    for (int o = 1; o < L.norders; ++o) {
      std::vector<OvReq> reqs;
      for (int d = 0; d < L.ndirs; ++d) reqs.push_back({d, 0, o, k0_[d].x, k0_[d].y});
      make_overlaps(reqs);
      CandSet cs{};
      cs.kc = kc_dev_ + (size_t)o * L.ndirs;
      cs.ks = slot_ids_;
      if (fit_fp32_)
        launch_modamp_fp32(m, cs, L.ndirs, L.nx, L.ny, P.dxy, part32_, stream_);
      else
        launch_modamp_fp64(m, cs, L.ndirs, L.nx, L.ny, P.dxy, part64_, stream_);
      launch_modamp_finish(fit_fp32_, m, slot_ids_, L.ndirs, part64_, part32_, nparts, zsum_blocks(L.nxy),
                           amps_dev_, corr_dev_, L.norders, o, stream_);
    }
    AmpPolicy a{};
    a.ndirs = L.ndirs;
    a.norders = L.norders;
    a.chain = L.nz == 1;
    a.nforceamp = P.nforceamp;
    for (int o = 0; o < SIMR_MAX_ORDERS; ++o) a.forceamp[o] = P.forceamp[o];
    a.fit_all_phases = P.bFitallphases;
    launch_amp_policy(a, amps_dev_, stream_);
    filter_dirs(0, L.ndirs, amps_dev_);
    assemble();
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(stream_, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  SIMR_HIP(hipStreamEndCapture(stream_, &graph_));
  SIMR_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
  graph_k0_ = k0_;
  graph_version_ = param_version_;
}

void Engine::run_graphed() {
  refresh_fit();
  bool same_k0 = graph_k0_.size() == k0_.size();
  for (size_t d = 0; same_k0 && d < k0_.size(); ++d)
    same_k0 = std::memcmp(&graph_k0_[d], &k0_[d], sizeof(float2)) == 0;
  if (!graph_exec_ || graph_version_ != param_version_ || !same_k0) build_graph();
  SIMR_HIP(hipGraphLaunch(graph_exec_, stream_));
  const size_t na = (size_t)L.ndirs * L.norders;
  SIMR_HIP(hipMemcpyAsync(fit_host_, amps_dev_, na * 8, hipMemcpyDeviceToHost, stream_));
  SIMR_HIP(hipMemcpyAsync((char*)fit_host_ + na * 8, corr_dev_, na * 4, hipMemcpyDeviceToHost, stream_));
  fit_pending_ = true;
  for (int d = 0; d < L.ndirs; ++d) {
    stats_[d] = simr_fit{};
    stats_[d].n_reductions = L.norders - 1;
  }
  fitted_ = processed_ = true;
}

// Moves the modamps of the last graphed time point into the host state.
void Engine::refresh_fit() {
  if (!fit_pending_) return;
  SIMR_HIP(hipStreamSynchronize(stream_));
  const size_t na = (size_t)L.ndirs * L.norders;
  const float2* a = (const float2*)fit_host_;
  const float* c = (const float*)((const char*)fit_host_ + na * 8);
  for (int d = 0; d < L.ndirs; ++d)
    for (int o = 1; o < L.norders; ++o) {
      amp_[d][o] = a[(size_t)d * L.norders + o];
      corr_[d][o] = c[(size_t)d * L.norders + o];
    }
  fit_pending_ = false;
}

void Engine::get_result(float* host) {
  if (!processed_) throw StateError("no processed volume");
  SIMR_HIP(hipMemcpyAsync(host, out_, result_elems() * 4, hipMemcpyDeviceToHost, stream_));
  sync();
}

simr_fit Engine::fit(int dir) {
  SIMR_REQUIRE(dir >= 0 && dir < L.ndirs, "direction out of range");
  refresh_fit();
  simr_fit f = stats_[dir];
  f.k0x = k0_[dir].x;
  f.k0y = k0_[dir].y;
  f.k0guess_x = k0guess_[dir].x;
  f.k0guess_y = k0guess_[dir].y;
  for (int o = 0; o < SIMR_MAX_ORDERS; ++o) {
    f.amp_re[o] = o < L.norders ? amp_[dir][o].x : 0.f;
    f.amp_im[o] = o < L.norders ? amp_[dir][o].y : 0.f;
    f.corr[o] = corr_[dir][o];
  }
  return f;
}

void Engine::set_fit(int dir, const simr_fit& f) {
  SIMR_REQUIRE(dir >= 0 && dir < L.ndirs, "direction out of range");
  SIMR_REQUIRE(std::isfinite(f.k0x) && std::isfinite(f.k0y), "k0 must be finite");
  refresh_fit();
  k0_[dir] = make_float2(f.k0x, f.k0y);
  for (int o = 0; o < L.norders; ++o) amp_[dir][o] = make_float2(f.amp_re[o], f.amp_im[o]);
  fitted_ = true;
}

// Time-0 state: [ndirs] float2 k0 fitted at t = 0, then [nz] double
// bleach-correction reference sums of t = 0.
size_t Engine::time0_state_size() const { return (size_t)L.ndirs * sizeof(float2) + (size_t)L.nz * sizeof(double); }

void Engine::get_time0_state(void* host, size_t bytes) {
  SIMR_REQUIRE(host && bytes == time0_state_size(), "time-0 state buffer has the wrong size");
  std::memcpy(host, k0_time0_.data(), (size_t)L.ndirs * sizeof(float2));
  SIMR_HIP(hipMemcpyAsync((char*)host + (size_t)L.ndirs * sizeof(float2), sdp_, (size_t)L.nz * sizeof(double),
                          hipMemcpyDeviceToHost, stream_));
  sync();
}

void Engine::set_time0_state(const void* host, size_t bytes) {
  SIMR_REQUIRE(host && bytes == time0_state_size(), "time-0 state buffer has the wrong size");
  std::memcpy(k0_time0_.data(), host, (size_t)L.ndirs * sizeof(float2));
  k0_ = k0_time0_;  // what a sequential context carries into t > 0
  SIMR_HIP(hipMemcpyAsync(sdp_, (const char*)host + (size_t)L.ndirs * sizeof(float2), (size_t)L.nz * sizeof(double),
                          hipMemcpyHostToDevice, stream_));
  sync();
}

void Engine::get_band(int dir, int band, float* host) {
  SIMR_REQUIRE(dir >= 0 && dir < L.ndirs && band >= 0 && band < L.nbands, "band index out of range");
  SIMR_HIP(hipMemcpyAsync(host, bands_ + ((size_t)dir * L.nbands + band) * L.H, L.H * 8, hipMemcpyDeviceToHost,
                          stream_));
  sync();
}

void Engine::set_band(int dir, int band, const float* host) {
  SIMR_REQUIRE(dir >= 0 && dir < L.ndirs && band >= 0 && band < L.nbands, "band index out of range");
  SIMR_HIP(hipMemcpyAsync(bands_ + ((size_t)dir * L.nbands + band) * L.H, host, L.H * 8, hipMemcpyHostToDevice,
                          stream_));
  sync();
  preprocessed_ = true;
}

void Engine::debug_overlaps(int dir, int o1, int o2, float k0x, float k0y, double* Ph, double* sx, double* sy) {
  make_overlaps({{dir, o1, o2, k0x, k0y}}, true, &dbg_);
  SIMR_HIP(hipMemcpyAsync(Ph, dbg_.P, L.nxy * 16, hipMemcpyDeviceToHost, stream_));
  SIMR_HIP(hipMemcpyAsync(sx, dbg_.SX, L.nxy * 8, hipMemcpyDeviceToHost, stream_));
  SIMR_HIP(hipMemcpyAsync(sy, dbg_.SY, L.nxy * 8, hipMemcpyDeviceToHost, stream_));
  sync();
}

// The overlap pair in real space, [nz][ny][nx] complex each (the reference's
// overlap0/overlap1 after makeoverlaps' inverse 3D FFT).  Built in the dumps'
// own scratch, then transformed into the assembly staging region, which no
// fit or assembly uses between find_modulation and filter_assemble.
void Engine::overlap_volumes(int dir, int o1, int o2, float k0x, float k0y, float* ov0, float* ov1) {
  SIMR_REQUIRE(dir >= 0 && dir < L.ndirs, "direction out of range");
  SIMR_REQUIRE(o1 >= 0 && o2 > o1 && o2 < L.norders, "overlap orders out of range");
  const size_t vbytes = (size_t)L.nz * L.nxy * 8;
  SIMR_REQUIRE(2 * vbytes <= stage_bytes_, "overlap volumes do not fit the staging region");
  sync();
  make_overlaps({{dir, o1, o2, k0x, k0y}}, false, &dbg_);
  launch_overlap_volumes(dbg_.ovx, dbg_.zd, nzs_stride_, L.nz, L.nxy, stageZ_, stream_);
  SIMR_HIP(hipMemcpyAsync(ov0, stageZ_, vbytes, hipMemcpyDeviceToHost, stream_));
  SIMR_HIP(hipMemcpyAsync(ov1, (const char*)stageZ_ + vbytes, vbytes, hipMemcpyDeviceToHost, stream_));
  sync();
}

}  // namespace simr
