// Host orchestration of one reconstruction context.  See engine.hpp.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <thread>

#include "engine.hpp"
#include "fft_api.hpp"
#include "sizes.hpp"

namespace simr {

namespace {
constexpr double kPI = 3.14159265358979323846;
constexpr float K0_WARNING_THRESH = 0.05f;  // cudaSireconImpl.h:54

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr int kMaxCand = 64;  // line-search candidates reduced per launch

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
// Bytes of one direction's fit scratch, in the order Engine::carve_fit lays it out.
static size_t fit_set_bytes(const Layout& L, int ncand) {
  size_t b = 0;
  b += align256((size_t)2 * (2 * L.zd_ov_max + 1) * L.nxy * 8);  // overlap x/y stage
  b += align256((size_t)L.nxy * 16) + 2 * align256((size_t)L.nxy * 8) + align256((size_t)L.nxy * 8);  // P, SX, SY, cc
  b += 4 * align256(4096) + align256((size_t)ncand * 8);  // peak search, candidates
  b += align256((size_t)ncand * modamp_partials_fp64(L.nx, L.ny) * 16);
  b += align256((size_t)ncand * modamp_partials_fp32(L.nx, L.ny) * 16);
  b += align256((size_t)2 * kZsumMaxBlocks * 8);
  return b;
}

Layout Layout::of(const simr_params& p) {
  Layout L{};
  SIMR_REQUIRE(p.nx > 0 && p.ny > 0 && p.nz > 0, "image dimensions must be positive");
  SIMR_REQUIRE(p.ndirs >= 1 && p.ndirs <= SIMR_MAX_DIRS, "ndirs must be in 1..8");
  L.norders = norders_of(p);
  SIMR_REQUIRE(L.norders >= 1 && L.norders <= SIMR_MAX_ORDERS, "norders must be in 1..3");
  SIMR_REQUIRE(2 * L.norders - 1 == p.nphases, "nphases must equal 2*norders-1");
  SIMR_REQUIRE(p.bRadAvgOTF, "only rotationally averaged OTFs are supported (otfRA)");
  SIMR_REQUIRE(p.nxotf >= 2 && p.nzotf >= 1 && p.nyotf == 1, "invalid radially averaged OTF geometry");
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
  const int ncand = kMaxCand;
  size_t b = 0;
  b += align256((size_t)L.nimg * L.nxy_raw * 4);        // raw (float or uint16)
  b += 2 * align256((size_t)L.nxy_raw * 4);              // bg, slope
  b += align256((size_t)L.nimg * kSumSplit * 8) + align256((size_t)L.nimg * 4) + align256((size_t)p.nz * 8) + align256(256);
  b += align256((size_t)L.ndirs * L.nbands * L.H * 8);   // bands
  b += align256((size_t)L.nDirsOTF * L.norders * L.otfsize * 8);
  b += 3 * align256((size_t)L.ndirs * SIMR_MAX_ORDERS * 8);  // device modamps of graphed time points
  const size_t fit = fit_set_bytes(L, ncand);  // direction 0's fit scratch
  b += fit;
  // assembly staging; the fit scratch of directions 1.. aliases it
  const size_t stage = align256((size_t)L.nasm * L.zdim * L.nxy * 8) +            // stageZ
                       align256((size_t)L.nasm * L.zdim * (size_t)L.ydim * p.nx * 8);  // stageY
  // (and the resampled float planes, which live there while the bands are formed)
  b += std::max(std::max(stage, (size_t)(L.ndirs - 1) * fit), L.resample ? (size_t)L.nimg * (size_t)L.nxy * 4 : (size_t)0);
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
  const size_t fitb = fit_set_bytes(L, kMaxCand);
  char* fit0 = (char*)arena_.take(fitb);
  const size_t zb = align256((size_t)L.nasm * L.zdim * L.nxy * 8);
  const size_t yb = align256((size_t)L.nasm * L.zdim * (size_t)L.ydim * L.nx * 8);
  stage_bytes_ = std::max(std::max(zb + yb, (size_t)(L.ndirs - 1) * fitb),
                          L.resample ? (size_t)L.nimg * (size_t)L.nxy * 4 : (size_t)0);
  char* stage = (char*)arena_.take(stage_bytes_);
  stageZ_ = (float2*)stage;
  resampled_ = (float*)stage;
  stageY_ = (float2*)(stage + zb);
  out_ = (float*)arena_.take((size_t)L.zdim * L.ydim * (size_t)L.xdim * 4);

  hstage_bytes_ = std::max<size_t>((size_t)kMaxCand * modamp_partials_fp32(L.nx, L.ny) * 16,
                                   (size_t)kMaxCand * modamp_partials_fp64(L.nx, L.ny) * 16) +
                  (size_t)2 * kZsumMaxBlocks * 8 + 8192;
  fw_.resize(L.ndirs);
  for (int d = 0; d < L.ndirs; ++d) {
    FitWork& w = fw_[d];
    char* at = d == 0 ? fit0 : stage + (size_t)(d - 1) * fitb;
    size_t unused = 0;
    carve_fit(w, at, false, unused);
    if (d == 0) {
      w.stream = stream_;
    } else {
      SIMR_HIP(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
      w.own_stream = true;
    }
    SIMR_HIP(hipHostMalloc(&w.hstage, hstage_bytes_, hipHostMallocDefault));
    SIMR_HIP(hipEventCreateWithFlags(&w.done, hipEventDisableTiming));
  }

  SIMR_HIP(hipHostMalloc(&fit_host_, (size_t)L.ndirs * L.norders * 12, hipHostMallocDefault));
  SIMR_HIP(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
  join_ev_.assign(L.ndirs, nullptr);
  for (auto& e : join_ev_) SIMR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
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
  if (fork_ev_) (void)hipEventDestroy(fork_ev_);
  for (hipEvent_t e : join_ev_)
    if (e) (void)hipEventDestroy(e);
  if (fit_host_) (void)hipHostFree(fit_host_);
  for (auto& e : events_) {
    (void)hipEventDestroy(e.second.first);
    (void)hipEventDestroy(e.second.second);
  }
  for (hipEvent_t e : ev_pool_) (void)hipEventDestroy(e);
  for (auto& w : fw_) {
    if (w.own_stream) {
      (void)hipStreamSynchronize(w.stream);
      (void)hipStreamDestroy(w.stream);
    }
    if (w.done) (void)hipEventDestroy(w.done);
    if (w.hstage) (void)hipHostFree(w.hstage);
  }
  arena_.release();
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::sync() { SIMR_HIP(hipStreamSynchronize(stream_)); }

// Waits for the work queued on a fit stream.  The fit's read-backs are small
// and on its critical path, so the event is polled: the wait ends within
// microseconds of the copy, where a blocking synchronize adds tens of them.
// After 2 ms of polling (a busy device, e.g. other contexts' work) it falls
// back to the blocking wait.
void Engine::wait_fit(FitWork& w) {
  SIMR_HIP(hipEventRecord(w.done, w.stream));
  const auto t0 = std::chrono::steady_clock::now();
  for (int i = 0;; ++i) {
    const hipError_t e = hipEventQuery(w.done);
    if (e == hipSuccess) return;
    if (e != hipErrorNotReady) SIMR_HIP(e);
    if ((i & 15) == 15 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) break;
    for (int k = 0; k < 32; ++k) __builtin_ia32_pause();  // a few microseconds between queries
  }
  SIMR_HIP(hipEventSynchronize(w.done));
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

// Stage timing: a pair of events around the stage's launches on its stream
// (the fit's directions record on their own streams, from their threads).
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

// Fit scratch of one direction at `at` (the layout fit_set_bytes counts).
void Engine::carve_fit(FitWork& w, char*& at, bool count_only, size_t& bytes) const {
  auto take = [&](size_t n) {
    char* p = at;
    at += align256(n);
    bytes += align256(n);
    return count_only ? nullptr : p;
  };
  w.ovx = (float2*)take((size_t)2 * (2 * L.zd_ov_max + 1) * L.nxy * 8);
  w.P = (double2*)take((size_t)L.nxy * 16);
  w.SX = (double*)take((size_t)L.nxy * 8);
  w.SY = (double*)take((size_t)L.nxy * 8);
  w.cc = (float2*)take((size_t)L.nxy * 8);
  w.pv = (float*)take(4096);
  w.pi = (int*)take(4096);
  w.peakvals = (float*)take(4096);
  w.peakidx = (int*)take(4096);
  w.kcand = (float2*)take((size_t)kMaxCand * 8);
  w.part64 = (double2*)take((size_t)kMaxCand * modamp_partials_fp64(L.nx, L.ny) * 16);
  w.part32 = (float4*)take((size_t)kMaxCand * modamp_partials_fp32(L.nx, L.ny) * 16);
  w.sumparts = (double*)take((size_t)2 * kZsumMaxBlocks * 8);
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

OtfRA Engine::otf_table(int dir, int order) const {
  const int d = P.bOneOTFperAngle ? dir : 0;
  OtfRA o;
  o.t = otf_ + ((size_t)d * L.norders + order) * L.otfsize;
  o.nzotf = P.nzotf;
  o.nxotf = P.nxotf;
  o.dkrotf = P.dkrotf;
  return o;
}

// ------------------------------------------------------------ overlaps / fit
void Engine::make_overlaps(FitWork& w, int dir, int o1, int o2, float k0x, float k0y, bool per_elem) {
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
  w.zd = p.zd;
  int st = stage_begin("ov_x", w.stream);
  launch_overlaps_x(p, w.ovx, w.stream);
  stage_end(st, w.stream);
  st = stage_begin("ov_y", w.stream);
  launch_overlaps_y(p, w.ovx, w.stream);
  stage_end(st, w.stream);
  st = stage_begin("ov_zsum", w.stream);
  launch_overlaps_z_reduce(p, w.ovx, w.P, w.SX, w.SY, w.sumparts, fit_fp32_, per_elem, w.stream);
  stage_end(st, w.stream);
  w.ov_totals_valid = false;
  w.cache.clear();
  w.cache_o1 = o1;
  w.cache_o2 = o2;
}

std::vector<Engine::Eval> Engine::eval_batch(FitWork& w, const std::vector<float2>& kvecs) {
  std::vector<Eval> res;
  for (size_t c0 = 0; c0 < kvecs.size(); c0 += kMaxCand) {
    const int nc = (int)std::min<size_t>(kMaxCand, kvecs.size() - c0);
    SIMR_HIP(hipMemcpyAsync(w.kcand, kvecs.data() + c0, nc * 8, hipMemcpyHostToDevice, w.stream));
    const int st = stage_begin("modamp", w.stream);
    if (fit_fp32_) {
      const int nb = modamp_partials_fp32(L.nx, L.ny);
      launch_modamp_fp32(w.P, w.SX, w.SY, w.kcand, nc, L.nx, L.ny, P.dxy, w.part32, w.stream);
      stage_end(st, w.stream);
      SIMR_HIP(hipMemcpyAsync(w.hstage, w.part32, (size_t)nc * nb * 16, hipMemcpyDeviceToHost, w.stream));
      wait_fit(w);
      const float4* h = (const float4*)w.hstage;
      for (int c = 0; c < nc; ++c) {
        float xr = 0, xi = 0, sx = 0, sy = 0;  // cpuReduce order (gpuFunctionsImpl_hh.cu:162-180)
        for (int b = 0; b < nb; ++b) {
          const float4 v = h[(size_t)c * nb + b];
          xr += v.x;
          xi += v.y;
          sx += v.z;
          sy += v.w;
        }
        res.push_back({xr, xi, sx, sy});
      }
    } else {
      const int nb = modamp_partials_fp64(L.nx, L.ny);
      launch_modamp_fp64(w.P, w.kcand, nc, L.nx, L.ny, P.dxy, w.part64, w.stream);
      stage_end(st, w.stream);
      char* h = (char*)w.hstage;
      SIMR_HIP(hipMemcpyAsync(h, w.part64, (size_t)nc * nb * 16, hipMemcpyDeviceToHost, w.stream));
      const size_t soff = (size_t)kMaxCand * nb * 16;
      const int zb = zsum_blocks(L.nxy);
      if (!w.ov_totals_valid)
        SIMR_HIP(hipMemcpyAsync(h + soff, w.sumparts, (size_t)2 * zb * 8, hipMemcpyDeviceToHost, w.stream));
      wait_fit(w);
      if (!w.ov_totals_valid) {
        const double* sp = (const double*)(h + soff);
        double sx = 0, sy = 0;
        for (int b = 0; b < zb; ++b) sx += sp[b];
        for (int b = 0; b < zb; ++b) sy += sp[zb + b];
        w.ov_sx = sx;
        w.ov_sy = sy;
        w.ov_totals_valid = true;
      }
      const double2* hp = (const double2*)h;
      for (int c = 0; c < nc; ++c) {
        double xr = 0, xi = 0;
        for (int b = 0; b < nb; ++b) {
          xr += hp[(size_t)c * nb + b].x;
          xi += hp[(size_t)c * nb + b].y;
        }
        res.push_back({xr, xi, w.ov_sx, w.ov_sy});
      }
    }
  }
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
  const float kx = mag * cosf(angle), ky = mag * sinf(angle);
  return make_float2(kx * (float)(o2 - o1), ky * (float)(o2 - o1));
}

void Engine::prefetch(FitWork& w, int o1, int o2, const std::vector<std::pair<float, float>>& am) {
  std::vector<float2> kv;
  std::vector<std::pair<uint32_t, uint32_t>> keys;
  for (auto& x : am) {
    auto k = key_of(x.first, x.second);
    if (w.cache.count(k) || std::find(keys.begin(), keys.end(), k) != keys.end()) continue;
    keys.push_back(k);
    kv.push_back(kvec_of(x.first, x.second, o1, o2));
  }
  if (kv.empty()) return;
  auto ev = eval_batch(w, kv);
  for (size_t i = 0; i < keys.size(); ++i) w.cache[keys[i]] = ev[i];
}

// getmodamp (gpuFunctionsImpl.cu:1046-1072) on the current overlap set.
double Engine::getmodamp(FitWork& w, int dir, float angle, float mag, int o1, int o2, float2* modamp,
                         float* corr) {
  auto k = key_of(angle, mag);
  if (!w.cache.count(k)) prefetch(w, o1, o2, {{angle, mag}});
  double a2;
  modamp_of(w.cache[k], modamp, &a2, corr);
  stats_[dir].n_reductions++;
  return a2;
}

double Engine::realspacemodamp(FitWork& w, int dir, int o1, int o2, float2 kv, bool redo, float2* modamp,
                               float* corr) {
  if (redo) make_overlaps(w, dir, o1, o2, kv.x, kv.y);
  const float2 k = make_float2(kv.x * (float)(o2 - o1), kv.y * (float)(o2 - o1));
  auto ev = eval_batch(w, {k});
  stats_[dir].n_reductions++;
  double a2;
  modamp_of(ev[0], modamp, &a2, corr);
  return a2;
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

void Engine::findk0(FitWork& w, int dir) {
  const int fo2 = fit_order2();
  make_overlaps(w, dir, 0, fo2, k0_[dir].x, k0_[dir].y);
  const int st = stage_begin("findk0_peak", w.stream);
  launch_findpeak(w.P, w.cc, w.pv, w.pi, w.peakvals, w.peakidx, L.nx, L.ny, w.stream);
  stage_end(st, w.stream);
  char* h = (char*)w.hstage;
  SIMR_HIP(hipMemcpyAsync(h, w.peakvals, 5 * 4, hipMemcpyDeviceToHost, w.stream));
  SIMR_HIP(hipMemcpyAsync(h + 64, w.peakidx, 4, hipMemcpyDeviceToHost, w.stream));
  wait_fit(w);
  const float* v = (const float*)h;
  const int idx = *(const int*)(h + 64);
  const int xc = idx % L.nx, yc = idx / L.nx;
  float2 pk;
  pk.x = fitparabola(v[1], v[0], v[2]) + (float)xc;
  pk.y = fitparabola(v[3], v[0], v[4]) + (float)yc;
  const Geo g = geometry();
  const float2 old = k0_[dir];
  if (old.x / g.dkx < pk.x - (float)(L.nx / 2)) pk.x -= (float)L.nx;
  if (old.x / g.dkx > pk.x + (float)(L.nx / 2)) pk.x += (float)L.nx;
  if (old.y / g.dky < pk.y - (float)(L.ny / 2)) pk.y -= (float)L.ny;
  if (old.y / g.dky > pk.y + (float)(L.ny / 2)) pk.y += (float)L.ny;
  pk.x *= g.dkx / (float)fo2;
  pk.y *= g.dky / (float)fo2;
  k0_[dir] = pk;
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

// fitk0andmodamps (gpuFunctionsImpl.cu:874-1020).  The reference evaluates the
// line-search steps one at a time with a synchronous reduction each; the
// step values it will visit are deterministic (float accumulation of the
// step), so each walk's next kLook steps in both directions are reduced in
// one launch and the reference's step rule is replayed on the results.
void Engine::fitk0andmodamps(FitWork& w, int dir) {
  constexpr int kLook = 6;
  const int fo1 = 0, fo2 = fit_order2();
  const float2 kv = k0_[dir];
  const float k0mag = sqrtf(kv.x * kv.x + kv.y * kv.y);
  const float k0angle = atan2f(kv.y, kv.x);
  auto parab = [&](float x1, double y1, float x2, double y2, float x3, double y3) -> float {
    if (!fit_fp32_) return (float)fitxyparabola<double>(x1, y1, x2, y2, x3, y3);
    return fitxyparabola<float>(x1, (float)y1, x2, (float)y2, x3, (float)y3);
  };
  const float deltaangle = 0.001f;
  const float deltamag = (float)(0.1 / (double)((float)std::max(L.nx, L.ny) * P.dxy));
  auto walk = [&](float start, float mag_or_angle, bool is_angle, float delta, int sign) {
    std::vector<std::pair<float, float>> v;
    float a = start;
    for (int m = 0; m < kLook; ++m) {
      v.push_back(is_angle ? std::make_pair(a, mag_or_angle) : std::make_pair(mag_or_angle, a));
      a = sign > 0 ? a + delta : a - delta;
    }
    return v;
  };
  auto ensure = [&](float angle, float mag, bool is_angle, float delta, int sign) {
    if (w.cache.count(key_of(angle, mag))) return;
    prefetch(w, fo1, fo2, is_angle ? walk(angle, mag, true, delta, sign) : walk(mag, angle, false, delta, sign));
  };

  // first evaluation recomputes the overlaps (redoarrays = recalcarrays >= 1)
  {
    const float2 k1 = make_float2(k0mag * cosf(k0angle), k0mag * sinf(k0angle));
    make_overlaps(w, dir, fo1, fo2, k1.x, k1.y);
  }
  // angle walk: both directions' first kLook steps in one batch
  {
    auto up = walk(k0angle + deltaangle, k0mag, true, deltaangle, +1);
    auto dn = walk(k0angle - deltaangle, k0mag, true, deltaangle, -1);
    std::vector<std::pair<float, float>> all{{k0angle, k0mag}};
    all.insert(all.end(), up.begin(), up.end());
    all.insert(all.end(), dn.begin(), dn.end());
    prefetch(w, fo1, fo2, all);
  }
  float2 modamp;
  float x2 = k0angle;
  double amp2 = getmodamp(w, dir, k0angle, k0mag, fo1, fo2, &modamp, nullptr);
  stats_[dir].n_angle_evals++;
  float angle = k0angle + deltaangle;
  float x3 = angle;
  double amp3 = getmodamp(w, dir, angle, k0mag, fo1, fo2, &modamp, nullptr);
  stats_[dir].n_angle_evals++;
  double amp1 = 0;
  float x1 = 0;
  if (amp3 > amp2) {
    while (amp3 > amp2) {
      amp1 = amp2;
      x1 = x2;
      amp2 = amp3;
      x2 = x3;
      angle += deltaangle;
      x3 = angle;
      ensure(angle, k0mag, true, deltaangle, +1);
      amp3 = getmodamp(w, dir, angle, k0mag, fo1, fo2, &modamp, nullptr);
      stats_[dir].n_angle_evals++;
    }
  } else {
    angle = k0angle;
    double a = amp3;
    amp3 = amp2;
    amp2 = a;
    float ax = x3;
    x3 = x2;
    x2 = ax;
    while (amp3 > amp2) {
      amp1 = amp2;
      x1 = x2;
      amp2 = amp3;
      x2 = x3;
      angle -= deltaangle;
      x3 = angle;
      ensure(angle, k0mag, true, deltaangle, -1);
      amp3 = getmodamp(w, dir, angle, k0mag, fo1, fo2, &modamp, nullptr);
      stats_[dir].n_angle_evals++;
    }
  }
  angle = parab(x1, amp1, x2, amp2, x3, amp3);

  // magnitude walk at the fitted angle
  {
    auto up = walk(k0mag + deltamag, angle, false, deltamag, +1);
    auto dn = walk(k0mag - deltamag, angle, false, deltamag, -1);
    std::vector<std::pair<float, float>> all{{angle, k0mag}};
    all.insert(all.end(), up.begin(), up.end());
    all.insert(all.end(), dn.begin(), dn.end());
    prefetch(w, fo1, fo2, all);
  }
  x2 = k0mag;
  amp2 = getmodamp(w, dir, angle, k0mag, fo1, fo2, &modamp, nullptr);
  stats_[dir].n_mag_evals++;
  float mag = k0mag + deltamag;
  x3 = mag;
  amp3 = getmodamp(w, dir, angle, mag, fo1, fo2, &modamp, nullptr);
  stats_[dir].n_mag_evals++;
  if (amp3 > amp2) {
    while (amp3 > amp2) {
      amp1 = amp2;
      x1 = x2;
      amp2 = amp3;
      x2 = x3;
      mag += deltamag;
      x3 = mag;
      ensure(angle, mag, false, deltamag, +1);
      amp3 = getmodamp(w, dir, angle, mag, fo1, fo2, &modamp, nullptr);
      stats_[dir].n_mag_evals++;
    }
  } else {
    mag = k0mag;
    double a = amp3;
    amp3 = amp2;
    amp2 = a;
    float ax = x3;
    x3 = x2;
    x2 = ax;
    while (amp3 > amp2) {
      amp1 = amp2;
      x1 = x2;
      amp2 = amp3;
      x2 = x3;
      mag -= deltamag;
      x3 = mag;
      ensure(angle, mag, false, deltamag, -1);
      amp3 = getmodamp(w, dir, angle, mag, fo1, fo2, &modamp, nullptr);
      stats_[dir].n_mag_evals++;
    }
  }
  mag = parab(x1, amp1, x2, amp2, x3, amp3);

  float cc = 0;
  getmodamp(w, dir, angle, mag, fo1, fo2, &modamp, &cc);
  corr_[dir][fo2] = cc;
  k0_[dir] = make_float2(mag * cosf(angle), mag * sinf(angle));
  amp_[dir][fo2] = modamp;
  if (L.nz == 1) {
    for (int order = 2; order < L.norders; ++order) {
      realspacemodamp(w, dir, order - 1, order, make_float2(mag * cosf(angle), mag * sinf(angle)), true, &modamp, &cc);
      amp_[dir][order] = modamp;
      corr_[dir][order] = cc;
    }
  } else {
    for (int order = 1; order < L.norders; ++order)
      if (order != fo2) {
        realspacemodamp(w, dir, 0, order, make_float2(mag * cosf(angle), mag * sinf(angle)), true, &modamp, &cc);
        amp_[dir][order] = modamp;
        corr_[dir][order] = cc;
      }
  }
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

// k0 policy of findModulationVectorsAndPhasesForAllDirections (cudaSirecon.cpp:296-492)
void Engine::modulation(FitWork& w, int dir) {
  const bool reuse = P.ntimes > 1 && cur_t_ > 0 && P.bUseTime0k0;
  stats_[dir] = simr_fit{};
  if (!reuse) k0_[dir] = k0guess_[dir];
  if (P.bSearchforvector && !reuse) {
    findk0(w, dir);
    require_finite_k0(dir, "the k0 search");
    fitk0andmodamps(w, dir);
    require_finite_k0(dir, "the k0 fit");
    if (cur_t_ == 0) k0_time0_[dir] = k0_[dir];
    const float dx = k0_[dir].x - k0guess_[dir].x, dy = k0_[dir].y - k0guess_[dir].y;
    const float dist = sqrtf(dx * dx + dy * dy);
    if (P.ntimes > 1 && cur_t_ > 0 && dist > 2 * K0_WARNING_THRESH) {
      k0_[dir] = k0_time0_[dir];
      for (int order = 1; order < L.norders; ++order) {
        float2 m;
        float cc;
        const int o1 = L.nz > 1 ? 0 : order - 1;
        realspacemodamp(w, dir, o1, order, k0_[dir], true, &m, &cc);
        amp_[dir][order] = m;
        corr_[dir][order] = cc;
      }
    }
  } else {
    for (int order = 1; order < L.norders; ++order) {
      float2 m;
      float cc;
      realspacemodamp(w, dir, 0, order, k0_[dir], true, &m, &cc);
      amp_[dir][order] = m;
      corr_[dir][order] = cc;
    }
  }
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

// The directions' fits are independent (findModulationVectorsAndPhasesForAllDirections
// loops over them): direction 0 runs on the calling thread and the context
// stream, every other one on a thread and stream of its own after the bands
// are ready, so one direction's kernels fill the GPU while another waits for
// its line-search read-back.  Each thread ends with its stream drained.
void Engine::find_modulation() {
  if (!preprocessed_) throw StateError("find_modulation before load_and_rescale");
  refresh_fit();
  if (L.ndirs > 1) {
    hipEvent_t ready;
    SIMR_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    SIMR_HIP(hipEventRecord(ready, stream_));
    for (int dir = 1; dir < L.ndirs; ++dir) SIMR_HIP(hipStreamWaitEvent(fw_[dir].stream, ready, 0));
    SIMR_HIP(hipEventDestroy(ready));
  }
  std::vector<std::exception_ptr> err(L.ndirs);
  std::vector<std::thread> th;
  std::atomic<int> running(L.ndirs - 1);
  for (int dir = 1; dir < L.ndirs; ++dir)
    th.emplace_back([this, dir, &err, &running] {
      try {
        SIMR_HIP(hipSetDevice(device_));
        modulation(fw_[dir], dir);
        wait_fit(fw_[dir]);
      } catch (...) {
        err[dir] = std::current_exception();
      }
      running.fetch_sub(1, std::memory_order_release);
    });
  try {
    modulation(fw_[0], 0);
  } catch (...) {
    err[0] = std::current_exception();
  }
  // the filter that follows needs every direction's fit: wait for the last
  // one by polling (a blocking join wakes up tens of microseconds late)
  const auto t0 = std::chrono::steady_clock::now();
  while (running.load(std::memory_order_acquire) > 0 &&
         std::chrono::steady_clock::now() - t0 < std::chrono::milliseconds(2))
    for (int k = 0; k < 32; ++k) __builtin_ia32_pause();
  for (auto& t : th) t.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
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

void Engine::filter_dir(int dir) {
  refresh_fit();
  for (int order = 0; order < L.norders; ++order) {
    const int st = stage_begin(order == 0 ? "filter0" : "filter", stream_);
    launch_filter(filter_params(dir, order), stream_);
    stage_end(st, stream_);
  }
}

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
  for (int dir = 0; dir < L.ndirs; ++dir) filter_dir(dir);
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
// every direction's overlaps and modamp reduction on its own stream (fork
// and join events), the modamp result arithmetic and the policy in one-thread
// kernels, and the filter reading the modamps from device memory.  The
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
  // candidate vectors of findrealspacemodamp (gpuFunctionsImpl.cu:1093-1094), order1 = 0
  std::vector<float2> kc((size_t)L.ndirs * L.norders, make_float2(0.f, 0.f));
  for (int d = 0; d < L.ndirs; ++d)
    for (int o = 1; o < L.norders; ++o)
      kc[(size_t)d * L.norders + o] = make_float2(k0_[d].x * (float)o, k0_[d].y * (float)o);
  SIMR_HIP(hipMemcpyAsync(kc_dev_, kc.data(), kc.size() * 8, hipMemcpyHostToDevice, stream_));
  sync();
  SIMR_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  try {
    if (L.ndirs > 1) {
      SIMR_HIP(hipEventRecord(fork_ev_, stream_));
      for (int d = 1; d < L.ndirs; ++d) SIMR_HIP(hipStreamWaitEvent(fw_[d].stream, fork_ev_, 0));
    }
    const int nparts = fit_fp32_ ? modamp_partials_fp32(L.nx, L.ny) : modamp_partials_fp64(L.nx, L.ny);
    for (int d = 0; d < L.ndirs; ++d) {
      FitWork& w = fw_[d];
      for (int o = 1; o < L.norders; ++o) {
        const size_t idx = (size_t)d * L.norders + o;
        make_overlaps(w, d, 0, o, k0_[d].x, k0_[d].y);
        if (fit_fp32_)
          launch_modamp_fp32(w.P, w.SX, w.SY, kc_dev_ + idx, 1, L.nx, L.ny, P.dxy, w.part32, w.stream);
        else
          launch_modamp_fp64(w.P, kc_dev_ + idx, 1, L.nx, L.ny, P.dxy, w.part64, w.stream);
        launch_modamp_finish(fit_fp32_, w.part64, w.part32, nparts, w.sumparts, zsum_blocks(L.nxy), amps_dev_ + idx,
                             corr_dev_ + idx, w.stream);
      }
    }
    if (L.ndirs > 1)
      for (int d = 1; d < L.ndirs; ++d) {
        SIMR_HIP(hipEventRecord(join_ev_[d], fw_[d].stream));
        SIMR_HIP(hipStreamWaitEvent(stream_, join_ev_[d], 0));
      }
    AmpPolicy a{};
    a.ndirs = L.ndirs;
    a.norders = L.norders;
    a.chain = L.nz == 1;
    a.nforceamp = P.nforceamp;
    for (int o = 0; o < SIMR_MAX_ORDERS; ++o) a.forceamp[o] = P.forceamp[o];
    a.fit_all_phases = P.bFitallphases;
    launch_amp_policy(a, amps_dev_, stream_);
    for (int d = 0; d < L.ndirs; ++d)
      for (int o = 0; o < L.norders; ++o) {
        FilterParams f = filter_params(d, o);
        f.amps = amps_dev_;
        launch_filter(f, stream_);
      }
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
  FitWork& w = fw_[0];
  make_overlaps(w, dir, o1, o2, k0x, k0y, true);
  SIMR_HIP(hipMemcpyAsync(Ph, w.P, L.nxy * 16, hipMemcpyDeviceToHost, stream_));
  SIMR_HIP(hipMemcpyAsync(sx, w.SX, L.nxy * 8, hipMemcpyDeviceToHost, stream_));
  SIMR_HIP(hipMemcpyAsync(sy, w.SY, L.nxy * 8, hipMemcpyDeviceToHost, stream_));
  sync();
}

// The overlap pair in real space, [nz][ny][nx] complex each (the reference's
// overlap0/overlap1 after makeoverlaps' inverse 3D FFT).  Built with
// direction 0's fit scratch into the assembly staging region, which no fit
// or assembly uses between find_modulation and filter_assemble.
void Engine::overlap_volumes(int dir, int o1, int o2, float k0x, float k0y, float* ov0, float* ov1) {
  SIMR_REQUIRE(dir >= 0 && dir < L.ndirs, "direction out of range");
  SIMR_REQUIRE(o1 >= 0 && o2 > o1 && o2 < L.norders, "overlap orders out of range");
  const size_t vbytes = (size_t)L.nz * L.nxy * 8;
  const size_t stage_bytes = align256((size_t)L.nasm * L.zdim * L.nxy * 8) +
                             align256((size_t)L.nasm * L.zdim * (size_t)L.ydim * L.nx * 8);
  SIMR_REQUIRE(2 * vbytes <= stage_bytes, "overlap volumes do not fit the staging region");
  sync();  // every direction's fit has finished with the staging region
  FitWork& w = fw_[0];
  make_overlaps(w, dir, o1, o2, k0x, k0y);
  launch_overlap_volumes(w.ovx, w.zd, L.nz, L.nxy, stageZ_, stream_);
  SIMR_HIP(hipMemcpyAsync(ov0, stageZ_, vbytes, hipMemcpyDeviceToHost, stream_));
  SIMR_HIP(hipMemcpyAsync(ov1, (const char*)stageZ_ + vbytes, vbytes, hipMemcpyDeviceToHost, stream_));
  sync();
}

}  // namespace simr
