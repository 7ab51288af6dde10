// extern "C" stage-level ABI (include/simr.h).  Nothing throws across it:
// every entry point maps exceptions to a simr_status and a thread-local message.
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/simr.h"
#include "engine.hpp"
#include "fft_api.hpp"

namespace simr {
void otf_params_default(simr_otf_params* p);
size_t otf_elems(const simr_otf_params& p, int nx, int ny, int nsections);
void make_otf(const simr_otf_params& p, const float* psf, int nx, int ny, int nsec, const float* section_bg,
              float* otf_out, float* band0_out, simr_otf_info* info, bool i2m = false);
}  // namespace simr

struct simr_ctx {
  simr::Engine* eng;
};

namespace {
thread_local std::string g_err;

template <typename F>
int guard(F&& f) {
  try {
    f();
    return SIMR_OK;
  } catch (const simr::ArgError& e) {
    g_err = e.what();
    return SIMR_ERR_ARG;
  } catch (const simr::NoMemError& e) {
    g_err = e.what();
    return SIMR_ERR_NOMEM;
  } catch (const simr::HipError& e) {
    g_err = e.what();
    return SIMR_ERR_HIP;
  } catch (const simr::StateError& e) {
    g_err = e.what();
    return SIMR_ERR_STATE;
  } catch (const simr::IoError& e) {
    g_err = e.what();
    return SIMR_ERR_IO;
  } catch (const simr::FormatError& e) {
    g_err = e.what();
    return SIMR_ERR_FORMAT;
  } catch (const simr::UnsupportedError& e) {
    g_err = e.what();
    return SIMR_ERR_UNSUPPORTED;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return SIMR_ERR_NOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return SIMR_ERR_ARG;
  }
}

simr::Engine& E(simr_ctx* c) {
  if (!c || !c->eng) throw simr::ArgError("null context");
  return *c->eng;
}
}  // namespace

namespace {
void check_fft_dims(const float* data, int nx, int ny, int nz) {
  if (!data) throw simr::ArgError("FFT data pointer is null");
  if (nx < 1 || ny < 1 || nz < 1 || nx > 4096 || ny > 4096 || nz > 4096)
    throw simr::ArgError("FFT lengths must be 1 to 4096");
}

// Device memory that stays allocated until the scope ends (the synchronous
// entry points' staging and direct-DFT scratch).
struct DevMem {
  void* p = nullptr;
  explicit DevMem(size_t bytes) {
    if (bytes) SIMR_HIP(hipMalloc(&p, bytes));
  }
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
};
}  // namespace

// ---- FFT without allocation (simr_fft3d_dev_noalloc)
//
// The device memory simr_fft3d_dev takes from the library (its twiddle tables
// and the direct-DFT scratch volume) comes here from one file-scope device
// array per GPU, so the call makes no allocation at all.  Its layout: three
// twiddle tables of up to 4096 entries (x, y, z), then the scratch volume.
// Every call on a GPU uses that GPU's one array: calls that overlap on one
// GPU (other streams or threads) race.
namespace {
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr size_t kNoAllocTwBytes = (size_t)4096 * sizeof(float2);
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kNoAllocMaxDevices] = {};

// The current device's address of `scratch` (the host-side name is only a
// stand-in), looked up on the first call on each device.
unsigned char* noalloc_scratch() {
  int dev = 0;
  SIMR_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) throw simr::ArgError("device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SIMR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw simr::HipError("hipGetSymbolAddress returned no address for the FFT scratch");
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

void simr_set_last_error(const char* msg) { g_err = msg ? msg : ""; }

const char* simr_last_error(void) { return g_err.c_str(); }

const char* simr_version(void) { return "1.2.0"; }
size_t simr_params_size(void) { return sizeof(simr_params); }
size_t simr_fit_size(void) { return sizeof(simr_fit); }

void simr_params_default(simr_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  // SetDefaultParams (cudaSirecon.cpp:16-72) overlaid with the option defaults (918-1005)
  p->ndirs = 3;
  p->nphases = 5;
  p->norders_output = 0;
  p->dxy = 0.1f;
  p->dz = 0.2f;
  p->wave = 530;
  p->ntimes = 1;
  p->k0startangle = 1.648f;
  p->linespacing = 0.172f;
  p->na = 1.2f;
  p->nimm = 1.33f;
  p->otfcutoff = 0.006f;
  p->zoomfact = 2.0f;
  p->z_zoom = 1;
  p->explodefact = 1.0f;
  p->bFilteroverlaps = 1;
  p->recalcarrays = 1;
  p->napodize = 10;
  p->bSearchforvector = 1;
  p->bUseTime0k0 = 1;
  p->apodizeoutput = 2;  // gammaApo always has a value (cudaSirecon.cpp:1040-1045)
  p->apoGamma = 1.0f;
  p->bSuppress_singularities = 1;
  p->suppression_radius = 10;
  p->bDampenOrder0 = 0;
  p->bFitallphases = 1;
  p->do_rescale = 1;
  p->wiener = 0.01f;
  p->fit_fp64 = 1;
  p->BesselNA = 0.144f;       // --besselNA default (cudaSirecon.cpp:983)
  p->BesselLambdaEx = 0.488f;  // --besselExWave default (cudaSirecon.cpp:981)
  p->use_graphs = 1;
}

size_t simr_otf_size(const simr_params* p) {
  if (!p) return 0;
  return (size_t)p->nzotf * p->nxotf * p->nyotf;
}

size_t simr_otf_total(const simr_params* p) {
  if (!p) return 0;
  const int norders = p->norders_output != 0 ? p->norders_output : p->nphases / 2 + 1;
  return (size_t)(p->bOneOTFperAngle ? p->ndirs : 1) * (size_t)(norders > 0 ? norders : 0) * simr_otf_size(p);
}

size_t simr_footprint(const simr_params* p) {
  size_t n = 0;
  int rc = guard([&] { n = simr::Layout::of(*p).bytes_total; });
  return rc ? 0 : n;
}

int simr_create(const simr_params* p, const float* otf, int device, simr_ctx** out) {
  return guard([&] {
    if (!p || !out) throw simr::ArgError("null argument");
    *out = nullptr;
    simr_ctx* c = new simr_ctx{nullptr};
    try {
      c->eng = new simr::Engine(*p, otf, device);
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

int simr_destroy(simr_ctx* ctx) {
  return guard([&] {
    if (!ctx) return;
    delete ctx->eng;
    delete ctx;
  });
}

int simr_get_params(const simr_ctx* ctx, simr_params* out) {
  return guard([&] { *out = E(const_cast<simr_ctx*>(ctx)).params(); });
}

void simr_otf_params_default(simr_otf_params* p) {
  if (p) simr::otf_params_default(p);
}

size_t simr_otf_elems(const simr_otf_params* p, int nx, int ny, int nsections) {
  return p ? simr::otf_elems(*p, nx, ny, nsections) : 0;
}

int simr_makeotf(const simr_otf_params* p, const float* psf, int nx, int ny, int nsections, const float* section_bg,
                 float* otf, float* band0, simr_otf_info* info) {
  return guard([&] {
    if (!p) throw simr::ArgError("null makeotf parameter block");
    simr::make_otf(*p, psf, nx, ny, nsections, section_bg, otf, band0, info);
  });
}

int simr_makeotf_i2m(const simr_otf_params* p, const float* i2m, int nx, int ny, int nz, float* otf,
                     simr_otf_info* info) {
  return guard([&] {
    if (!p) throw simr::ArgError("null makeotf parameter block");
    simr::make_otf(*p, i2m, nx, ny, nz, nullptr, otf, nullptr, info, true);
  });
}

int simr_update_params(simr_ctx* ctx, const simr_params* p) {
  return guard([&] {
    if (!p) throw simr::ArgError("null parameter block");
    E(ctx).update_params(*p);
  });
}

int simr_set_raw_f32_host(simr_ctx* ctx, const float* raw) { return guard([&] { E(ctx).set_raw(raw, false, false); }); }
int simr_set_raw_u16_host(simr_ctx* ctx, const uint16_t* raw) { return guard([&] { E(ctx).set_raw(raw, true, false); }); }
int simr_set_raw_f32_dev(simr_ctx* ctx, const float* raw) { return guard([&] { E(ctx).set_raw(raw, false, true); }); }
int simr_set_raw_u16_dev(simr_ctx* ctx, const uint16_t* raw) { return guard([&] { E(ctx).set_raw(raw, true, true); }); }
int simr_set_flatfield(simr_ctx* ctx, const float* bg, const float* slope) {
  return guard([&] { E(ctx).set_flatfield(bg, slope); });
}

int simr_load_and_rescale(simr_ctx* ctx, int it) { return guard([&] { E(ctx).load_and_rescale(it); }); }
int simr_set_time_index(simr_ctx* ctx, int it) { return guard([&] { E(ctx).set_time(it); }); }
int simr_find_modulation(simr_ctx* ctx) { return guard([&] { E(ctx).find_modulation(); }); }
int simr_filter_assemble(simr_ctx* ctx) { return guard([&] { E(ctx).filter_assemble(); }); }
int simr_process_volume(simr_ctx* ctx) { return guard([&] { E(ctx).process_volume(); }); }

size_t simr_result_elems(const simr_ctx* ctx) { return ctx && ctx->eng ? ctx->eng->result_elems() : 0; }
void* simr_host_alloc(size_t bytes) {
  void* p = nullptr;
  const int rc = guard([&] {
    if (bytes == 0) throw simr::ArgError("simr_host_alloc: zero bytes");
    SIMR_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
  });
  return rc == SIMR_OK ? p : nullptr;
}

int simr_host_free(void* p) {
  return guard([&] {
    if (p) SIMR_HIP(hipHostFree(p));
  });
}

int simr_get_result_host(simr_ctx* ctx, float* out) { return guard([&] { E(ctx).get_result(out); }); }
int simr_get_result_dev(simr_ctx* ctx, float* out_dev) {
  return guard([&] {
    if (!out_dev) throw simr::ArgError("null output pointer");
    simr::Engine& e = E(ctx);
    SIMR_HIP(hipMemcpyAsync(out_dev, e.result_dev(), e.result_elems() * sizeof(float), hipMemcpyDeviceToDevice,
                            e.stream()));
    e.sync();
  });
}
const float* simr_result_dev(simr_ctx* ctx) { return ctx && ctx->eng ? ctx->eng->result_dev() : nullptr; }
int simr_get_fit(const simr_ctx* ctx, int dir, simr_fit* out) {
  return guard([&] { *out = E(const_cast<simr_ctx*>(ctx)).fit(dir); });
}
int simr_set_fit(simr_ctx* ctx, int dir, const simr_fit* in) { return guard([&] { E(ctx).set_fit(dir, *in); }); }
size_t simr_time0_state_size(const simr_ctx* ctx) {
  return ctx && ctx->eng ? ctx->eng->time0_state_size() : 0;
}
int simr_get_time0_state(simr_ctx* ctx, void* host, size_t bytes) {
  return guard([&] { E(ctx).get_time0_state(host, bytes); });
}
int simr_set_time0_state(simr_ctx* ctx, const void* host, size_t bytes) {
  return guard([&] { E(ctx).set_time0_state(host, bytes); });
}
int simr_get_band_host(simr_ctx* ctx, int dir, int band, float* out) {
  return guard([&] { E(ctx).get_band(dir, band, out); });
}
int simr_set_band_host(simr_ctx* ctx, int dir, int band, const float* in) {
  return guard([&] { E(ctx).set_band(dir, band, in); });
}
int simr_filter_dir(simr_ctx* ctx, int dir) {
  return guard([&] {
    E(ctx).filter_dir(dir);
    E(ctx).sync();
  });
}
int simr_debug_overlaps(simr_ctx* ctx, int dir, int o1, int o2, float k0x, float k0y, double* P, double* sx,
                        double* sy) {
  return guard([&] { E(ctx).debug_overlaps(dir, o1, o2, k0x, k0y, P, sx, sy); });
}
int simr_debug_overlap_volumes(simr_ctx* ctx, int dir, int o1, int o2, float k0x, float k0y, float* ov0,
                               float* ov1) {
  return guard([&] {
    if (!ov0 || !ov1) throw simr::ArgError("overlap volume pointers are null");
    E(ctx).overlap_volumes(dir, o1, o2, k0x, k0y, ov0, ov1);
  });
}

int simr_sync(simr_ctx* ctx) { return guard([&] { E(ctx).sync(); }); }
void* simr_stream(simr_ctx* ctx) { return ctx && ctx->eng ? (void*)ctx->eng->stream() : nullptr; }
int simr_enable_timing(simr_ctx* ctx, int on) { return guard([&] { E(ctx).enable_timing(on != 0); }); }
int simr_time_stage(simr_ctx* ctx, const char* stage) { return guard([&] { E(ctx).time_only(stage); }); }
int simr_stage_timings(simr_ctx* ctx, const char** names, float* ms, int cap) {
  int n = 0;
  int rc = guard([&] { n = E(ctx).timings(names, ms, cap); });
  return rc ? rc : n;
}

int simr_fft_size_supported(int n) { return simr::fft_size_supported(n) ? 1 : 0; }


int simr_fft3d_dev(float* data, int nx, int ny, int nz, int inverse, void* stream) {
  return guard([&] {
    check_fft_dims(data, nx, ny, nz);
    hipStream_t s = (hipStream_t)stream;
    if (!simr::fft3d_needs_scratch(nx, ny, nz)) {
      simr::fft3d((float2*)data, nx, ny, nz, inverse != 0, s);
      return;
    }
    // a length without a transform kernel: direct-DFT passes through a
    // scratch volume allocated for the call; the call then waits for them
    DevMem scratch((size_t)nx * ny * nz * sizeof(float2));
    simr::fft3d_any((float2*)data, (float2*)scratch.p, nx, ny, nz, inverse != 0, s);
    SIMR_HIP(hipStreamSynchronize(s));
  });
}


size_t simr_fft3d_noalloc_max_elems(void) { return (kNoAllocScratchBytes - 3 * kNoAllocTwBytes) / sizeof(float2); }

int simr_fft3d_dev_noalloc(float* data, int nx, int ny, int nz, int inverse, void* stream) {
  return guard([&] {
    check_fft_dims(data, nx, ny, nz);
    const bool need = simr::fft3d_needs_scratch(nx, ny, nz);
    const size_t elems = (size_t)nx * ny * nz;
    if (need && elems > simr_fft3d_noalloc_max_elems())
      throw simr::ArgError("simr_fft3d_dev_noalloc: " + std::to_string(elems) +
                           " complex values exceed its scratch of " + std::to_string(simr_fft3d_noalloc_max_elems()) +
                           " (a length without a transform kernel needs a scratch volume of the array's size)");
    hipStream_t s = (hipStream_t)stream;
    unsigned char* base = noalloc_scratch();
    float2* tw[3];
    const int len[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {
      tw[a] = (float2*)(base + a * kNoAllocTwBytes);
      if (len[a] > 1) simr::twiddles_into(tw[a], len[a], s);  // rewritten every call
    }
    const float2* twc[3] = {tw[0], tw[1], tw[2]};
    simr::fft3d_any((float2*)data, need ? (float2*)(base + 3 * kNoAllocTwBytes) : nullptr, nx, ny, nz,
                    inverse != 0, s, twc);
  });
}

int simr_fft3d_host(float* data, int nx, int ny, int nz, int inverse) {
  return guard([&] {
    check_fft_dims(data, nx, ny, nz);
    const size_t bytes = (size_t)nx * ny * nz * sizeof(float2);
    DevMem d(bytes);
    DevMem scratch(simr::fft3d_needs_scratch(nx, ny, nz) ? bytes : 0);
    SIMR_HIP(hipMemcpy(d.p, data, bytes, hipMemcpyHostToDevice));
    simr::fft3d_any((float2*)d.p, (float2*)scratch.p, nx, ny, nz, inverse != 0, nullptr);
    SIMR_HIP(hipMemcpy(data, d.p, bytes, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
