// extern "C" SR_* entry points (include/cudasirecon.h), the drop-in for
// src/cudaSirecon/interface.cpp:10-70.  No exception crosses this boundary:
// each call maps it to a simr_status and a thread-local message.
#include <cstring>
#include <new>
#include <string>

#include "../../include/cudasirecon.h"
#include "errors.hpp"
#include "sr_options.hpp"
#include "sr_reconstructor.hpp"

struct SIM_Reconstructor {
  simr::Reconstructor* r;
};

namespace {

thread_local std::string g_sr_err;

template <class F>
int guard(F&& f) {
  try {
    f();
    g_sr_err.clear();
    return SIMR_OK;
  } catch (const simr::Reconstructor::HelpRequest& e) {
    g_sr_err = e.what();
    return SIMR_ERR_ARG;
  } catch (const simr::ArgError& e) {
    g_sr_err = e.what();
    return SIMR_ERR_ARG;
  } catch (const simr::NoMemError& e) {
    g_sr_err = e.what();
    return SIMR_ERR_NOMEM;
  } catch (const simr::HipError& e) {
    g_sr_err = e.what();
    return SIMR_ERR_HIP;
  } catch (const simr::StateError& e) {
    g_sr_err = e.what();
    return SIMR_ERR_STATE;
  } catch (const simr::IoError& e) {
    g_sr_err = e.what();
    return SIMR_ERR_IO;
  } catch (const simr::FormatError& e) {
    g_sr_err = e.what();
    return SIMR_ERR_FORMAT;
  } catch (const simr::UnsupportedError& e) {
    g_sr_err = e.what();
    return SIMR_ERR_UNSUPPORTED;
  } catch (const std::bad_alloc&) {
    g_sr_err = "host allocation failed";
    return SIMR_ERR_NOMEM;
  } catch (const std::exception& e) {
    g_sr_err = e.what();
    return SIMR_ERR_ARG;
  }
}

simr::Reconstructor& R(SIM_Reconstructor* sr) {
  if (!sr || !sr->r) throw simr::ArgError("null SIM_Reconstructor handle");
  return *sr->r;
}

SIM_Reconstructor* wrap(simr::Reconstructor* r) {
  SIM_Reconstructor* h = new (std::nothrow) SIM_Reconstructor{r};
  if (!h) {
    delete r;
    throw std::bad_alloc();
  }
  return h;
}

}  // namespace

extern "C" {

SIM_Reconstructor* SR_new_from_shape(int nx, int ny, int nImages, const char* configFileName) {
  SIM_Reconstructor* out = nullptr;
  guard([&] { out = wrap(new simr::Reconstructor(nx, ny, nImages, configFileName ? configFileName : "")); });
  return out;
}

int SR_setRaw(SIM_Reconstructor* sr, const float* raw, int nx, int ny, int nz) {
  return guard([&] { R(sr).set_raw(raw, nx, ny, nz); });
}

int SR_loadAndRescaleImage(SIM_Reconstructor* sr, int it, int iw) {
  return guard([&] { R(sr).load_and_rescale(it, iw); });
}

int SR_setCurTimeIdx(SIM_Reconstructor* sr, int it) { return guard([&] { R(sr).set_cur_time(it); }); }

int SR_processOneVolume(SIM_Reconstructor* sr) {
  int done = 1;
  const int rc = guard([&] { done = R(sr).process_one_volume(); });
  return rc != 0 ? rc : (done ? 0 : 1);
}

int SR_getResult(SIM_Reconstructor* sr, float* result) { return guard([&] { R(sr).get_result(result); }); }

SR_ReconParams* SR_getReconParams(SIM_Reconstructor* sr) {
  SR_ReconParams* out = nullptr;
  guard([&] { out = &R(sr).recon(); });
  return out;
}

SR_ImageParams* SR_getImageParams(SIM_Reconstructor* sr) {
  SR_ImageParams* out = nullptr;
  guard([&] { out = &R(sr).image(); });
  return out;
}

SIM_Reconstructor* SR_new_from_argv(int argc, char** argv) {
  SIM_Reconstructor* out = nullptr;
  guard([&] { out = wrap(new simr::Reconstructor(argc, argv)); });
  return out;
}

int SR_setFile(SIM_Reconstructor* sr, int it, int iw) { return guard([&] { R(sr).set_file(it, iw); }); }

int SR_writeResult(SIM_Reconstructor* sr, int it, int iw) { return guard([&] { R(sr).write_result(it, iw); }); }

int SR_closeFiles(SIM_Reconstructor* sr) { return guard([&] { R(sr).close_files(); }); }

void SR_delete(SIM_Reconstructor* sr) {
  if (!sr) return;
  delete sr->r;
  delete sr;
}

const char* SR_lastError(void) { return g_sr_err.c_str(); }

int SR_getNTimes(SIM_Reconstructor* sr) {
  int n = -1;
  guard([&] { n = R(sr).ntimes(); });
  return n;
}

int SR_parseArgs(int argc, char** argv, SR_ReconParams* recon, SR_ImageParams* img) {
  int kind = 0;
  const int rc = guard([&] {
    if (!recon || !img) throw simr::ArgError("null output struct");
    const simr::cli::Parsed p = simr::cli::parse_argv(argc, argv, *recon, *img);
    kind = p.help ? 1 : (p.version ? 2 : 0);
  });
  return rc != SIMR_OK ? rc : kind;
}

int SR_parseConfig(const char* configFileName, SR_ReconParams* recon, SR_ImageParams* img) {
  return guard([&] {
    if (!recon || !img) throw simr::ArgError("null output struct");
    simr::cli::parse_config(configFileName ? configFileName : "", *recon, *img);
  });
}

}  // extern "C"
