/*
 * cudasirecon.h - drop-in C ABI of the cudasirecon shared library.
 *
 * Every SR_* entry point the reference exports from libcudasirecon.so
 * (src/cudaSirecon/interface.cpp:10-70) is declared here with the same name
 * and argument meaning, backed by the MI355X engine of simr.h.  Differences a
 * caller can see, all additive:
 *   - functions the reference declares void return int: 0 on success, a
 *     negative simr_status on failure (the reference throws C++ exceptions
 *     across extern "C" or calls exit(-1)); callers that ignore the return
 *     value keep working;
 *   - constructors return NULL on failure; SR_lastError() says why;
 *   - SR_getReconParams / SR_getImageParams return pointers to plain C structs
 *     (the reference returns C++ references to structs holding std::vector and
 *     std::string); the fields keep the reference's names and meaning;
 *   - SR_delete frees a handle (the reference has no destructor entry point);
 *   - every handle owns its own HIP stream and device arena, so any number of
 *     handles may live in one process (the reference keeps process-global
 *     state and allows one).
 */
#ifndef CUDASIRECON_H
#define CUDASIRECON_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SR_PATH_MAX 4096
#define SR_MAX_LIST 8

typedef struct SIM_Reconstructor SIM_Reconstructor;

/* Mirrors ReconParams (src/cudaSirecon/cudaSireconImpl.h:104-180): same field
 * names and meaning; bool fields are int32; std::vector<float> fields are a
 * count plus a fixed array; std::string fields are NUL-terminated buffers. */
typedef struct SR_ReconParams {
  float k0startangle, linespacing;
  float na, nimm;
  int32_t ndirs, nphases, norders_output;
  int32_t norders;
  int32_t bTwolens, bFastSIM, bBessel;
  float BesselNA, BesselLambdaEx;
  float deskewAngle;
  int32_t extraShift, bNoRecon;
  uint32_t cropXYto;
  int32_t bWriteTitle;
  float otfcutoff, zoomfact;
  int32_t z_zoom, nzPadTo;
  float explodefact;
  int32_t bFilteroverlaps, recalcarrays, napodize, bSearchforvector, bUseTime0k0, apodizeoutput;
  float apoGamma;
  int32_t bSuppress_singularities, suppression_radius, bDampenOrder0, bFitallphases, do_rescale;
  int32_t equalizez, equalizet, bNoKz0;
  float wiener, wienerInr;
  int32_t nforceamp;
  float forceamp[SR_MAX_LIST];
  int32_t nk0angles;
  float k0angles[SR_MAX_LIST];
  int32_t nxotf, nyotf, nzotf;
  float dzPSF, dkzotf, dkrotf;
  int32_t bRadAvgOTF, bOneOTFperAngle;
  int32_t bFixdrift;
  float drift_filter_fact;
  float constbkgd;
  int32_t bBgInExtHdr, bUsecorr;
  char corrfiles[SR_PATH_MAX];
  float readoutNoiseVar, electrons_per_bit;
  int32_t bMakemodel, bSaveSeparated;
  char fileSeparated[SR_PATH_MAX];
  int32_t bSaveAlignedRaw;
  char fileRawAligned[SR_PATH_MAX];
  int32_t bSaveOverlaps;
  char fileOverlaps[SR_PATH_MAX];
  int32_t bTIFF;
  char ifiles[SR_PATH_MAX], ofiles[SR_PATH_MAX], otffiles[SR_PATH_MAX];
} SR_ReconParams;

/* Mirrors ImageParams (src/cudaSirecon/cudaSireconImpl.h:181-195). */
typedef struct SR_ImageParams {
  int32_t nx, nx_raw, ny, nz, nz0;
  int16_t nwaves, wave[5], ntimes;
  uint16_t curTimeIdx;
  float dxy, dz, dz_raw, inscale;
} SR_ImageParams;

/* interface.cpp:14-18.  In-memory reconstructor: nImages raw sections of
 * nx x ny per volume; parameters from a config file of key=value lines (the
 * cudasirecon option names); the OTF is read from the TIFF named by its
 * otf-file key.  Pixel sizes come from xyres/zres and the emission
 * wavelength from wavelength (reference defaults 0.1, 0.2, 530). */
SIM_Reconstructor* SR_new_from_shape(int nx, int ny, int nImages, const char* configFileName);

/* interface.cpp:20-26.  raw: float32 [nz][ny][nx] (nz = nImages), borrowed and
 * copied; section order Angle->Z->Phase, or Z->Angle->Phase with fastSI. */
int SR_setRaw(SIM_Reconstructor* sr, const float* raw, int nx, int ny, int nz);

/* interface.cpp:27-30.  Bleach correction (rescale) of the loaded stack for time it. */
int SR_loadAndRescaleImage(SIM_Reconstructor* sr, int it, int iw);

/* interface.cpp:31-34. */
int SR_setCurTimeIdx(SIM_Reconstructor* sr, int it);

/* interface.cpp:35-38.  k0/modamp fit, Wiener filter and assembly.  Returns 0
 * after a reconstruction, or 1 when --saveprefiltered, --savealignedraw or
 * --saveoverlaps wrote their debug dumps and stopped before the filter (the
 * reference's processOneVolume returns 0 then and its driver skips
 * writeResult, cudaSirecon.cpp:1286-1288, cudaSireconDriver.cpp:15-16). */
int SR_processOneVolume(SIM_Reconstructor* sr);

/* interface.cpp:39-42.  result: caller-allocated
 * (zoomfact*nx)*(zoomfact*ny)*(z_zoom*nz0) floats, layout [z][y][x]. */
int SR_getResult(SIM_Reconstructor* sr, float* result);

/* interface.cpp:43-51.  Live parameter blocks of the handle.  Fields changed by
 * the caller take effect at the next SR_setRaw / SR_setFile (the handle
 * rebuilds its device state then); image sizes must not be changed. */
SR_ReconParams* SR_getReconParams(SIM_Reconstructor* sr);
SR_ImageParams* SR_getImageParams(SIM_Reconstructor* sr);

/* interface.cpp:55-58.  Command-line reconstructor (the cudasirecon CLI): argv
 * as main() receives it.  --help / --version print and return NULL with
 * SR_lastError() == "help" / "version" (the reference exits the process). */
SIM_Reconstructor* SR_new_from_argv(int argc, char** argv);

/* interface.cpp:60-63.  Load time point it (wave iw) from the input file(s). */
int SR_setFile(SIM_Reconstructor* sr, int it, int iw);

/* interface.cpp:64-67.  Write the current result of time point it. */
int SR_writeResult(SIM_Reconstructor* sr, int it, int iw);

/* interface.cpp:68-71.  Finish the output file (MRC header min/max, titles). */
int SR_closeFiles(SIM_Reconstructor* sr);

/* Extensions (no counterpart in the reference). */
void SR_delete(SIM_Reconstructor* sr);
const char* SR_lastError(void);
/* Number of time points of the input (SIM_Reconstructor::getNTimes,
 * src/cudaSirecon/SIM_reconstructor.hpp:85). */
int SR_getNTimes(SIM_Reconstructor* sr);
/* Parse a command line and its config file exactly as SR_new_from_argv does,
 * without opening data files or touching a GPU.  Returns 0, 1 for --help,
 * 2 for --version, or a negative status. */
int SR_parseArgs(int argc, char** argv, SR_ReconParams* recon, SR_ImageParams* img);
/* Parse a config file as SR_new_from_shape does (no OTF or GPU access). */
int SR_parseConfig(const char* configFileName, SR_ReconParams* recon, SR_ImageParams* img);

#ifdef __cplusplus
}
#endif
#endif /* CUDASIRECON_H */
