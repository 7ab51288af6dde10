// The cudasirecon executable: the reference's driver loop
// (src/cudaSirecon/cudaSireconDriver.cpp:5-27) over the SR_* ABI.
//
//   cudasirecon [--options] input output otf          (MRC files)
//   cudasirecon [--options] folder pattern otf.tif     (TIFF folder)
//
// As in the reference, errors print "!!Error occurred: ..." and the process
// still exits with status 0; --help and --version exit with status 0.
#include <cstdio>
#include <cstring>

#include "../../include/cudasirecon.h"

int main(int argc, char** argv) {
  SIM_Reconstructor* sr = SR_new_from_argv(argc, argv);
  if (!sr) {
    const char* e = SR_lastError();
    if (std::strcmp(e, "help") == 0 || std::strcmp(e, "version") == 0) return 0;
    std::fprintf(stderr, "\n!!Error occurred: %s\n", e);
    return 0;
  }
  int rc = 0;
  const int nt = SR_getNTimes(sr);
  for (int it = 0; it < nt && rc == 0; ++it) {
    for (int iw = 0; iw < 1 && rc == 0; ++iw) {
      if ((rc = SR_setFile(sr, it, iw)) != 0) break;
      if ((rc = SR_loadAndRescaleImage(sr, it, iw)) != 0) break;
      if ((rc = SR_setCurTimeIdx(sr, it)) != 0) break;
      // 1: a debug dump stopped processing before the filter; nothing to write
      if ((rc = SR_processOneVolume(sr)) == 1) {
        rc = 0;
        continue;
      }
      if (rc != 0) break;
      rc = SR_writeResult(sr, it, iw);
    }
  }
  if (rc == 0) rc = SR_closeFiles(sr);
  if (rc != 0) std::fprintf(stderr, "\n!!Error occurred: %s\n", SR_lastError());
  std::fflush(stdout);
  SR_delete(sr);
  return 0;
}
