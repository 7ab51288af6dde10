// Command-line and config-file options of cudasirecon.
//
// Restates the option table of SIM_Reconstructor::setupProgramOptions and
// setParams (src/cudaSirecon/cudaSirecon.cpp:913-1101) with the semantics of
// the Boost.Program_options calls it makes (store/notify order, defaults,
// implicit values, switches, positional arguments), written as a small parser
// of its own.  Options are resolved in this order: command line, then the
// config file named by -c/--config (a value given on the command line wins),
// then the option defaults; fields without an option default keep the values
// of SetDefaultParams (cudaSirecon.cpp:16-72).
#pragma once
#include <string>
#include <vector>

#include "../../include/cudasirecon.h"

namespace simr {
namespace cli {

struct Parsed {
  bool help = false, version = false;
  std::string config_file;  // -c/--config as given
  bool config_missing = false;
};

// SetDefaultParams (cudaSirecon.cpp:16-72) plus the ImageParams fields the
// options bind (dxy, dz, wave[0]).
void set_default_params(SR_ReconParams& r, SR_ImageParams& img);

// Full cudasirecon command line (argv[0] is the program name): CLI, then the
// config file, then setParams.  Throws ArgError with Boost-style messages on
// invalid input; a help/version request returns early with the flag set.
Parsed parse_argv(int argc, char** argv, SR_ReconParams& r, SR_ImageParams& img);

// Config file only, as the in-memory constructor does (cudaSirecon.cpp:855-887):
// a missing file leaves the defaults.
void parse_config(const std::string& file, SR_ReconParams& r, SR_ImageParams& img);

std::string usage_text();

void set_str(char* dst, const std::string& s);

}  // namespace cli
}  // namespace simr
