// Host-side reconstructor behind the SR_* ABI and the cudasirecon executable.
//
// Mirrors SIM_Reconstructor (src/cudaSirecon/SIM_reconstructor.hpp:19-141;
// cudaSirecon.cpp:789-1700): the command-line and in-memory constructors,
// file discovery (MRC file or a folder of TIFFs), header handling, OTF
// loading, flat-field maps, the per-time-point load / process / write cycle
// and the output file.  The device work is delegated to a simr context
// (include/simr.h); this class owns one and never touches HIP directly.
#pragma once
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cudasirecon.h"
#include "../../include/simr.h"
#include "io_mrc.hpp"
#include "io_tiff.hpp"

namespace simr {

class Reconstructor {
 public:
  // Command-line constructor (cudaSirecon.cpp:792-856).  Throws on error;
  // help/version requests throw HelpRequest after printing.
  Reconstructor(int argc, char** argv);
  // In-memory constructor (cudaSirecon.cpp:860-887).
  Reconstructor(int nx, int ny, int nimages, const std::string& config);
  ~Reconstructor();

  void set_raw(const float* raw, int nx, int ny, int nz);
  void set_file(int it, int iw);
  void load_and_rescale(int it, int iw);
  void set_cur_time(int it);
  int process_one_volume();
  void get_result(float* out);
  void write_result(int it, int iw);
  void close_files();
  int ntimes() const { return I_.ntimes; }
  SR_ReconParams& recon() { return P_; }
  SR_ImageParams& image() { return I_; }
  size_t result_elems() const;

  struct HelpRequest : std::runtime_error {
    using std::runtime_error::runtime_error;
  };

 private:
  void check_supported() const;
  void setup_common();
  void determine_otf_dimensions();
  void load_otfs();
  void load_flatfield();
  simr_params engine_params() const;
  void sync_engine();  // create the context, or bring it up to date with P_/I_
  void upload_raw(const float* f32, const uint16_t* u16);
  std::string output_path(const std::string& input) const;
  void open_dumps();
  void write_separated();
  void write_overlaps();
  void close_dumps();

  SR_ReconParams P_{};
  SR_ImageParams I_{};
  SR_ReconParams built_P_{};
  SR_ImageParams built_I_{};
  simr_ctx* ctx_ = nullptr;
  bool loaded_ = false, processed_ = false;

  std::vector<std::string> argv_;
  std::vector<std::string> tiff_files_;
  std::string out_dir_;
  std::unique_ptr<mrc::Reader> in_mrc_;
  std::unique_ptr<mrc::Writer> out_mrc_;
  // debug dumps (--saveprefiltered, --savealignedraw, --saveoverlaps), MRC mode
  std::unique_ptr<mrc::Writer> sep_mrc_, aligned_mrc_, ovl_mrc_;
  mrc::Header io_header_;
  std::unique_ptr<tiff::Reader> otf_tiff_;
  std::unique_ptr<mrc::Reader> otf_mrc_;
  bool otf_valid_ = false;
  std::vector<float> otf_;  // [nDirsOTF][norders][nxotf*nzotf] complex
  std::vector<float> bg_, slope_;
  std::vector<float> host_out_;
  bool minmax_done_ = false;
};

}  // namespace simr
