// Host-side reconstructor (see sr_reconstructor.hpp).
#include "sr_reconstructor.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <regex>

#include "errors.hpp"
#include "sr_options.hpp"

namespace fs = std::filesystem;

namespace simr {

namespace {

constexpr const char* kVersion = "1.2.0";

void check(int rc) {
  if (rc == SIMR_OK) return;
  const std::string msg = simr_last_error();
  switch (rc) {
    case SIMR_ERR_HIP: throw HipError(msg);
    case SIMR_ERR_NOMEM: throw NoMemError(msg);
    case SIMR_ERR_IO: throw IoError(msg);
    case SIMR_ERR_FORMAT: throw FormatError(msg);
    case SIMR_ERR_STATE: throw StateError(msg);
    case SIMR_ERR_UNSUPPORTED: throw UnsupportedError(msg);
    default: throw ArgError(msg);
  }
}

// gatherMatchingFiles (src/cudaSirecon/boostfs.cpp:9-55): regular files of
// `folder` whose full path matches ".*<pattern>" (".*\.[tT][iI][fF]" is
// appended unless the pattern names .tif/.TIF), sorted by name; the output
// folder <folder>/GPUsirecon is created in any case.
std::vector<std::string> gather_matching_files(const std::string& folder, std::string pattern, std::string& out_dir) {
  if (pattern.rfind(".tif") == std::string::npos && pattern.rfind(".TIF") == std::string::npos)
    pattern.append(".*\\.[tT][iI][fF]");
  pattern.insert(0, ".*");
  const std::regex filter(pattern);
  std::vector<std::string> files;
  for (const auto& e : fs::directory_iterator(folder)) {
    if (!e.is_regular_file()) continue;
    const std::string name = e.path().string();
    if (std::regex_match(name, filter)) files.push_back(name);
  }
  std::sort(files.begin(), files.end());
  const fs::path out = fs::path(folder) / "GPUsirecon";
  if (!fs::exists(out)) fs::create_directory(out);
  out_dir = out.string();
  return files;
}

}  // namespace

// ------------------------------------------------------------ constructors

Reconstructor::Reconstructor(int argc, char** argv) {
  for (int i = 0; i < argc; i++) argv_.push_back(argv[i] ? argv[i] : "");
  const cli::Parsed parsed = cli::parse_argv(argc, argv, P_, I_);
  if (parsed.help) {
    std::printf("cudasirecon v%s (MI355X engine)\n\n%s\n", kVersion, cli::usage_text().c_str());
    throw HelpRequest("help");
  }
  if (parsed.version) {
    std::printf("Version %s\n", kVersion);
    throw HelpRequest("version");
  }
  std::printf("nphases=%d, ndirs=%d\n", P_.nphases, P_.ndirs);

  // TIFF mode: the input is a folder holding TIFFs whose names match the
  // output-file pattern (cudaSirecon.cpp:845-862)
  std::error_code ec;
  if (P_.ifiles[0] && fs::is_directory(P_.ifiles, ec)) {
    tiff_files_ = gather_matching_files(P_.ifiles, P_.ofiles, out_dir_);
    P_.bTIFF = !tiff_files_.empty();
  }
  if (P_.bTIFF) I_.ntimes = (int16_t)tiff_files_.size();
  check_supported();

  // openFiles (cudaSirecon.cpp:1073-1104)
  if (!P_.ifiles[0] || !P_.ofiles[0])
    throw ArgError("Calling openFiles() but no input or output files were specified");
  if (!P_.bTIFF) {
    try {
      in_mrc_ = std::make_unique<mrc::Reader>(P_.ifiles);
    } catch (const IoError&) {
      throw IoError("No matching input TIFF or MRC files not found");
    }
    try {
      otf_mrc_ = std::make_unique<mrc::Reader>(P_.otffiles);
    } catch (const IoError&) {
      throw IoError("OTF file not found");
    }
  } else {
    otf_tiff_ = std::make_unique<tiff::Reader>(P_.otffiles);
  }
  otf_valid_ = true;

  // setup (cudaSirecon.cpp:1118-1135)
  if (P_.bTIFF) {
    const tiff::Reader t0(tiff_files_[0]);
    I_.nx = 0;
    I_.nx_raw = t0.width();
    I_.ny = t0.height();
    I_.nz = t0.pages();
    I_.nwaves = (int16_t)t0.samples();
    I_.nz /= I_.nwaves * P_.nphases * P_.ndirs;
  } else {
    // loadHeader (src/cudaSirecon/mrc.h:18-45)
    const mrc::Header& h = in_mrc_->header();
    I_.nx = 0;
    I_.nx_raw = h.nx;
    I_.ny = h.ny;
    I_.nz = h.nz / (h.num_waves * h.num_times);
    I_.nz /= P_.nphases * P_.ndirs;
    I_.nwaves = h.num_waves;
    I_.ntimes = h.num_times;
    I_.wave[0] = h.iwav1;
    I_.wave[1] = h.iwav2;
    I_.wave[2] = h.iwav3;
    I_.wave[3] = h.iwav4;
    I_.wave[4] = h.iwav5;
    I_.dxy = h.ylen;
    I_.dz = h.zlen;
  }
  std::printf("nx_raw=%d, ny=%d, nz=%d\n", I_.nx_raw, I_.ny, I_.nz);
  setup_common();
  open_dumps();

  if (!P_.bTIFF) {
    // setOutputHeader (src/cudaSirecon/mrc.h:47-60)
    io_header_ = in_mrc_->header();
    io_header_.mode = mrc::kFloat;
    io_header_.nz = I_.nz * I_.nwaves * I_.ntimes * P_.z_zoom;
    io_header_.nx = (int)((float)I_.nx * P_.zoomfact);
    io_header_.ny = (int)((float)io_header_.ny * P_.zoomfact);
    io_header_.xlen /= P_.zoomfact;
    io_header_.ylen /= P_.zoomfact;
    io_header_.zlen /= (float)P_.z_zoom;
    io_header_.next = 0;
    try {
      out_mrc_ = std::make_unique<mrc::Writer>(P_.ofiles, io_header_);
    } catch (const IoError&) {
      std::fprintf(stderr, "File %s can not be created.\n", P_.ofiles);
      throw IoError("Error creating output file");
    }
  }
  load_flatfield();
  sync_engine();
}

Reconstructor::Reconstructor(int nx, int ny, int nimages, const std::string& config) {
  cli::parse_config(config, P_, I_);
  if (!P_.otffiles[0]) throw IoError("no OTF file: the config file names no otf-file");
  otf_tiff_ = std::make_unique<tiff::Reader>(P_.otffiles);
  otf_valid_ = true;
  P_.bTIFF = 1;  // in-memory data: no MRC files (cudaSirecon.cpp:879)
  check_supported();
  std::printf("zoomz: %g\n", (double)P_.zoomfact);
  if (nx <= 0 || ny <= 0 || nimages <= 0) throw ArgError("image shape must be positive");
  I_.nx = nx;
  I_.nx_raw = nx;
  I_.ny = ny;
  I_.nz = nimages;
  I_.nwaves = 1;
  I_.ntimes = 1;
  I_.nz /= I_.nwaves * P_.nphases * P_.ndirs;
  setup_common();
  load_flatfield();
  sync_engine();
}

Reconstructor::~Reconstructor() {
  if (ctx_) simr_destroy(ctx_);
}

// The reference's optional branches that this engine does not implement are
// refused up front instead of producing a partial result.
// notGoodDimension / findOptimalDimension (cudaSirecon.cpp:77-97): step down
// to the next size whose only prime factors are 2, 3, 5 and 7.
static bool not_good_dimension(unsigned n) {
  if (n == 2 || n == 3 || n == 5 || n == 7) return false;
  if (n % 2 == 0) return not_good_dimension(n / 2);
  if (n % 3 == 0) return not_good_dimension(n / 3);
  if (n % 5 == 0) return not_good_dimension(n / 5);
  if (n % 7 == 0) return not_good_dimension(n / 7);
  return true;
}
static unsigned find_optimal_dimension(unsigned n) {
  while (n > 1 && not_good_dimension(n)) n -= 1;
  return n;
}

void Reconstructor::check_supported() const {
  // --noRecon only acts together with --bessel in the reference: loadAndRescaleImage
  // (cudaSirecon.cpp:1325) then returns before loading the stack, and the
  // reconstruction that follows runs on whatever the device buffers held.
  if (P_.bBessel && P_.bNoRecon)
    throw UnsupportedError("--noRecon together with --bessel is not supported (the reference skips loading the "
                           "data and reconstructs uninitialised buffers)");
  if (P_.bTwolens) throw UnsupportedError("--2lenses is not supported by this engine");
  if (P_.bUsecorr && (std::fabs(P_.deskewAngle) > 0.f || (P_.cropXYto > 0 && (int)P_.cropXYto < I_.nx_raw)))
    throw UnsupportedError(
        "--usecorr together with --deskew or --cropXY is not supported (the reference indexes the "
        "calibration maps with the raw width but sizes them with the resampled one)");
}

// setup_common (cudaSirecon.cpp:1139-1261): the host part.  The separation
// matrix, inscale and k0 guesses are derived inside the engine context.
void Reconstructor::setup_common() {
  if (I_.nz < 1) throw ArgError("fewer raw images than ndirs*nphases");
  I_.nz0 = I_.nz;  // nzPadTo is never set
  // deskewing widens nx to the next smaller 2/3/5/7-smooth size that holds the
  // sheared stack and scales dz by sin(angle) (cudaSirecon.cpp:1153-1165);
  // cropping then narrows nx (1168-1171)
  if (std::fabs(P_.deskewAngle) > 0.f) {
    if (P_.deskewAngle < 0) P_.deskewAngle += 180.f;
    const double a = (double)P_.deskewAngle * M_PI / 180.;
    I_.nx = (int)find_optimal_dimension(
        (unsigned)((double)I_.nx_raw + (double)I_.nz0 * std::fabs(std::cos(a)) * (double)I_.dz / (double)I_.dxy));
    I_.dz_raw = I_.dz;
    I_.dz = (float)((double)I_.dz * std::fabs(std::sin(a)));
  } else {
    I_.nx = I_.nx_raw;
    I_.dz_raw = 0;
  }
  if (P_.cropXYto > 0 && (int)P_.cropXYto < I_.nx) I_.nx = (int)P_.cropXYto;
  std::printf("nx=%d, ny=%d, nz=%d, nz0 = %d, nwaves=%d\n", I_.nx, I_.ny, I_.nz, I_.nz0, I_.nwaves);
  std::printf("dxy=%f, dz=%f um\n", I_.dxy, I_.dz);
  P_.norders = P_.norders_output != 0 ? P_.norders_output : P_.nphases / 2 + 1;
  std::printf("nphases=%d, norders=%d, ndirs=%d\n", P_.nphases, P_.norders, P_.ndirs);
  if (otf_valid_) {
    determine_otf_dimensions();
    load_otfs();
  }
  I_.inscale = (float)(1.0 / ((float)(I_.nx * I_.ny * I_.nz0) * P_.zoomfact * P_.zoomfact * (float)P_.z_zoom *
                              (float)P_.ndirs));
}

// determine_otf_dimensions (cudaSirecon.cpp:1444-1530)
void Reconstructor::determine_otf_dimensions() {
  if (P_.bTIFF) {
    const int nxotf = otf_tiff_->width() / 2, nyotf = otf_tiff_->height(), nzotf = otf_tiff_->pages();
    if (I_.nz == 1) {
      P_.nxotf = nxotf;
      P_.nyotf = P_.bRadAvgOTF ? 1 : nyotf;
      P_.nzotf = 1;
      P_.dkrotf = 1.f / (I_.dxy * (float)(nxotf - 1) * 2.f);
      P_.dkzotf = 1;
    } else if (P_.bRadAvgOTF) {
      P_.nzotf = nxotf;
      P_.nxotf = nyotf;
      P_.nyotf = 1;
      P_.dkzotf = 1.f / (P_.dzPSF * (float)P_.nzotf);
      P_.dkrotf = 1.f / (I_.dxy * (float)(P_.nxotf - 1) * 2.f);
    } else {
      P_.nzotf = nzotf / P_.norders;
      P_.nxotf = nxotf;
      P_.nyotf = nyotf;
      P_.dkzotf = 1.f / (P_.dzPSF * (float)P_.nzotf);
      P_.dkrotf = 1.f / (I_.dxy * (float)P_.nxotf);
    }
  } else {
    const mrc::Header& h = otf_mrc_->header();
    if (I_.nz == 1) {
      P_.nxotf = h.nx;
      P_.nyotf = P_.bRadAvgOTF ? 1 : h.ny;
      P_.nzotf = 1;
      P_.dkrotf = h.xlen;
      P_.dkzotf = 1;
    } else if (P_.bRadAvgOTF) {
      P_.nzotf = h.nx;
      P_.nxotf = h.ny;
      P_.nyotf = 1;
      P_.dkzotf = h.xlen;
      P_.dkrotf = h.ylen;
    } else {
      P_.nzotf = h.nz / P_.norders;
      if (P_.bOneOTFperAngle) P_.nzotf /= P_.ndirs;
      P_.nxotf = h.nx;
      P_.nyotf = h.ny;
      P_.dkzotf = h.zlen;
      P_.dkrotf = h.xlen;
    }
  }
  std::printf("nzotf=%d, dkzotf=%f, nxotf=%d, nyotf=%d, dkrotf=%f\n", P_.nzotf, P_.dkzotf, P_.nxotf, P_.nyotf,
              P_.dkrotf);
  if (I_.nz > 1 && !P_.bRadAvgOTF && (P_.nzotf < 1 || !(P_.dkzotf > 0) || !(P_.dkrotf > 0)))
    throw ArgError("the OTF file does not hold a full 3D OTF (norders x nzotf sections with positive kx and kz "
                   "steps); pass --otfRA for a rotationally averaged OTF");
}

// loadOTFs (cudaSirecon.cpp:1532-1617).  Rotationally averaged or 2D:
// section (dir*norders + order) of an MRC file, or page `order` of a TIFF
// (the TIFF branch reads page `order` for every direction); when the file has
// fewer sections the order-0 OTF of direction 0 is repeated.  A full 3D OTF
// (no --otfRA): nzotf consecutive MRC sections per (direction, order), or
// TIFF pages order*nzotf .. order*nzotf + nzotf - 1 for every direction.
void Reconstructor::load_otfs() {
  const int ndo = P_.bOneOTFperAngle ? P_.ndirs : 1;
  const size_t size = (size_t)P_.nzotf * P_.nxotf * P_.nyotf;
  otf_.assign((size_t)ndo * P_.norders * size * 2, 0.f);
  auto slot = [&](int dir, int order) { return otf_.data() + ((size_t)dir * P_.norders + order) * size * 2; };
  std::vector<float> page;
  long long next_sec = 0;
  const bool full3d = I_.nz > 1 && !P_.bRadAvgOTF;
  for (int dir = 0; dir < ndo; dir++) {
    for (int i = 0; i < P_.norders; i++) {
      if (full3d) {
        const size_t plane = (size_t)P_.nxotf * P_.nyotf;
        for (int z = 0; z < P_.nzotf; z++) {
          if (P_.bTIFF) {
            const long long pg = (long long)i * P_.nzotf + z;
            if (pg >= otf_tiff_->pages()) throw FormatError("OTF TIFF has fewer pages than norders * nzotf");
            page.resize((size_t)otf_tiff_->width() * otf_tiff_->height());
            if (page.size() < 2 * plane) throw FormatError("OTF TIFF page smaller than an OTF plane");
            otf_tiff_->read_f32(pg, page.data());
          } else {
            const mrc::Header& h = otf_mrc_->header();
            if (next_sec >= h.nz) throw FormatError("OTF MRC file has fewer sections than the OTF it should hold");
            page.resize((size_t)h.nx * h.ny * 2);
            otf_mrc_->read_c64(next_sec++, page.data());
          }
          std::memcpy(slot(dir, i) + z * plane * 2, page.data(), plane * 2 * sizeof(float));
        }
        continue;
      }
      const long long nsec = P_.bTIFF ? otf_tiff_->pages() : otf_mrc_->header().nz;
      if (nsec > i + dir * P_.norders) {
        if (P_.bTIFF) {
          const size_t n = (size_t)otf_tiff_->width() * otf_tiff_->height();
          if (n < 2 * size) throw FormatError("OTF TIFF page smaller than the OTF it should hold");
          page.resize(n);
          otf_tiff_->read_f32(i, page.data());
          std::memcpy(slot(dir, i), page.data(), size * 2 * sizeof(float));
        } else {
          const mrc::Header& h = otf_mrc_->header();
          if ((size_t)h.nx * h.ny < size) throw FormatError("OTF MRC section smaller than the OTF it should hold");
          page.resize((size_t)h.nx * h.ny * 2);
          otf_mrc_->read_c64(next_sec++, page.data());
          std::memcpy(slot(dir, i), page.data(), size * 2 * sizeof(float));
        }
      } else {
        std::memcpy(slot(dir, i), slot(0, 0), size * 2 * sizeof(float));
      }
    }
  }
  otf_mrc_.reset();
}

// bgAndSlope / getbg_and_slope (cudaSirecon.cpp:139-192): constant
// background and unit slope, or the two sections of an MRC calibration file.
void Reconstructor::load_flatfield() {
  bg_.clear();
  slope_.clear();
  if (!P_.bUsecorr) return;
  std::printf("loading CCD calibration file\n");
  std::unique_ptr<mrc::Reader> r;
  try {
    r = std::make_unique<mrc::Reader>(P_.corrfiles);
  } catch (const IoError&) {
    throw IoError(std::string("File ") + P_.corrfiles + " does not exist");
  }
  const mrc::Header& h = r->header();
  if (h.nx != I_.nx || h.ny != I_.ny)
    throw FormatError(std::string("calibration file ") + P_.corrfiles + " has different dimension than data file");
  if (h.nz < 2) throw FormatError(std::string("calibration file ") + P_.corrfiles + " needs two sections");
  bg_.resize((size_t)I_.nx * I_.ny);
  slope_.resize(bg_.size());
  r->read_f32(0, bg_.data());
  r->read_f32(1, slope_.data());
}

// ------------------------------------------------------------ engine link

simr_params Reconstructor::engine_params() const {
  simr_params p;
  simr_params_default(&p);
  p.nx = I_.nx;
  p.nx_raw = I_.nx_raw;
  // loadImageData (cudaSirecon.cpp:1355-1360): cos(angle) dz_raw / dxy in double, stored as float
  p.deskew_factor = std::fabs(P_.deskewAngle) > 0.f
                        ? (float)(std::cos((double)P_.deskewAngle * M_PI / 180.) * (double)I_.dz_raw / (double)I_.dxy)
                        : 0.f;
  p.deskew_shift = P_.extraShift;
  p.ny = I_.ny;
  p.nz = I_.nz0;
  p.ndirs = P_.ndirs;
  p.nphases = P_.nphases;
  p.norders_output = P_.norders_output;
  p.dxy = I_.dxy;
  p.dz = I_.dz;
  p.wave = I_.wave[0];
  p.ntimes = I_.ntimes;
  p.k0startangle = P_.k0startangle;
  p.linespacing = P_.linespacing;
  p.na = P_.na;
  p.nimm = P_.nimm;
  p.nk0angles = std::min(P_.nk0angles, (int32_t)SIMR_MAX_DIRS);
  for (int i = 0; i < p.nk0angles; i++) p.k0angles[i] = P_.k0angles[i];
  p.otfcutoff = P_.otfcutoff;
  p.zoomfact = P_.zoomfact;
  p.z_zoom = P_.z_zoom;
  p.explodefact = P_.explodefact;
  p.bFilteroverlaps = P_.bFilteroverlaps;
  p.recalcarrays = P_.recalcarrays;
  p.napodize = P_.napodize;
  p.bSearchforvector = P_.bSearchforvector;
  p.bUseTime0k0 = P_.bUseTime0k0;
  p.apodizeoutput = P_.apodizeoutput;
  p.apoGamma = P_.apoGamma;
  p.bSuppress_singularities = P_.bSuppress_singularities;
  p.suppression_radius = P_.suppression_radius;
  p.bDampenOrder0 = P_.bDampenOrder0;
  p.bFitallphases = P_.bFitallphases;
  p.do_rescale = P_.do_rescale;
  p.equalizez = P_.equalizez;
  p.equalizet = P_.equalizet;
  p.bNoKz0 = P_.bNoKz0;
  p.wiener = P_.wiener;
  p.nforceamp = std::min(P_.nforceamp, (int32_t)SIMR_MAX_ORDERS);
  for (int i = 0; i < p.nforceamp; i++) p.forceamp[i] = P_.forceamp[i];
  p.bRadAvgOTF = P_.bRadAvgOTF;
  p.bOneOTFperAngle = P_.bOneOTFperAngle;
  p.nxotf = P_.nxotf;
  p.nyotf = P_.nyotf;
  p.nzotf = P_.nzotf;
  p.dkrotf = P_.dkrotf;
  p.dkzotf = P_.dkzotf;
  p.constbkgd = P_.constbkgd;
  p.bFastSIM = P_.bFastSIM;
  p.bBessel = P_.bBessel ? 1 : 0;
  p.BesselNA = P_.BesselNA;
  p.BesselLambdaEx = P_.BesselLambdaEx;
  return p;
}

void Reconstructor::sync_engine() {
  check_supported();
  const simr_params p = engine_params();
  if (!ctx_) {
    check(simr_create(&p, otf_.data(), 0, &ctx_));
    if (P_.bUsecorr) check(simr_set_flatfield(ctx_, bg_.data(), slope_.data()));
    return;
  }
  simr_params cur;
  check(simr_get_params(ctx_, &cur));
  if (std::memcmp(&cur, &p, sizeof(p)) == 0) return;
  if (simr_update_params(ctx_, &p) == SIMR_OK) return;
  if (loaded_) throw ArgError(std::string("parameter change needs a new context; reload the raw data first: ") +
                              simr_last_error());
  simr_destroy(ctx_);
  ctx_ = nullptr;
  check(simr_create(&p, otf_.data(), 0, &ctx_));
  if (P_.bUsecorr) check(simr_set_flatfield(ctx_, bg_.data(), slope_.data()));
}

void Reconstructor::upload_raw(const float* f32, const uint16_t* u16) {
  loaded_ = false;
  sync_engine();
  if (u16)
    check(simr_set_raw_u16_host(ctx_, u16));
  else
    check(simr_set_raw_f32_host(ctx_, f32));
  loaded_ = true;
  processed_ = false;
}

// ------------------------------------------------------------ per time point

void Reconstructor::set_raw(const float* raw, int nx, int ny, int nz) {
  if (!raw) throw ArgError("raw data pointer is null");
  const long long nimg = (long long)I_.nz * P_.nphases * P_.ndirs;
  if (nx != I_.nx_raw || ny != I_.ny || nz < nimg)
    throw ArgError("raw stack " + std::to_string(nx) + "x" + std::to_string(ny) + "x" + std::to_string(nz) +
                   " does not match the reconstructor's " + std::to_string(I_.nx_raw) + "x" + std::to_string(I_.ny) +
                   "x" + std::to_string(nimg));
  upload_raw(raw, nullptr);
}

// setFile (cudaSirecon.cpp:1337-1352): one time point of one wave.
void Reconstructor::set_file(int it, int iw) {
  if (it < 0 || it >= I_.ntimes) throw ArgError("time index out of range");
  if (iw < 0 || iw >= I_.nwaves) throw ArgError("wave index out of range");
  const long long nimg = (long long)I_.nz * P_.nphases * P_.ndirs;
  const size_t plane = (size_t)I_.nx_raw * I_.ny;
  if (P_.bTIFF) {
    const tiff::Reader t(tiff_files_.at(it));
    if (t.width() != I_.nx_raw || t.height() != I_.ny || t.pages() < nimg)
      throw FormatError("TIFF stack " + tiff_files_[it] + " does not match the first file's size");
    if (t.is_u16()) {
      std::vector<uint16_t> buf(plane * nimg);
      for (long long s = 0; s < nimg; s++) t.read_u16((int)s, buf.data() + plane * s);
      upload_raw(nullptr, buf.data());
    } else {
      std::vector<float> buf(plane * nimg);
      for (long long s = 0; s < nimg; s++) t.read_f32((int)s, buf.data() + plane * s);
      upload_raw(buf.data(), nullptr);
    }
  } else {
    const mrc::Header& h = in_mrc_->header();
    if (h.mode == mrc::kUint16) {
      std::vector<uint16_t> buf(plane * nimg);
      for (long long s = 0; s < nimg; s++) in_mrc_->read_u16(mrc::section_of(h, (int)s, iw, it), buf.data() + plane * s);
      upload_raw(nullptr, buf.data());
    } else {
      std::vector<float> buf(plane * nimg);
      for (long long s = 0; s < nimg; s++) in_mrc_->read_f32(mrc::section_of(h, (int)s, iw, it), buf.data() + plane * s);
      upload_raw(buf.data(), nullptr);
    }
  }
}

void Reconstructor::load_and_rescale(int it, int /*iw*/) {
  if (!loaded_) throw StateError("no raw data loaded: call SR_setRaw or SR_setFile first");
  check(simr_load_and_rescale(ctx_, it));
}

void Reconstructor::set_cur_time(int it) {
  I_.curTimeIdx = (uint16_t)it;
  if (ctx_) check(simr_set_time_index(ctx_, it));
}

// processOneVolume (cudaSirecon.cpp:1264-1321).  Returns 1 when the volume
// was reconstructed, 0 when a debug dump was requested: the dumps are written
// and processing stops before the Wiener filter, as saveIntermediateDataForDebugging
// makes it (src/cudaSirecon/mrc.h:76-96; cudaSirecon.cpp:1286-1288), and
// there is no result to write.
int Reconstructor::process_one_volume() {
  if (!loaded_) throw StateError("no raw data loaded: call SR_setRaw or SR_setFile first");
  sync_engine();
  const bool dumps = P_.bSaveSeparated || P_.bSaveAlignedRaw || P_.bSaveOverlaps;
  if (P_.bSaveSeparated && !P_.bTIFF) {
    // the separated bands replace the k0 search of every direction (cudaSirecon.cpp:269-287)
    write_separated();
  } else {
    if (P_.bSaveSeparated)
      for (int d = 0; d < P_.ndirs; d++) std::printf("No unmixed raw images were saved in TIFF mode\n");
    check(dumps ? simr_find_modulation(ctx_) : simr_process_volume(ctx_));
    const float dkx = 1.f / ((float)I_.nx * I_.dxy), dky = 1.f / ((float)I_.ny * I_.dxy);
    for (int d = 0; d < P_.ndirs; d++) {
      simr_fit f;
      check(simr_get_fit(ctx_, d, &f));
      std::printf("direction %d: k0 = (%f, %f) pixels, guess (%f, %f)\n", d, f.k0x / dkx, f.k0y / dky,
                  f.k0guess_x / dkx, f.k0guess_y / dky);
      for (int o = 1; o < P_.norders; o++)
        std::printf("  order %d: modamp mag=%f, phase=%f, correlation coeff=%f\n", o,
                    std::sqrt(f.amp_re[o] * f.amp_re[o] + f.amp_im[o] * f.amp_im[o]),
                    std::atan2(f.amp_im[o], f.amp_re[o]), f.corr[o]);
    }
    if (P_.bSaveOverlaps) write_overlaps();
  }
  if (dumps) {
    std::printf("\nQuit processing because either bSaveSeparated, bSaveAlignedRaw, or bSaveOverlaps is TRUE\n");
    processed_ = false;
    return 0;
  }
  processed_ = true;
  return 1;
}

// Debug dump files, MRC mode (setup_common, cudaSirecon.cpp:1179-1209): each
// header is the input file's with the changes below; no extended header.
//   savealignedraw:  float, nx = the reconstruction width.  Drift correction
//                    is unreachable in the reference, so no section is ever
//                    written to it: the file holds the header only.
//   saveprefiltered: complex, nx = nx/2+1: per time point, per direction,
//                    nphases separated 3D spectra (band 0, 1re, 1im, 2re, 2im),
//                    nz sections each.
//   saveoverlaps:    complex, nx = nx, nz = nz*2*ndirs*ntimes*nwaves, two
//                    waves (overlap0, overlap1) interleaved WZT.
// TIFF mode writes only the overlaps, as one float TIFF per call (below).
void Reconstructor::open_dumps() {
  if (P_.bTIFF || !in_mrc_) return;
  const mrc::Header& in = in_mrc_->header();
  auto open = [&](const char* path, mrc::Header h) {
    h.next = 0;
    try {
      return std::make_unique<mrc::Writer>(path, h);
    } catch (const IoError&) {
      throw IoError(std::string("Error creating debug output file ") + path);
    }
  };
  if (P_.bSaveAlignedRaw) {
    mrc::Header h = in;
    h.mode = mrc::kFloat;
    h.nx = I_.nx;
    aligned_mrc_ = open(P_.fileRawAligned, h);
  }
  if (P_.bSaveSeparated) {
    mrc::Header h = in;
    h.nx = I_.nx / 2 + 1;
    h.mode = mrc::kComplex;
    sep_mrc_ = open(P_.fileSeparated, h);
  }
  if (P_.bSaveOverlaps) {
    mrc::Header h = in;
    h.nz = I_.nz * 2 * P_.ndirs * I_.ntimes * I_.nwaves;
    h.nx = I_.nx;
    h.num_waves = 2;
    h.interleaved = mrc::kWZT;
    h.mode = mrc::kComplex;
    ovl_mrc_ = open(P_.fileOverlaps, h);
  }
}

// The separated bands of every direction, each [z][y][nx/2+1] complex
// (cudaSirecon.cpp:275-286; there is no z padding, so zoffset is 0).  A
// phase index past the last band (even nphases) gets zero sections.
void Reconstructor::write_separated() {
  if (!sep_mrc_) throw StateError("the separated-bands file is not open");
  const int nb = 2 * P_.norders - 1;
  const size_t plane = (size_t)(I_.nx / 2 + 1) * I_.ny * 2;
  std::vector<float> band(plane * I_.nz0, 0.f);
  for (int d = 0; d < P_.ndirs; d++)
    for (int ph = 0; ph < P_.nphases; ph++) {
      if (ph < nb)
        check(simr_get_band_host(ctx_, d, ph, band.data()));
      else
        std::fill(band.begin(), band.end(), 0.f);
      for (int z = 0; z < I_.nz; z++) sep_mrc_->write_c64(band.data() + plane * z);
    }
}

// The overlap pair of each direction in real space.  With a k0 search, the
// pair findk0 builds: order 0 against order 2 (3D) or 1 (2D, Bessel) at the
// k0 guess (cudaSirecon.cpp:318-352; gpuFunctionsImpl.cu:306-330); MRC:
// sections overlap0 z, overlap1 z for every z; TIFF: one float file of width
// 2*nx (re, im), height ny, 2*nz pages (overlap0's planes, then
// overlap1's), rewritten by every direction.  With a known k0 (no search,
// or a later time point reusing time 0's k0), the order 0 / order 1 pair at
// that k0, MRC only (cudaSirecon.cpp:427-445; the reference steps through
// those buffers by nx*ny floats, half a complex plane, per section; here
// every section is one whole plane).
void Reconstructor::write_overlaps() {
  const bool search = P_.bSearchforvector && !(I_.ntimes > 1 && I_.curTimeIdx > 0 && P_.bUseTime0k0);
  const size_t plane = (size_t)I_.nx * I_.ny * 2;
  const size_t vol = plane * I_.nz0;
  std::vector<float> ov(2 * vol);
  auto put = [&]() {
    if (P_.bTIFF) {
      tiff::write_f32(P_.fileOverlaps, ov.data(), 2 * I_.nx, I_.ny, 2 * I_.nz0);
      return;
    }
    if (!ovl_mrc_) throw StateError("the overlaps file is not open");
    for (int z = 0; z < I_.nz0; z++) {
      ovl_mrc_->write_c64(ov.data() + plane * z);
      ovl_mrc_->write_c64(ov.data() + vol + plane * z);
    }
  };
  for (int d = 0; d < P_.ndirs; d++) {
    simr_fit f;
    check(simr_get_fit(ctx_, d, &f));
    if (search) {
      const int o2 = (I_.nz0 > 1 && !P_.bBessel) ? 2 : 1;
      check(simr_debug_overlap_volumes(ctx_, d, 0, o2, f.k0guess_x, f.k0guess_y, ov.data(), ov.data() + vol));
      put();
    } else {
      for (int o = 1; o < P_.norders; o++) {
        if (P_.bTIFF) {
          std::printf("No overlaps were saved in TIFF mode");
        } else if (o == 1) {
          check(simr_debug_overlap_volumes(ctx_, d, 0, 1, f.k0x, f.k0y, ov.data(), ov.data() + vol));
          put();
        }
      }
    }
  }
}

// saveIntermediateDataForDebugging (src/cudaSirecon/mrc.h:76-89): each dump
// gets its title appended to the labels (IMWrHdr with ntflag 1) and the
// header's min/max/mean set to 0/1/0 (the aligned-raw file keeps the input's).
// Closed once, after the last time point.
void Reconstructor::close_dumps() {
  auto finish = [](std::unique_ptr<mrc::Writer>& w, const char* title, bool range) {
    if (!w) return;
    mrc::Header& h = w->header();
    int n = std::max(0, std::min<int>(h.nlab, 10));
    if (n == 10) {
      for (int l = 0; l < 9; l++) std::memcpy(h.label[l], h.label[l + 1], 80);
      n = 9;
    }
    std::memset(h.label[n], ' ', 80);
    std::memcpy(h.label[n], title, std::min<size_t>(80, std::strlen(title)));
    h.nlab = n + 1;
    if (range) {
      h.amin = 0.f;
      h.amax = 1.f;
      h.amean = 0.f;
    }
    w->close();
    w.reset();
  };
  finish(sep_mrc_, "separated bands of all directions", true);
  finish(aligned_mrc_, "drift-corrected raw images", false);
  finish(ovl_mrc_, "overlaps in real space", true);
}

size_t Reconstructor::result_elems() const {
  return ctx_ ? simr_result_elems(ctx_) : 0;
}

void Reconstructor::get_result(float* out) {
  if (!out) throw ArgError("result pointer is null");
  if (!processed_) throw StateError("no processed volume: call SR_processOneVolume first");
  check(simr_get_result_host(ctx_, out));
}

// makeOutputFilePath (boostfs.cpp:58-72) inserts "_proc" three characters
// before the last of any of the characters ".tif"; that is the extension
// only for lower-case ".tif" names.  Here "_proc" goes before the last '.',
// which gives the same name for "*.tif" and a sane one for ".tiff"/".TIF".
std::string Reconstructor::output_path(const std::string& input) const {
  std::string base = fs::path(input).filename().string();
  const size_t dot = base.find_last_of('.');
  base.insert(dot == std::string::npos ? base.size() : dot, "_proc");
  const std::string out = (fs::path(out_dir_) / base).string();
  std::printf("Output: %s\n", out.c_str());
  return out;
}

// writeResult (cudaSirecon.cpp:1636-1700)
void Reconstructor::write_result(int it, int iw) {
  if (!processed_) throw StateError("no processed volume to write");
  host_out_.resize(simr_result_elems(ctx_));
  check(simr_get_result_host(ctx_, host_out_.data()));
  const int xdim = (int)std::lrint((double)P_.zoomfact * I_.nx), ydim = (int)std::lrint((double)P_.zoomfact * I_.ny);
  const int zdim = P_.z_zoom * I_.nz0;
  if (P_.bTIFF) {
    if (argv_.empty()) throw StateError("SR_writeResult needs a reconstructor made from a command line");
    tiff::write_f32(output_path(tiff_files_.at(it)), host_out_.data(), xdim, ydim, zdim);
  } else {
    if (!out_mrc_) throw StateError("output file already closed");
    // min/max of time point 0, with the reference's else-if scan
    float maxval = -FLT_MAX, minval = FLT_MAX;
    const size_t plane = (size_t)xdim * ydim;
    for (int i = 0; i < I_.nz * P_.z_zoom; ++i) {
      const float* ptr = host_out_.data() + plane * i;
      out_mrc_->write_f32(ptr);
      if (it == 0) {
        for (size_t j = 0; j < plane; ++j) {
          if (ptr[j] > maxval)
            maxval = ptr[j];
          else if (ptr[j] < minval)
            minval = ptr[j];
        }
      }
    }
    if (it == 0 && iw == 0) {
      out_mrc_->header().amin = minval;
      out_mrc_->header().amax = maxval;
    }
  }
  std::printf("Time point %d, wave %d done\n", it, iw);
}

// closeFiles + saveCommandLineToHeader (cudaSirecon.cpp:1703-1713;
// src/cudaSirecon/mrc.h:62-75).  The command line from argv[3] on is the
// title; with --writeTitle it replaces the labels.  IVE's IMWrHdr(..., title
// = header.label, ntflag = 1, ...) then appends the header's first label.
void Reconstructor::close_files() {
  close_dumps();
  if (P_.bTIFF || !out_mrc_) return;
  in_mrc_.reset();
  mrc::Header& h = out_mrc_->header();
  char first[80];
  std::memcpy(first, io_header_.label[0], 80);
  if (P_.bWriteTitle) {
    std::string titles;
    for (size_t i = 3; i < argv_.size(); ++i) titles += argv_[i] + " ";
    const int nlab = std::min<int>(10, (int)(titles.size() / 80 + 1));
    std::memset(h.label, ' ', sizeof(h.label));
    for (int l = 0; l < nlab; l++) {
      const std::string part = titles.size() > (size_t)l * 80 ? titles.substr((size_t)l * 80, 80) : "";
      std::memcpy(h.label[l], part.data(), part.size());
    }
    h.nlab = nlab;
  }
  int n = std::max(0, std::min<int>(h.nlab, 10));
  if (n == 10) {
    for (int l = 0; l < 9; l++) std::memcpy(h.label[l], h.label[l + 1], 80);
    n = 9;
  }
  std::memcpy(h.label[n], first, 80);
  h.nlab = n + 1;
  out_mrc_->close();
  out_mrc_.reset();
}

}  // namespace simr
