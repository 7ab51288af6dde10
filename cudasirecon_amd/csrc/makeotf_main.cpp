// The makeotf executable: PSF stack (TIFF or MRC/DV) -> rotationally averaged
// OTF, computed on the GPU by simr_makeotf (include/simr.h).
//
// Command line, file handling and outputs follow the reference's makeotf
// (src/otf/radialft.cpp:82-434, commandline 1381-1553, outputdata 1037-1151):
//   makeotf psf otf [-angle a] [-ls s] [-na n] [-nimm n] [-fixorigin k1 k2]
//           [-leavekz a b c] [-nphases n] [-beaddiam d] [-background b]
//           [-bgInExtHdr] [-5bands] [-nocompen] [-conj] [-rescale] [-2lenses]
//           [-bessel] [-PIshift] [-ifixkz n] [-ifixkr n] [-wavelength nm]
//           [-xyres um] [-zres um] [-gen_order0 file] [-I2M file]
// The input is read as TIFF when it is one, else as MRC.  MRC input takes the
// pixel sizes and wavelength from its header; TIFF input from -xyres, -zres
// and -wavelength.  The output format follows the input's.  With -I2M the
// MRC input holds one extra I2M section after the phases of every z, whose
// OTF (simr_makeotf_i2m) is written to the -I2M file as an MRC complex
// [kr][kz] section (radialft.cpp:425-430, mrc_file_write mode 2).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "../../include/simr.h"
#include "errors.hpp"
#include "io_mrc.hpp"
#include "io_tiff.hpp"

namespace {

void usage() {
  std::printf(
      "Usage:\n makeotf [[input file] [output file] [options]]\n\n"
      "Options:\n"
      "  -angle k0angle         k0 angle of the pattern in radians (1.57)\n"
      "  -ls line_spacing       line spacing of the pattern in microns (0.2)\n"
      "  -na NA                 detection numerical aperture (1.4)\n"
      "  -nimm n                refractive index of the immersion medium (1.515)\n"
      "  -fixorigin kr1 kr2     extrapolate the order-0 OTF near the origin from kr1..kr2\n"
      "  -leavekz kz1 kz2 kz3   keep these kz pixels of orders 1 and 2; clean up the rest\n"
      "  -nphases n             number of phases (5)\n"
      "  -beaddiam d            bead diameter in microns (0.12)\n"
      "  -background b          camera background (TIFF input; default: estimated from the borders)\n"
      "  -bgInExtHdr            per-section background from the MRC extended header (3rd float)\n"
      "  -5bands                write the re and im bands of every order separately\n"
      "  -nocompen              no bead-size compensation\n"
      "  -conj                  conjugate the OTF\n"
      "  -rescale               normalise every order by its own origin value\n"
      "  -2lenses               I5S PSF\n"
      "  -bessel                coherent Bessel-SIM (no out-of-band cleanup)\n"
      "  -PIshift               add pi to the order-1 phase\n"
      "  -ifixkz n              replace the kr=0 line by the kr=1 line (1)\n"
      "  -ifixkr n              replace kz=0 from this kr on by the mean of its neighbours (0: off)\n"
      "  -wavelength nm         emission wavelength in nm (525; TIFF input)\n"
      "  -xyres um              x-y pixel size (0.106; TIFF input)\n"
      "  -zres um               z step (0.1; TIFF input)\n"
      "  -gen_order0 file       also write the separated order-0 band\n"
      "  -I2M file              the input holds an I2M section after each z's phases (MRC);\n"
      "                         write its OTF to this file\n"
      "  -help or -h            this message\n");
}

bool getfloat(const char* s, float* out) {
  if (!s) return false;
  char* end = nullptr;
  errno = 0;
  const double v = std::strtod(s, &end);
  if (errno || end == s || *end != '\0') return false;
  *out = (float)v;
  return true;
}

bool getint(const char* s, int* out) {
  if (!s) return false;
  char* end = nullptr;
  errno = 0;
  const long v = std::strtol(s, &end, 10);
  if (errno || end == s || *end != '\0') return false;
  *out = (int)v;
  return true;
}

struct Cli {
  simr_otf_params p;
  std::string in, out, order0, i2m_out;
  bool in_given = false, out_given = false, bg_in_ext = false, i2m = false;
  float user_dr = 0, user_dz = 0;
  int wavelength = 525;
};

// commandline (radialft.cpp:1381-1553): options in any order; the first two
// arguments are the file names when they are not options.
bool commandline(int argc, char** argv, Cli& c) {
  int n = 1;
  auto need = [&](int k) { return n + k < argc; };
  while (n < argc) {
    const std::string a = argv[n];
    if (a == "-angle" || a == "-ls" || a == "-beaddiam" || a == "-background" || a == "-na" || a == "-nimm" ||
        a == "-xyres" || a == "-zres") {
      float v;
      if (!need(1) || !getfloat(argv[n + 1], &v)) {
        std::printf("Invalid input for switch %s\n", a.c_str());
        return false;
      }
      if (a == "-angle") c.p.k0angle = v;
      if (a == "-ls") c.p.linespacing = v;
      if (a == "-beaddiam") c.p.bead_diameter = v;
      if (a == "-background") c.p.background = v;
      if (a == "-na") c.p.na = v;
      if (a == "-nimm") c.p.nimm = v;
      if (a == "-xyres") c.user_dr = v;
      if (a == "-zres") c.user_dz = v;
      n += 2;
    } else if (a == "-fixorigin") {
      if (!need(2) || !getint(argv[n + 1], &c.p.interpkr[0]) || !getint(argv[n + 2], &c.p.interpkr[1])) {
        std::printf("Invalid input for switch -fixorigin\n");
        return false;
      }
      n += 3;
    } else if (a == "-leavekz") {
      if (!need(3) || !getint(argv[n + 1], &c.p.leavekz[0]) || !getint(argv[n + 2], &c.p.leavekz[1]) ||
          !getint(argv[n + 3], &c.p.leavekz[2])) {
        std::printf("Invalid input for switch -leavekz\n");
        return false;
      }
      n += 4;
    } else if (a == "-nphases" || a == "-ifixkz" || a == "-ifixkr") {
      int v;
      if (!need(1) || !getint(argv[n + 1], &v)) {
        std::printf("Invalid input for switch %s\n", a.c_str());
        return false;
      }
      if (a == "-nphases") c.p.nphases = v;
      if (a == "-ifixkz") c.p.ifixz = v;
      if (a == "-ifixkr") c.p.ifixr = v;
      n += 2;
    } else if (a == "-wavelength") {
      try {
        if (!need(1)) throw std::invalid_argument("missing value");
        c.wavelength = std::stoi(argv[n + 1]);
      } catch (std::exception& e) {
        std::cout << "Invalid input for switch -wavelength: " << e.what() << std::endl;
        return false;
      }
      n += 2;
    } else if (a == "-2lenses") {
      c.p.twolens = 1, n++;
    } else if (a == "-rescale") {
      c.p.dorescale = 1, n++;
    } else if (a == "-5bands") {
      c.p.five_bands = 1, n++;
    } else if (a == "-conj") {
      c.p.conjugate = 1, n++;
    } else if (a == "-nocompen") {
      c.p.do_compen = 0, n++;
    } else if (a == "-bgInExtHdr") {
      c.bg_in_ext = true, n++;
    } else if (a == "-bessel") {
      c.p.coherent_bsim = 1, n++;
    } else if (a == "-PIshift") {
      c.p.forced_pi_shift = 1, n++;
    } else if (a == "-I2M") {
      if (!need(1)) {
        std::printf("Invalid input for switch -I2M\n");
        return false;
      }
      c.i2m = true;
      c.i2m_out = argv[n + 1];
      n += 2;
    } else if (a == "-gen_order0") {
      if (!need(1)) {
        std::printf("Invalid input for switch -gen_order0\n");
        return false;
      }
      c.order0 = argv[n + 1];
      n += 2;
    } else if (a == "-help" || a == "-h") {
      usage();
      return false;
    } else if (n < 3) {
      if (n == 1) {
        c.in = a;
        c.in_given = true;
      } else {
        c.out = a;
        c.out_given = true;
      }
      n++;
    } else {
      std::printf("Invalid command line option %s\n", a.c_str());
      usage();
      return false;
    }
  }
  return true;
}

bool is_tiff(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  unsigned char b[4] = {0, 0, 0, 0};
  const size_t got = std::fread(b, 1, 4, f);
  std::fclose(f);
  return got == 4 && ((b[0] == 'I' && b[1] == 'I' && (b[2] == 42 || b[2] == 43) && b[3] == 0) ||
                      (b[0] == 'M' && b[1] == 'M' && b[2] == 0 && (b[3] == 42 || b[3] == 43)));
}

int run(int argc, char** argv) {
  Cli c;
  simr_otf_params_default(&c.p);
  if (!commandline(argc, argv, c)) return -1;
  std::printf("ileavekz[0]=%d, ifixz[0]=%d\n", c.p.leavekz[0], c.p.ifixz);
  if (!c.in_given) {
    std::cout << " PSF dataset file name: ";
    std::getline(std::cin, c.in);
  }
  const bool tiff = is_tiff(c.in);
  std::unique_ptr<simr::mrc::Reader> mrc_in;
  std::unique_ptr<simr::tiff::Reader> tiff_in;
  if (tiff) {
    tiff_in = std::make_unique<simr::tiff::Reader>(c.in);
  } else {
    std::cout << "Not a TIFF file; now try MRC\n";
    try {
      mrc_in = std::make_unique<simr::mrc::Reader>(c.in);
    } catch (const std::exception&) {
      std::cerr << "File " << c.in << " does not exist.\n";
      return -1;
    }
  }
  if (!c.out_given) {
    std::cout << "Output OTF file name: ";
    std::getline(std::cin, c.out);
  }
  int nx, ny, nsec;
  if (tiff) {
    nx = tiff_in->width();
    ny = tiff_in->height();
    nsec = tiff_in->pages();
    c.p.dr = c.user_dr > 0 ? c.user_dr : 0.106f;
    c.p.dz = c.user_dz > 0 ? c.user_dz : 0.1f;
    c.p.wavelength = c.wavelength;
  } else {
    const simr::mrc::Header& h = mrc_in->header();
    nx = h.nx;
    ny = h.ny;
    nsec = h.nz;
    c.p.dr = h.ylen;
    c.p.dz = h.zlen;
    c.p.wavelength = h.iwav1;
  }
  if (c.p.nphases < 1) {
    std::fprintf(stderr, "nphases must be positive\n");
    return -1;
  }
  if (c.i2m && tiff) {
    // the reference reads no I2M section from TIFF input (radialft.cpp:268-271)
    std::fprintf(stderr, "-I2M needs MRC input\n");
    return -1;
  }
  if (c.i2m && nx != ny) {
    std::fprintf(stderr, "-I2M needs square sections (its OTF is written with nx/2+1 radial samples)\n");
    return -1;
  }
  // sections per z: the phases, then the I2M section
  const int per_z = c.p.nphases + (c.i2m ? 1 : 0);
  const int nz = nsec / per_z;
  std::printf("nx=%d, ny=%d, nz=%d, dr=%f, dz=%f, wavelength=%d\n", nx, ny, nz, c.p.dr, c.p.dz, c.p.wavelength);
  const int used = nz * c.p.nphases;
  const size_t nxy = (size_t)nx * ny;
  std::vector<float> psf(nxy * used);
  std::vector<float> bg;
  std::vector<float> i2m(c.i2m ? nxy * nz : 0);
  for (int z = 0; z < nz; z++) {
    for (int ph = 0; ph < c.p.nphases; ph++) {
      const int s = z * c.p.nphases + ph;
      if (tiff)
        tiff_in->read_f32(s, psf.data() + nxy * s);
      else
        mrc_in->read_f32((long long)z * per_z + ph, psf.data() + nxy * s);
    }
    if (c.i2m) mrc_in->read_f32((long long)z * per_z + c.p.nphases, i2m.data() + nxy * z);
  }
  const float* bgp = nullptr;
  if (!tiff) {
    // MRC input: the per-section background comes from the extended header
    // with -bgInExtHdr and is 0 otherwise (-background applies to TIFF input).
    // Phase section s takes the extended header of file section s, as the
    // reference's section counter skips the I2M sections (radialft.cpp:244-252).
    bg.assign(used, 0.f);
    if (c.bg_in_ext) {
      std::vector<float> ext;
      for (int s = 0; s < used; s++) {
        mrc_in->ext_floats(s, ext);
        bg[s] = ext.size() > 2 ? ext[2] : 0.f;
      }
    }
    bgp = bg.data();
  }
  std::vector<float> otf(2 * simr_otf_elems(&c.p, nx, ny, used));
  std::vector<float> band0(c.order0.empty() ? 0 : nxy * nz);
  simr_otf_info info;
  if (simr_makeotf(&c.p, psf.data(), nx, ny, used, bgp, otf.data(), band0.empty() ? nullptr : band0.data(), &info) !=
      SIMR_OK) {
    std::fprintf(stderr, "makeotf failed: %s\n", simr_last_error());
    return -1;
  }
  std::printf("Center of mass is (%.3f, %.3f, %.3f)\n\n", info.xc, info.yc, info.zc);
  std::vector<float> otf_i2m;
  simr_otf_info info_i2m;
  if (c.i2m) {
    otf_i2m.assign((size_t)2 * info.nkr * nz, 0.f);
    if (simr_makeotf_i2m(&c.p, i2m.data(), nx, ny, nz, otf_i2m.data(), &info_i2m) != SIMR_OK) {
      std::fprintf(stderr, "makeotf -I2M failed: %s\n", simr_last_error());
      return -1;
    }
    std::printf("I2M psf's center of mass is (%.3f, %.3f, %.3f)\n\n", info_i2m.xc, info_i2m.yc, info_i2m.zc);
  }
  for (int o = 1; o < info.norders && !c.p.five_bands; o++) std::printf("  phi=%f\n", info.phi[o]);

  const int nbo = info.nbands_out, nkr = info.nkr;
  if (!c.order0.empty()) {
    if (tiff) {
      simr::tiff::write_f32(c.order0, band0.data(), nx, ny, nz);
    } else {
      simr::mrc::Header h;
      h.nx = nx;
      h.ny = ny;
      h.mx = nx;
      h.my = ny;
      h.mz = nz;
      h.xlen = c.p.dr;
      h.ylen = c.p.dr;
      h.zlen = c.p.dz;
      h.iwav1 = (int16_t)c.p.wavelength;
      h.interleaved = simr::mrc::kZWT;
      simr::mrc::write_volume(c.order0, h, band0.data(), nz);
    }
  }
  if (tiff) {
    // outputdata (radialft.cpp:1037-1071): pages = bands, rows = kr, 2*nz floats per row
    const int w = nz > 1 ? 2 * nz : 2 * nkr, hgt = nz > 1 ? nkr : 1;
    simr::tiff::write_f32(c.out, otf.data(), w, hgt, nbo);
    std::vector<float> amps((size_t)(w / 2) * hgt * nbo);
    for (size_t i = 0; i < amps.size(); i++) amps[i] = std::sqrt(std::hypot(otf[2 * i], otf[2 * i + 1]));
    std::string ampname = c.out;
    size_t pos = c.out.rfind(".tif");
    if (pos == std::string::npos) pos = c.out.rfind(".TIF");
    if (pos == std::string::npos) {
      std::fprintf(stderr, "output name has no .tif extension; the _OnlyForViewing amplitude file was not written\n");
      return -1;
    }
    ampname.insert(pos, "_OnlyForViewing");
    simr::tiff::write_f32(ampname, amps.data(), w / 2, hgt, nbo);
  } else {
    // outputdata (radialft.cpp:1075-1101)
    simr::mrc::Header h = mrc_in->header();
    h.nx = nz;
    h.ny = nkr;
    h.nz = nbo;
    h.mode = simr::mrc::kComplex;
    h.xlen = info.dkz;
    h.ylen = info.dkr;
    h.zlen = 0;
    h.amin = 0;
    h.amax = 1;
    h.nlab = 1;
    simr::mrc::Writer wr(c.out, h);
    for (int b = 0; b < nbo; b++) wr.write_c64(otf.data() + (size_t)2 * nkr * nz * b);
    wr.close();
  }
  if (c.i2m) {
    // mrc_file_write(..., nz, nx/2+1, 1, dkz, dkx, mode 2, wavelength, I2Mfiles) (radialft.cpp:1556-1633)
    std::printf("Writing output file: %s\n", c.i2m_out.c_str());
    simr::mrc::Header h;
    h.nx = h.mx = nz;
    h.ny = h.my = nkr;
    h.nz = h.mz = 1;
    h.mode = simr::mrc::kComplex;
    h.xlen = h.ylen = info.dkz;
    h.zlen = info.dkr;
    h.alpha = h.beta = h.gamma = 0;
    h.num_waves = 1;
    h.num_times = 1;
    h.iwav1 = (int16_t)c.p.wavelength;
    h.nres = 1;
    h.nzfact = 1;
    h.interleaved = simr::mrc::kZWT;
    h.amin = 0.f;
    h.amax = 1.f;
    h.amean = 0.1f;
    h.nlab = 1;
    std::memcpy(h.label[0], "Processed", 9);
    simr::mrc::Writer wr(c.i2m_out, h);
    wr.write_c64(otf_i2m.data());
    wr.close();
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "makeotf: %s\n", e.what());
    return -1;
  }
}
