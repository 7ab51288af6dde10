// MRC / DeltaVision ("Priism IVE") image files.
//
// The reference reads and writes these through the closed IVE library
// (IMOpen/IMRdHdr/IMPosnZWT/IMRdSec/IMWrSec/IMPutHdr/IMWrHdr, used in
// src/cudaSirecon/mrc.h:18-96 and cudaSirecon.cpp:1323-1335,1465-1617,1636-1699).
// This is an independent codec of the published 1024-byte header layout:
// both byte orders are read (the DV stamp -16224 at byte 96 tells the
// order), files are written little-endian with the stamp, and sections are
// addressed by (z, wave, time) through the header's interleave code.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace simr {
namespace mrc {

enum Mode : int32_t { kByte = 0, kInt16 = 1, kFloat = 2, kComplexInt16 = 3, kComplex = 4, kUint16 = 6 };
enum Interleave : int16_t { kZTW = 0, kWZT = 1, kZWT = 2 };

struct Header {
  int32_t nx = 0, ny = 0, nz = 0, mode = kFloat;
  int32_t nxst = 0, nyst = 0, nzst = 0;
  int32_t mx = 0, my = 0, mz = 0;
  float xlen = 0, ylen = 0, zlen = 0;
  float alpha = 90, beta = 90, gamma = 90;
  int32_t mapc = 1, mapr = 2, maps = 3;
  float amin = 0, amax = 0, amean = 0;
  int32_t ispg = 0, next = 0;
  int16_t dvid = -16224, nblank = 0;
  int32_t ntst = 0;
  char blank[24] = {0};
  int16_t nint = 0, nreal = 0, nres = 0, nzfact = 0;
  float min2 = 0, max2 = 0, min3 = 0, max3 = 0, min4 = 0, max4 = 0;
  int16_t file_type = 0, lens = 0, n1 = 0, n2 = 0, v1 = 0, v2 = 0;
  float min5 = 0, max5 = 0;
  int16_t num_times = 1, interleaved = kZTW;
  float tilt_x = 0, tilt_y = 0, tilt_z = 0;
  int16_t num_waves = 1, iwav1 = 0, iwav2 = 0, iwav3 = 0, iwav4 = 0, iwav5 = 0;
  float zorig = 0, xorig = 0, yorig = 0;
  int32_t nlab = 0;
  char label[10][80];
  Header();
};

size_t mode_bytes(int32_t mode);

// Number of sections per (wave, time) and the section index of (z, w, t).
int zsections(const Header& h);
long long section_of(const Header& h, int z, int w, int t);

class Reader {
 public:
  explicit Reader(const std::string& path);  // throws IoError / FormatError
  ~Reader();
  const Header& header() const { return h_; }
  bool swapped() const { return swap_; }
  // Section `sec` (0-based over all nz sections) as float, or raw uint16 when
  // the file mode is kUint16 (read_u16 throws for other modes).
  void read_f32(long long sec, float* out);
  void read_u16(long long sec, uint16_t* out);
  // Complex sections (mode 4) as interleaved floats.
  void read_c64(long long sec, float* out);
  // Extended-header floats of one section (nint ints then nreal floats per section).
  void ext_floats(long long sec, std::vector<float>& out);

 private:
  void read_raw(long long sec, std::vector<unsigned char>& buf);
  FILE* f_ = nullptr;
  Header h_;
  bool swap_ = false;
  std::string path_;
};

class Writer {
 public:
  Writer(const std::string& path, const Header& h);  // writes the header immediately
  ~Writer();
  Header& header() { return h_; }
  void write_f32(const float* sec);                   // appends one section
  void write_c64(const float* sec);                   // appends one complex section (2*nx*ny floats)
  void close();                                       // rewrites the header and closes

 private:
  FILE* f_ = nullptr;
  Header h_;
  std::string path_;
};

Header read_header(const std::string& path);
void write_volume(const std::string& path, Header h, const float* data, int nsections);

}  // namespace mrc
}  // namespace simr
