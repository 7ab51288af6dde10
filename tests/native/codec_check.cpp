// Test driver for the C++ file codecs (compiled by tests/test_io.py with g++):
//   codec_check tiff-read  <in.tif> <out.f32>          all pages as float32
//   codec_check tiff-u16   <in.tif> <out.u16>          all pages as uint16
//   codec_check tiff-write <in.f32> nx ny nz <out.tif>
//   codec_check mrc-read   <in.mrc> <out.f32> w t      sections of (wave w, time t) as float32
//   codec_check mrc-c64    <in.mrc> <out.c64>          all sections as complex64
//   codec_check mrc-write  <in.f32> nx ny nz <out.mrc>
// Prints "nx ny nz" of the input on stdout.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "io_mrc.hpp"
#include "io_tiff.hpp"

using namespace simr;

static void dump(const std::string& path, const void* p, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f || std::fwrite(p, 1, bytes, f) != bytes) throw std::runtime_error("write " + path);
  std::fclose(f);
}

static std::vector<float> slurp(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("read " + path);
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<float> v((size_t)n / 4);
  if (std::fread(v.data(), 4, v.size(), f) != v.size()) throw std::runtime_error("read " + path);
  std::fclose(f);
  return v;
}

int main(int argc, char** argv) {
  try {
    const std::string cmd = argc > 1 ? argv[1] : "";
    if (cmd == "tiff-read" || cmd == "tiff-u16") {
      tiff::Reader r(argv[2]);
      const size_t plane = (size_t)r.width() * r.height();
      if (cmd == "tiff-read") {
        std::vector<float> v(plane * r.pages());
        for (int p = 0; p < r.pages(); p++) r.read_f32(p, v.data() + plane * p);
        dump(argv[3], v.data(), v.size() * 4);
      } else {
        std::vector<uint16_t> v(plane * r.pages());
        for (int p = 0; p < r.pages(); p++) r.read_u16(p, v.data() + plane * p);
        dump(argv[3], v.data(), v.size() * 2);
      }
      std::printf("%d %d %d\n", r.width(), r.height(), r.pages());
    } else if (cmd == "tiff-write") {
      const std::vector<float> v = slurp(argv[2]);
      tiff::write_f32(argv[6], v.data(), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]));
      std::printf("%s %s %s\n", argv[3], argv[4], argv[5]);
    } else if (cmd == "mrc-read") {
      mrc::Reader r(argv[2]);
      const mrc::Header& h = r.header();
      const int w = std::atoi(argv[4]), t = std::atoi(argv[5]);
      const int nzz = mrc::zsections(h);
      const size_t plane = (size_t)h.nx * h.ny;
      std::vector<float> v(plane * nzz);
      for (int z = 0; z < nzz; z++) r.read_f32(mrc::section_of(h, z, w, t), v.data() + plane * z);
      dump(argv[3], v.data(), v.size() * 4);
      std::printf("%d %d %d\n", h.nx, h.ny, nzz);
    } else if (cmd == "mrc-c64") {
      mrc::Reader r(argv[2]);
      const mrc::Header& h = r.header();
      const size_t plane = (size_t)h.nx * h.ny;
      std::vector<float> v(2 * plane * h.nz);
      for (int z = 0; z < h.nz; z++) r.read_c64(z, v.data() + 2 * plane * z);
      dump(argv[3], v.data(), v.size() * 4);
      std::printf("%d %d %d\n", h.nx, h.ny, h.nz);
    } else if (cmd == "mrc-write") {
      const std::vector<float> v = slurp(argv[2]);
      mrc::Header h;
      h.nx = std::atoi(argv[3]);
      h.ny = std::atoi(argv[4]);
      h.xlen = 0.04f;
      h.ylen = 0.04f;
      h.zlen = 0.125f;
      h.iwav1 = 515;
      mrc::write_volume(argv[6], h, v.data(), std::atoi(argv[5]));
      std::printf("%s %s %s\n", argv[3], argv[4], argv[5]);
    } else {
      std::fprintf(stderr, "unknown command\n");
      return 2;
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
