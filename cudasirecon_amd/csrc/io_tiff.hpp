// Multi-page TIFF stacks.
//
// The reference reads raw stacks and OTFs, and writes results, through CImg's
// libtiff binding (src/cudaSirecon/cudaSirecon.cpp:1120-1128,1449-1452,
// 1540-1571,1661-1668).  This codec reads classic and BigTIFF files with one
// sample per pixel, 8/16/32-bit integer or 32/64-bit float pixels, strips or
// tiles, no compression, PackBits, LZW or Deflate (with the horizontal
// differencing predictor), and writes 32-bit float stacks uncompressed.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace simr {
namespace tiff {

struct PageInfo {
  uint32_t width = 0, height = 0;
  uint16_t bits = 0, sample_format = 1, spp = 1, compression = 1, predictor = 1, planar = 1;
  uint32_t rows_per_strip = 0, tile_w = 0, tile_h = 0;
  std::vector<uint64_t> offsets, counts;
  std::string description;
};

class Reader {
 public:
  explicit Reader(const std::string& path);  // parses every directory; throws IoError / FormatError
  int pages() const { return (int)pages_.size(); }
  const PageInfo& page(int i) const { return pages_.at(i); }
  int width() const { return pages_.empty() ? 0 : (int)pages_[0].width; }
  int height() const { return pages_.empty() ? 0 : (int)pages_[0].height; }
  int samples() const { return pages_.empty() ? 0 : (int)pages_[0].spp; }
  bool is_u16() const;
  // One page as float (width*height values, row 0 first).
  void read_f32(int page, float* out) const;
  void read_u16(int page, uint16_t* out) const;

 private:
  void decode(int page, std::vector<unsigned char>& pix) const;  // native-endian pixel bytes
  std::vector<unsigned char> file_;
  bool swap_ = false, big_ = false;
  std::vector<PageInfo> pages_;
  std::string path_;
};

// Writes nz pages of nx*ny float32 (row 0 first), uncompressed, little-endian;
// BigTIFF when the data exceed 2 GiB.
void write_f32(const std::string& path, const float* data, int nx, int ny, int nz);

}  // namespace tiff
}  // namespace simr
