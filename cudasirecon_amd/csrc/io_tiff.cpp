// Multi-page TIFF codec (see io_tiff.hpp).  Written from the TIFF 6.0 and
// BigTIFF specifications.
#include "io_tiff.hpp"

#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "errors.hpp"

namespace simr {
namespace tiff {

namespace {

template <class T>
T bswap(T v) {
  unsigned char b[sizeof(T)];
  std::memcpy(b, &v, sizeof(T));
  std::reverse(b, b + sizeof(T));
  std::memcpy(&v, b, sizeof(T));
  return v;
}

struct Cursor {
  const std::vector<unsigned char>& f;
  bool swap;
  const std::string& path;
  template <class T>
  T at(uint64_t off) const {
    if (off > f.size() || sizeof(T) > f.size() - off)
      throw FormatError("TIFF structure points past the end of " + path);
    T v;
    std::memcpy(&v, f.data() + off, sizeof(T));
    return swap ? bswap(v) : v;
  }
};

size_t type_size(uint16_t type) {
  switch (type) {
    case 1: case 2: case 6: case 7: return 1;   // BYTE ASCII SBYTE UNDEFINED
    case 3: case 8: return 2;                   // SHORT SSHORT
    case 4: case 9: case 11: case 13: return 4; // LONG SLONG FLOAT IFD
    case 5: case 10: case 12: case 16: case 17: case 18: return 8;  // RATIONAL ... LONG8 SLONG8 IFD8
    default: return 0;
  }
}

// Values of one directory entry as unsigned integers.
std::vector<uint64_t> entry_values(const Cursor& c, uint16_t type, uint64_t count, uint64_t value_off) {
  std::vector<uint64_t> v;
  const size_t ts = type_size(type);
  if (ts == 0) return v;
  v.reserve((size_t)std::min<uint64_t>(count, 1u << 24));
  for (uint64_t i = 0; i < count; i++) {
    const uint64_t o = value_off + i * ts;
    switch (type) {
      case 1: case 6: case 7: v.push_back(c.at<uint8_t>(o)); break;
      case 3: case 8: v.push_back(c.at<uint16_t>(o)); break;
      case 4: case 9: case 13: v.push_back(c.at<uint32_t>(o)); break;
      case 16: case 17: case 18: v.push_back(c.at<uint64_t>(o)); break;
      default: v.push_back(0); break;
    }
  }
  return v;
}

// TIFF LZW (MSB-first codes, 9-12 bits, early change).
void lzw_decode(const unsigned char* in, size_t n, std::vector<unsigned char>& out, size_t expect,
                const std::string& path) {
  std::vector<std::vector<unsigned char>> table;
  auto reset = [&]() {
    table.assign(258, {});
    for (int i = 0; i < 256; i++) table[i] = {(unsigned char)i};
  };
  reset();
  int width = 9;
  uint64_t bitpos = 0;
  const uint64_t nbits = (uint64_t)n * 8;
  auto next = [&]() -> int {
    if (bitpos + width > nbits) return 257;
    int code = 0;
    for (int i = 0; i < width; i++) {
      const uint64_t b = bitpos + i;
      code = (code << 1) | ((in[b >> 3] >> (7 - (b & 7))) & 1);
    }
    bitpos += width;
    return code;
  };
  std::vector<unsigned char> prev;
  while (out.size() < expect) {
    const int code = next();
    if (code == 257) break;
    if (code == 256) {
      reset();
      width = 9;
      prev.clear();
      continue;
    }
    std::vector<unsigned char> entry;
    if (code < (int)table.size()) {
      entry = table[code];
      if (!prev.empty()) {
        std::vector<unsigned char> add = prev;
        add.push_back(entry[0]);
        table.push_back(std::move(add));
      }
    } else if (code == (int)table.size() && !prev.empty()) {
      entry = prev;
      entry.push_back(prev[0]);
      table.push_back(entry);
    } else {
      throw FormatError("corrupt LZW data in " + path);
    }
    out.insert(out.end(), entry.begin(), entry.end());
    prev = std::move(entry);
    const size_t ts = table.size();
    if (ts + 1 >= 4096)
      width = 12;
    else if (ts + 1 >= 2048)
      width = 12;
    else if (ts + 1 >= 1024)
      width = 11;
    else if (ts + 1 >= 512)
      width = 10;
  }
}

void packbits_decode(const unsigned char* in, size_t n, std::vector<unsigned char>& out, size_t expect) {
  size_t i = 0;
  while (i < n && out.size() < expect) {
    const int8_t h = (int8_t)in[i++];
    if (h >= 0) {
      const size_t cnt = std::min<size_t>((size_t)h + 1, n - i);
      out.insert(out.end(), in + i, in + i + cnt);
      i += cnt;
    } else if (h != -128) {
      if (i >= n) break;
      out.insert(out.end(), (size_t)(1 - h), in[i++]);
    }
  }
}

void inflate_decode(const unsigned char* in, size_t n, std::vector<unsigned char>& out, size_t expect,
                    const std::string& path) {
  const size_t base = out.size();
  out.resize(base + expect);
  uLongf len = (uLongf)expect;
  const int rc = uncompress(out.data() + base, &len, in, (uLong)n);
  if (rc != Z_OK && rc != Z_BUF_ERROR) throw FormatError("corrupt Deflate data in " + path);
  out.resize(base + len);
}

}  // namespace

Reader::Reader(const std::string& path) : path_(path) {
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) throw IoError("cannot open TIFF file " + path);
  std::fseek(fp, 0, SEEK_END);
  const long long sz = ftello(fp);
  std::fseek(fp, 0, SEEK_SET);
  if (sz < 8) {
    std::fclose(fp);
    throw FormatError("not a TIFF file: " + path);
  }
  file_.resize((size_t)sz);
  const size_t got = std::fread(file_.data(), 1, file_.size(), fp);
  std::fclose(fp);
  if (got != file_.size()) throw IoError("cannot read TIFF file " + path);
  if (file_[0] == 'I' && file_[1] == 'I')
    swap_ = false;
  else if (file_[0] == 'M' && file_[1] == 'M')
    swap_ = true;
  else
    throw FormatError("not a TIFF file: " + path);
  const Cursor c{file_, swap_, path_};
  const uint16_t magic = c.at<uint16_t>(2);
  uint64_t ifd;
  if (magic == 42) {
    big_ = false;
    ifd = c.at<uint32_t>(4);
  } else if (magic == 43) {
    big_ = true;
    ifd = c.at<uint64_t>(8);
  } else {
    throw FormatError("not a TIFF file: " + path);
  }
  std::vector<uint64_t> seen;
  while (ifd != 0) {
    if (std::find(seen.begin(), seen.end(), ifd) != seen.end()) throw FormatError("TIFF directory loop in " + path);
    seen.push_back(ifd);
    const uint64_t nent = big_ ? c.at<uint64_t>(ifd) : c.at<uint16_t>(ifd);
    const uint64_t esz = big_ ? 20 : 12, first = ifd + (big_ ? 8 : 2);
    PageInfo p;
    for (uint64_t e = 0; e < nent; e++) {
      const uint64_t eo = first + e * esz;
      const uint16_t tag = c.at<uint16_t>(eo), type = c.at<uint16_t>(eo + 2);
      const uint64_t count = big_ ? c.at<uint64_t>(eo + 4) : c.at<uint32_t>(eo + 4);
      const size_t ts = type_size(type);
      const uint64_t inline_cap = big_ ? 8 : 4;
      uint64_t voff = eo + (big_ ? 12 : 8);
      // a count whose byte size exceeds the file cannot be valid (and ts * count could wrap)
      if (ts != 0 && count > file_.size() / ts) throw FormatError("TIFF entry count past the end of " + path);
      if (ts * count > inline_cap) voff = big_ ? c.at<uint64_t>(voff) : c.at<uint32_t>(voff);
      if (tag == 270 && type == 2) {
        if (voff <= file_.size() && count <= file_.size() - voff)
          p.description.assign((const char*)file_.data() + voff, (size_t)count);
        continue;
      }
      if (tag != 256 && tag != 257 && tag != 258 && tag != 259 && tag != 273 && tag != 277 && tag != 278 &&
          tag != 279 && tag != 284 && tag != 317 && tag != 322 && tag != 323 && tag != 324 && tag != 325 &&
          tag != 339)
        continue;
      const std::vector<uint64_t> v = entry_values(c, type, count, voff);
      if (v.empty()) continue;
      switch (tag) {
        case 256: p.width = (uint32_t)v[0]; break;
        case 257: p.height = (uint32_t)v[0]; break;
        case 258: p.bits = (uint16_t)v[0]; break;
        case 259: p.compression = (uint16_t)v[0]; break;
        case 273: case 324: p.offsets = v; break;
        case 277: p.spp = (uint16_t)v[0]; break;
        case 278: p.rows_per_strip = (uint32_t)v[0]; break;
        case 279: case 325: p.counts = v; break;
        case 284: p.planar = (uint16_t)v[0]; break;
        case 317: p.predictor = (uint16_t)v[0]; break;
        case 322: p.tile_w = (uint32_t)v[0]; break;
        case 323: p.tile_h = (uint32_t)v[0]; break;
        case 339: p.sample_format = (uint16_t)v[0]; break;
      }
    }
    if (p.bits == 0) p.bits = 1;
    if (p.rows_per_strip == 0 || p.rows_per_strip > p.height) p.rows_per_strip = p.height;
    pages_.push_back(std::move(p));
    const uint64_t nxt = first + nent * esz;
    ifd = big_ ? c.at<uint64_t>(nxt) : c.at<uint32_t>(nxt);
  }
  if (pages_.empty()) throw FormatError("TIFF file has no images: " + path);
  const PageInfo& p0 = pages_[0];
  for (const PageInfo& p : pages_) {
    if (p.width != p0.width || p.height != p0.height)
      throw FormatError("TIFF pages differ in size in " + path);
    if (p.spp != 1) throw UnsupportedError("TIFF with more than one sample per pixel: " + path);
    const bool ok_bits = (p.sample_format == 3 && (p.bits == 32 || p.bits == 64)) ||
                         ((p.sample_format == 1 || p.sample_format == 2) &&
                          (p.bits == 8 || p.bits == 16 || p.bits == 32));
    if (!ok_bits) throw UnsupportedError("TIFF pixel type not supported in " + path);
    if (p.compression != 1 && p.compression != 5 && p.compression != 8 && p.compression != 32946 &&
        p.compression != 32773)
      throw UnsupportedError("TIFF compression " + std::to_string(p.compression) + " not supported in " + path);
    if (p.predictor != 1 && p.predictor != 2)
      throw UnsupportedError("TIFF predictor " + std::to_string(p.predictor) + " not supported in " + path);
    if (p.offsets.empty() || p.offsets.size() != p.counts.size())
      throw FormatError("TIFF strip/tile table missing in " + path);
  }
}

bool Reader::is_u16() const {
  for (const PageInfo& p : pages_)
    if (!(p.bits == 16 && p.sample_format == 1)) return false;
  return true;
}

void Reader::decode(int page, std::vector<unsigned char>& pix) const {
  const PageInfo& p = pages_.at(page);
  const size_t bpp = p.bits / 8;
  const bool tiled = p.tile_w > 0 && p.tile_h > 0;
  const uint32_t bw = tiled ? p.tile_w : p.width;
  const uint32_t bh = tiled ? p.tile_h : p.rows_per_strip;
  const size_t block_bytes = (size_t)bw * bh * bpp;
  const uint32_t tiles_x = tiled ? (p.width + bw - 1) / bw : 1;
  pix.assign((size_t)p.width * p.height * bpp, 0);
  std::vector<unsigned char> blk;
  for (size_t b = 0; b < p.offsets.size(); b++) {
    const uint64_t off = p.offsets[b], cnt = p.counts[b];
    if (off > file_.size() || cnt > file_.size() - off) throw FormatError("TIFF data past the end of " + path_);
    const unsigned char* src = file_.data() + off;
    blk.clear();
    switch (p.compression) {
      case 1: blk.assign(src, src + std::min<uint64_t>(cnt, block_bytes)); break;
      case 5: lzw_decode(src, (size_t)cnt, blk, block_bytes, path_); break;
      case 8: case 32946: inflate_decode(src, (size_t)cnt, blk, block_bytes, path_); break;
      case 32773: packbits_decode(src, (size_t)cnt, blk, block_bytes); break;
    }
    blk.resize(block_bytes, 0);
    // native byte order before undoing the predictor
    if (swap_ && bpp > 1) {
      for (size_t i = 0; i + bpp <= blk.size(); i += bpp) std::reverse(blk.begin() + i, blk.begin() + i + bpp);
    }
    if (p.predictor == 2) {
      for (uint32_t r = 0; r < bh; r++) {
        unsigned char* row = blk.data() + (size_t)r * bw * bpp;
        for (uint32_t x = 1; x < bw; x++) {
          if (bpp == 1) {
            row[x] = (unsigned char)(row[x] + row[x - 1]);
          } else if (bpp == 2) {
            uint16_t a, b2;
            std::memcpy(&a, row + 2 * (x - 1), 2);
            std::memcpy(&b2, row + 2 * x, 2);
            b2 = (uint16_t)(a + b2);
            std::memcpy(row + 2 * x, &b2, 2);
          } else if (bpp == 4) {
            uint32_t a, b2;
            std::memcpy(&a, row + 4 * (x - 1), 4);
            std::memcpy(&b2, row + 4 * x, 4);
            b2 = a + b2;
            std::memcpy(row + 4 * x, &b2, 4);
          }
        }
      }
    }
    uint32_t x0 = 0, y0;
    if (tiled) {
      x0 = (uint32_t)(b % tiles_x) * bw;
      y0 = (uint32_t)(b / tiles_x) * bh;
    } else {
      y0 = (uint32_t)b * bh;
    }
    for (uint32_t r = 0; r < bh && y0 + r < p.height; r++) {
      const uint32_t w = std::min(bw, p.width - x0);
      std::memcpy(pix.data() + ((size_t)(y0 + r) * p.width + x0) * bpp, blk.data() + (size_t)r * bw * bpp,
                  (size_t)w * bpp);
    }
  }
}

void Reader::read_f32(int page, float* out) const {
  std::vector<unsigned char> pix;
  decode(page, pix);
  const PageInfo& p = pages_.at(page);
  const size_t n = (size_t)p.width * p.height;
  for (size_t i = 0; i < n; i++) {
    const unsigned char* q = pix.data() + i * (p.bits / 8);
    float v = 0;
    if (p.sample_format == 3) {
      if (p.bits == 32) {
        std::memcpy(&v, q, 4);
      } else {
        double d;
        std::memcpy(&d, q, 8);
        v = (float)d;
      }
    } else if (p.sample_format == 2) {
      if (p.bits == 8) v = (float)(int8_t)q[0];
      if (p.bits == 16) { int16_t t; std::memcpy(&t, q, 2); v = (float)t; }
      if (p.bits == 32) { int32_t t; std::memcpy(&t, q, 4); v = (float)t; }
    } else {
      if (p.bits == 8) v = (float)q[0];
      if (p.bits == 16) { uint16_t t; std::memcpy(&t, q, 2); v = (float)t; }
      if (p.bits == 32) { uint32_t t; std::memcpy(&t, q, 4); v = (float)t; }
    }
    out[i] = v;
  }
}

void Reader::read_u16(int page, uint16_t* out) const {
  const PageInfo& p = pages_.at(page);
  if (!(p.bits == 16 && p.sample_format == 1)) throw FormatError("TIFF page is not 16-bit unsigned: " + path_);
  std::vector<unsigned char> pix;
  decode(page, pix);
  std::memcpy(out, pix.data(), pix.size());
}

void write_f32(const std::string& path, const float* data, int nx, int ny, int nz) {
  if (nx <= 0 || ny <= 0 || nz <= 0) throw ArgError("TIFF write: empty volume");
  const uint64_t page_bytes = (uint64_t)nx * ny * 4;
  const bool big = page_bytes * nz >= (1ull << 31);
  // Layout (pixel data contiguous, as ImageJ stacks are): header, IFD 0, the
  // description, all pages, then IFDs 1..nz-1.
  char desc[128];
  const uint64_t dlen =
      (uint64_t)std::snprintf(desc, sizeof(desc), "ImageJ=1.11a\nimages=%d\nslices=%d\nloop=false\n", nz, nz) + 1;
  const int nent = 12;
  const uint64_t hdr_bytes = big ? 16 : 8;
  const uint64_t ifd_bytes = big ? (8 + (uint64_t)nent * 20 + 8) : (2 + (uint64_t)nent * 12 + 4);
  const uint64_t desc_off = hdr_bytes + ifd_bytes;
  const uint64_t pix0 = (desc_off + dlen + 7) & ~7ull;
  const uint64_t tail = pix0 + page_bytes * nz;
  auto ifd_off = [&](int z) { return z == 0 ? hdr_bytes : tail + (uint64_t)(z - 1) * ifd_bytes; };

  std::vector<unsigned char> buf;
  auto put = [&](const void* p, size_t n) {
    const unsigned char* c = (const unsigned char*)p;
    buf.insert(buf.end(), c, c + n);
  };
  auto u16 = [&](uint16_t v) { put(&v, 2); };
  auto u32 = [&](uint32_t v) { put(&v, 4); };
  auto u64 = [&](uint64_t v) { put(&v, 8); };
  auto ifd = [&](int z) {
    auto entry = [&](uint16_t tag, uint16_t type, uint64_t count, uint64_t value) {
      u16(tag);
      u16(type);
      if (big) {
        u64(count);
        u64(value);
      } else {
        u32((uint32_t)count);
        if (type == 3 && count == 1) {
          u16((uint16_t)value);
          u16(0);
        } else {
          u32((uint32_t)value);
        }
      }
    };
    if (big) u64(nent); else u16(nent);
    entry(256, 4, 1, (uint32_t)nx);
    entry(257, 4, 1, (uint32_t)ny);
    entry(258, 3, 1, 32);
    entry(259, 3, 1, 1);
    entry(262, 3, 1, 1);
    entry(270, 2, z == 0 ? dlen : 1, z == 0 ? desc_off : 0);
    entry(273, big ? 16 : 4, 1, pix0 + (uint64_t)z * page_bytes);
    entry(277, 3, 1, 1);
    entry(278, 4, 1, (uint32_t)ny);
    entry(279, big ? 16 : 4, 1, page_bytes);
    entry(284, 3, 1, 1);
    entry(339, 3, 1, 3);
    const uint64_t next = (z + 1 < nz) ? ifd_off(z + 1) : 0;
    if (big) u64(next); else u32((uint32_t)next);
  };

  put(big ? "II+\0" : "II*\0", 4);
  if (big) {
    u16(8);
    u16(0);
    u64(hdr_bytes);
  } else {
    u32((uint32_t)hdr_bytes);
  }
  ifd(0);
  put(desc, (size_t)dlen);
  buf.resize((size_t)pix0, 0);

  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw IoError("cannot create TIFF file " + path);
  bool ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  for (int z = 0; ok && z < nz; z++)
    ok = std::fwrite(data + (size_t)z * nx * ny, 1, page_bytes, f) == page_bytes;
  buf.clear();
  for (int z = 1; z < nz; z++) ifd(z);
  if (ok && !buf.empty()) ok = std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  const bool closed = std::fclose(f) == 0;
  if (!ok || !closed) throw IoError("cannot write TIFF file " + path);
}

}  // namespace tiff
}  // namespace simr
