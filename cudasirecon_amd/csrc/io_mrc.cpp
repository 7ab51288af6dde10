// MRC / DeltaVision codec (see io_mrc.hpp).  Field offsets follow the
// published Priism header layout; every field is (de)serialised explicitly so
// the in-memory struct layout never matters.
#include "io_mrc.hpp"

#include <cstring>

#include "errors.hpp"

namespace simr {
namespace mrc {

namespace {

constexpr int16_t kStamp = -16224;

template <class T>
T swap_bytes(T v) {
  unsigned char b[sizeof(T)];
  std::memcpy(b, &v, sizeof(T));
  for (size_t i = 0; i < sizeof(T) / 2; i++) std::swap(b[i], b[sizeof(T) - 1 - i]);
  std::memcpy(&v, b, sizeof(T));
  return v;
}

template <class T>
void get(const unsigned char* buf, int off, bool swap, T& v) {
  std::memcpy(&v, buf + off, sizeof(T));
  if (swap) v = swap_bytes(v);
}

template <class T>
void put(unsigned char* buf, int off, const T& v) {
  std::memcpy(buf + off, &v, sizeof(T));
}

// One table drives both directions: (offset, member) pairs.
template <class F>
void fields(Header& h, F&& f) {
  f(0, h.nx); f(4, h.ny); f(8, h.nz); f(12, h.mode);
  f(16, h.nxst); f(20, h.nyst); f(24, h.nzst);
  f(28, h.mx); f(32, h.my); f(36, h.mz);
  f(40, h.xlen); f(44, h.ylen); f(48, h.zlen);
  f(52, h.alpha); f(56, h.beta); f(60, h.gamma);
  f(64, h.mapc); f(68, h.mapr); f(72, h.maps);
  f(76, h.amin); f(80, h.amax); f(84, h.amean);
  f(88, h.ispg); f(92, h.next); f(96, h.dvid); f(98, h.nblank); f(100, h.ntst);
  f(128, h.nint); f(130, h.nreal); f(132, h.nres); f(134, h.nzfact);
  f(136, h.min2); f(140, h.max2); f(144, h.min3); f(148, h.max3); f(152, h.min4); f(156, h.max4);
  f(160, h.file_type); f(162, h.lens); f(164, h.n1); f(166, h.n2); f(168, h.v1); f(170, h.v2);
  f(172, h.min5); f(176, h.max5);
  f(180, h.num_times); f(182, h.interleaved);
  f(184, h.tilt_x); f(188, h.tilt_y); f(192, h.tilt_z);
  f(196, h.num_waves); f(198, h.iwav1); f(200, h.iwav2); f(202, h.iwav3); f(204, h.iwav4); f(206, h.iwav5);
  f(208, h.zorig); f(212, h.xorig); f(216, h.yorig); f(220, h.nlab);
}

Header parse(const unsigned char* buf, bool& swap) {
  int16_t stamp_le;
  std::memcpy(&stamp_le, buf + 96, 2);
  if (stamp_le == kStamp) {
    swap = false;
  } else if (swap_bytes(stamp_le) == kStamp) {
    swap = true;
  } else {
    // plain MRC without the DV stamp: a sane mode value tells the order
    int32_t mode;
    std::memcpy(&mode, buf + 12, 4);
    swap = !(mode >= 0 && mode <= 16);
  }
  Header h;
  fields(h, [&](int off, auto& v) { get(buf, off, swap, v); });
  std::memcpy(h.blank, buf + 104, 24);
  for (int i = 0; i < 10; i++) std::memcpy(h.label[i], buf + 224 + 80 * i, 80);
  if (h.num_waves <= 0) h.num_waves = 1;
  if (h.num_times <= 0) h.num_times = 1;
  return h;
}

void serialise(const Header& hin, unsigned char* buf) {
  Header h = hin;
  std::memset(buf, 0, 1024);
  fields(h, [&](int off, auto& v) { put(buf, off, v); });
  std::memcpy(buf + 104, h.blank, 24);
  for (int i = 0; i < 10; i++) std::memcpy(buf + 224 + 80 * i, h.label[i], 80);
}

}  // namespace

Header::Header() { std::memset(label, ' ', sizeof(label)); }

size_t mode_bytes(int32_t mode) {
  switch (mode) {
    case kByte: return 1;
    case kInt16: return 2;
    case kFloat: return 4;
    case kComplexInt16: return 4;
    case kComplex: return 8;
    case kUint16: return 2;
    default: return 0;
  }
}

int zsections(const Header& h) { return h.nz / (h.num_waves * h.num_times); }

long long section_of(const Header& h, int z, int w, int t) {
  const long long nzz = zsections(h), nw = h.num_waves, nt = h.num_times;
  switch (h.interleaved) {
    case kWZT: return ((long long)t * nzz + z) * nw + w;
    case kZWT: return ((long long)t * nw + w) * nzz + z;
    default: return ((long long)w * nt + t) * nzz + z;  // ZTW
  }
}

Reader::Reader(const std::string& path) : path_(path) {
  f_ = std::fopen(path.c_str(), "rb");
  if (!f_) throw IoError("cannot open MRC file " + path);
  unsigned char buf[1024];
  if (std::fread(buf, 1, 1024, f_) != 1024) {
    std::fclose(f_);
    f_ = nullptr;
    throw FormatError("MRC file shorter than its 1024-byte header: " + path);
  }
  h_ = parse(buf, swap_);
  if (mode_bytes(h_.mode) == 0) {
    std::fclose(f_);
    f_ = nullptr;
    throw FormatError("unsupported MRC pixel mode " + std::to_string(h_.mode) + " in " + path);
  }
  if (h_.nx <= 0 || h_.ny <= 0 || h_.nz <= 0 || h_.next < 0) {
    std::fclose(f_);
    f_ = nullptr;
    throw FormatError("MRC header has invalid dimensions in " + path);
  }
}

Reader::~Reader() {
  if (f_) std::fclose(f_);
}

void Reader::read_raw(long long sec, std::vector<unsigned char>& buf) {
  if (sec < 0 || sec >= h_.nz) throw ArgError("MRC section index out of range in " + path_);
  const size_t nb = (size_t)h_.nx * h_.ny * mode_bytes(h_.mode);
  buf.resize(nb);
  const long long off = 1024LL + h_.next + sec * (long long)nb;
  if (fseeko(f_, (off_t)off, SEEK_SET) != 0 || std::fread(buf.data(), 1, nb, f_) != nb)
    throw IoError("MRC file truncated: " + path_);
}

void Reader::read_f32(long long sec, float* out) {
  std::vector<unsigned char> b;
  read_raw(sec, b);
  const size_t n = (size_t)h_.nx * h_.ny;
  switch (h_.mode) {
    case kByte:
      for (size_t i = 0; i < n; i++) out[i] = (float)b[i];
      break;
    case kInt16:
      for (size_t i = 0; i < n; i++) {
        int16_t v;
        get(b.data(), (int)(2 * i), swap_, v);
        out[i] = (float)v;
      }
      break;
    case kUint16:
      for (size_t i = 0; i < n; i++) {
        uint16_t v;
        get(b.data(), (int)(2 * i), swap_, v);
        out[i] = (float)v;
      }
      break;
    case kFloat:
      for (size_t i = 0; i < n; i++) {
        float v;
        std::memcpy(&v, b.data() + 4 * i, 4);
        out[i] = swap_ ? swap_bytes(v) : v;
      }
      break;
    default:
      throw FormatError("MRC section is complex; a real section was expected in " + path_);
  }
}

void Reader::read_u16(long long sec, uint16_t* out) {
  if (h_.mode != kUint16) throw FormatError("MRC file is not 16-bit unsigned: " + path_);
  std::vector<unsigned char> b;
  read_raw(sec, b);
  const size_t n = (size_t)h_.nx * h_.ny;
  std::memcpy(out, b.data(), 2 * n);
  if (swap_)
    for (size_t i = 0; i < n; i++) out[i] = swap_bytes(out[i]);
}

void Reader::read_c64(long long sec, float* out) {
  std::vector<unsigned char> b;
  read_raw(sec, b);
  const size_t n = (size_t)h_.nx * h_.ny;
  if (h_.mode == kComplex) {
    for (size_t i = 0; i < 2 * n; i++) {
      float v;
      std::memcpy(&v, b.data() + 4 * i, 4);
      out[i] = swap_ ? swap_bytes(v) : v;
    }
  } else if (h_.mode == kComplexInt16) {
    for (size_t i = 0; i < 2 * n; i++) {
      int16_t v;
      get(b.data(), (int)(2 * i), swap_, v);
      out[i] = (float)v;
    }
  } else {
    throw FormatError("MRC section is real; a complex section was expected in " + path_);
  }
}

void Reader::ext_floats(long long sec, std::vector<float>& out) {
  const int nint = h_.nint, nreal = h_.nreal;
  out.assign(nreal, 0.f);
  const long long per = 4LL * (nint + nreal);
  if (per <= 0 || (sec + 1) * per > h_.next) return;
  std::vector<unsigned char> b((size_t)per);
  if (fseeko(f_, (off_t)(1024 + sec * per), SEEK_SET) != 0 || std::fread(b.data(), 1, b.size(), f_) != b.size())
    throw IoError("MRC extended header truncated: " + path_);
  for (int i = 0; i < nreal; i++) get(b.data(), 4 * (nint + i), swap_, out[i]);
}

Writer::Writer(const std::string& path, const Header& h) : h_(h), path_(path) {
  f_ = std::fopen(path.c_str(), "wb");
  if (!f_) throw IoError("cannot create MRC file " + path);
  h_.dvid = kStamp;
  h_.next = 0;
  unsigned char buf[1024];
  serialise(h_, buf);
  if (std::fwrite(buf, 1, 1024, f_) != 1024) throw IoError("cannot write MRC header to " + path);
}

Writer::~Writer() {
  if (f_) std::fclose(f_);
}

void Writer::write_f32(const float* sec) {
  if (!f_) throw StateError("MRC writer already closed: " + path_);
  const size_t n = (size_t)h_.nx * h_.ny;
  if (std::fwrite(sec, sizeof(float), n, f_) != n) throw IoError("cannot write MRC section to " + path_);
}

void Writer::write_c64(const float* sec) {
  if (!f_) throw StateError("MRC writer already closed: " + path_);
  const size_t n = (size_t)h_.nx * h_.ny * 2;
  if (std::fwrite(sec, sizeof(float), n, f_) != n) throw IoError("cannot write MRC section to " + path_);
}

void Writer::close() {
  if (!f_) return;
  unsigned char buf[1024];
  serialise(h_, buf);
  const bool ok = std::fseek(f_, 0, SEEK_SET) == 0 && std::fwrite(buf, 1, 1024, f_) == 1024;
  const bool closed = std::fclose(f_) == 0;
  f_ = nullptr;
  if (!ok || !closed) throw IoError("cannot finish MRC file " + path_);
}

Header read_header(const std::string& path) {
  Reader r(path);
  return r.header();
}

void write_volume(const std::string& path, Header h, const float* data, int nsections) {
  h.mode = kFloat;
  h.nz = nsections;
  Writer w(path, h);
  const size_t n = (size_t)h.nx * h.ny;
  for (int s = 0; s < nsections; s++) w.write_f32(data + n * s);
  w.close();
}

}  // namespace mrc
}  // namespace simr
