// TEST INFRASTRUCTURE ONLY.
//
// CPU oracle: a straight restatement of cudasirecon's per-stack reconstruction
// (src/cudaSirecon/cudaSirecon.cpp and src/cudaSirecon/gpuFunctionsImpl.cu),
// templated on the working precision T (float = "FP32 oracle", double = "FP64
// oracle").  Every CUDA kernel of the reference becomes a plain loop here, in
// the same arithmetic order, with the reference's implicit float->double
// promotions written out.  The cuFFT calls become fft_oracle.hpp.
//
// It is the checker for the HIP product (tests, smoke(), bench.py cpu_baseline
// only) and is never linked into or called by the product.  Build: oracle/Makefile.
//
// Fit modes for T=float:
//   fit64 = 0 : the k0/modamp reductions mirror the reference exactly: per-(x,y)
//               float z-sums, 64x4 block tree in float, host float sum over blocks
//               (gpuFunctionsImpl.cu:1074-1220, gpuFunctionsImpl_hh.cu:162-180).
//   fit64 = 1 : same line search, but the reductions are carried in double
//               (the MI355X engine's default precision plan).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include <omp.h>

#include "../include/simr.h"
#include "fft_oracle.hpp"

using orc::C2;

namespace {

constexpr double kPI = 3.14159265358979323846;
constexpr float SPOTRATIO = 0.1f;         // cudaSireconImpl.h:50
constexpr float K0_WARNING_THRESH = 0.05f;  // cudaSireconImpl.h:54
constexpr int RED_X = 64, RED_Y = 4;      // cudaSireconImpl.h:45-46

template <typename T>
inline T mag2(C2<T> a) { return a.x * a.x + a.y * a.y; }

// float "sqrt"/"cos"/... of the working type
inline float Tsqrt(float v) { return sqrtf(v); }
inline double Tsqrt(double v) { return std::sqrt(v); }
inline float Tcos(float v) { return cosf(v); }
// The line search's angles and candidate vectors: cos, sin and atan2 in double
// rounded once to float, where the reference calls cosf/sinf/atan2f (at most
// one ulp apart, in a small fraction of arguments).  The engine forms them the
// same way, bit-reproducibly on host and device.
inline float fcos1(float v) { return (float)std::cos((double)v); }
inline float fsin1(float v) { return (float)std::sin((double)v); }
inline float fatan2_1(float y, float x) { return (float)std::atan2((double)y, (double)x); }
inline double Tcos(double v) { return std::cos(v); }
inline float Tsin(float v) { return sinf(v); }
inline double Tsin(double v) { return std::sin(v); }
inline float Tasin(float v) { return asinf(v); }
inline double Tasin(double v) { return std::asin(v); }
inline float Tacos(float v) { return acosf(v); }
inline double Tacos(double v) { return std::acos(v); }
inline float Tatan2(float a, float b) { return atan2f(a, b); }
inline double Tatan2(double a, double b) { return std::atan2(a, b); }

struct FitStats {
  int n_angle = 0, n_mag = 0, n_red = 0;
};

template <typename T>
struct Oracle {
  simr_params P{};
  int nx = 0, ny = 0, nz = 0, nz0 = 0, ndirs = 0, nphases = 0, norders = 0;
  int nx_raw = 0;     // raw section width (deskew / crop resample it to nx)
  int xdim = 0;       // (nx/2+1)*2 floats per padded row
  long long H = 0;    // complex elements of one half-complex band
  T dxy = 0, dz = 0;
  int wave = 0;
  T inscale = 0;
  bool fit64 = false;
  int curTimeIdx = 0;

  std::vector<T> sepMatrix, noiseVarFactors;
  std::vector<std::vector<std::vector<C2<T>>>> otf;  // [dirOTF][order][size]
  std::vector<std::vector<std::vector<T>>> bands;    // [dir][phase], 2H floats
  std::vector<C2<T>> k0, k0_time0, k0guess;
  std::vector<std::vector<C2<T>>> amp;
  std::vector<double> sum_dir0_phase0;
  std::vector<T> background, slope;
  std::vector<T> outbuffer;
  std::vector<FitStats> stats;
  std::vector<std::vector<float>> corr;

  // z-reduced overlap products for the current overlap set
  std::vector<C2<T>> Pf;     // per-(x,y) sum_z conj(X) Y in T (reference thread sums)
  std::vector<T> SXf, SYf;
  std::vector<C2<double>> Pd;  // same in double (fit64)
  std::vector<double> SXd, SYd;

  C2<T>* band(int dir, int b) { return reinterpret_cast<C2<T>*>(bands[dir][b].data()); }

  void setup(const simr_params& p, const float* otf_in, bool fit64_) {
    P = p;
    fit64 = fit64_;
    nx = p.nx; ny = p.ny; nz = p.nz; nz0 = nz;
    ndirs = p.ndirs; nphases = p.nphases;
    dxy = (T)p.dxy; dz = (T)p.dz; wave = p.wave;
    xdim = (nx / 2 + 1) * 2;
    H = (long long)(nx / 2 + 1) * ny * nz0;
    // setup_common (cudaSirecon.cpp:1212-1217)
    norders = p.norders_output != 0 ? p.norders_output : p.nphases / 2 + 1;
    if (norders > SIMR_MAX_ORDERS || 2 * norders - 1 != nphases)
      throw std::runtime_error("oracle: need nphases == 2*norders-1 <= 5");
    // OTFs (loadOTFs, cudaSirecon.cpp:1532-1617): caller passes the per-order tables
    const size_t sz = (size_t)p.nzotf * p.nxotf * p.nyotf;
    const int nDirsOTF = p.bOneOTFperAngle ? ndirs : 1;
    otf.assign(nDirsOTF, std::vector<std::vector<C2<T>>>(norders, std::vector<C2<T>>(sz)));
    for (int d = 0; d < nDirsOTF; ++d)
      for (int o = 0; o < norders; ++o)
        for (size_t i = 0; i < sz; ++i) {
          const float* s = otf_in + 2 * ((size_t)(d * norders + o) * sz + i);
          otf[d][o][i] = {(T)s[0], (T)s[1]};
        }
    makematrix();
    noiseVarFactors.assign((size_t)ndirs * norders, (T)1);  // allocSepMatrixAndNoiseVarFactors 725-729
    bands.assign(ndirs, std::vector<std::vector<T>>(nphases, std::vector<T>(2 * H, (T)0)));
    // inscale (cudaSirecon.cpp:1231-1232): int products, then float factors, then 1.0/x in double
    {
      int ip = nx * ny * nz0;
      T f = (T)ip * (T)p.zoomfact;
      f = f * (T)p.zoomfact;
      f = f * (T)p.z_zoom;
      f = f * (T)ndirs;
      inscale = (T)(1.0 / (double)f);
    }
    k0.assign(ndirs, {0, 0});
    k0_time0.assign(ndirs, {0, 0});
    k0guess.assign(ndirs, {0, 0});
    T delta_angle = (T)(kPI / ndirs);
    T k0magguess = (T)(1.0 / (double)(T)p.linespacing);
    if (nz > 1) {
      int nordersIn = nphases / 2 + 1;
      k0magguess = k0magguess / (T)(nordersIn - 1);
    }
    for (int i = 0; i < ndirs; ++i) {
      T ang;
      if (p.nk0angles < ndirs)
        ang = (T)p.k0startangle + (T)i * delta_angle;
      else
        ang = (T)p.k0angles[i];
      k0guess[i].x = (T)((double)k0magguess * std::cos((double)ang));
      k0guess[i].y = (T)((double)k0magguess * std::sin((double)ang));
    }
    sum_dir0_phase0.assign(nz, 0.0);
    amp.assign(ndirs, std::vector<C2<T>>(norders, {0, 0}));
    for (int i = 0; i < ndirs; ++i) amp[i][0] = {(T)1, (T)0};
    nx_raw = p.nx_raw > 0 ? p.nx_raw : nx;
    background.assign((size_t)std::max(nx, nx_raw) * ny, (T)p.constbkgd);  // bgAndSlope 139-161
    slope.assign((size_t)std::max(nx, nx_raw) * ny, (T)1);
    stats.assign(ndirs, FitStats());
    corr.assign(ndirs, std::vector<float>(SIMR_MAX_ORDERS, 0.f));
  }

  // makematrix, equally spaced phases (cudaSirecon.cpp:591-617)
  void makematrix() {
    sepMatrix.assign((size_t)(2 * norders - 1) * nphases, (T)0);
    T phi = (T)(2.0 * kPI / nphases);
    for (int j = 0; j < nphases; ++j) {
      sepMatrix[0 * nphases + j] = (T)1;
      for (int order = 1; order < norders; ++order) {
        T arg = (T)(j * order) * phi;
        sepMatrix[(2 * order - 1) * nphases + j] = (T)std::cos((double)arg);
        sepMatrix[2 * order * nphases + j] = (T)std::sin((double)arg);
      }
    }
  }

  // deskewOneSection (cudaSirecon.cpp:1733-1760): the flat-fielded section
  // `in` (nx_raw wide, float as the reference's CImg) into dst (xdim floats
  // per row).  in[indin + 1] past the end of the section, which the reference
  // reads beyond its buffer, is taken as 0.
  void deskew_section(const std::vector<float>& in, T* dst, int z) const {
    const float factor = P.deskew_factor;
    for (int xout = 0; xout < xdim; ++xout) {
      float xin = (float)xout;
      if (std::fabs(factor) > 0)
        xin = (float)(xout - nx / 2. + P.deskew_shift - factor * (z - nz / 2.) + nx_raw / 2.);
      for (int y = 0; y < ny; ++y) {
        const size_t indout = (size_t)y * xdim + xout;
        if (xin >= 0 && xin < nx_raw) {
          const size_t indin = (size_t)y * nx_raw + (size_t)std::floor(xin);
          const float offset = xin - (float)std::floor(xin);
          const float b = indin + 1 < in.size() ? in[indin + 1] : 0.f;
          dst[indout] = (T)((1 - offset) * in[indin] + offset * b);
        } else {
          dst[indout] = (T)0;
        }
      }
    }
  }

  // loadImageData + load_and_flatfield + deskewOneSection
  // (cudaSirecon.cpp:1354-1435, 1715-1760).  raw: [nImages][ny][nx_raw] floats.
  void load(const float* raw) {
    if (nx_raw != nx || P.deskew_factor != 0.f) {
      std::vector<float> sec((size_t)nx_raw * ny);
      for (int dir = 0; dir < ndirs; ++dir)
        for (int z = 0; z < nz; ++z) {
          int zsec = P.bFastSIM ? (z * ndirs * nphases + dir * nphases) : (dir * nz * nphases + z * nphases);
          for (int phase = 0; phase < nphases; ++phase) {
            const float* src = raw + (size_t)zsec * nx_raw * ny;
            for (size_t e = 0; e < sec.size(); ++e)  // load_and_flatfield in float, raw-width indices
              sec[e] = (((float)src[e] - (float)background[e]) - 0.0f) * (float)slope[e] * (float)inscale;
            deskew_section(sec, bands[dir][phase].data() + (size_t)z * xdim * ny, z);
            ++zsec;
          }
        }
      return;
    }
    for (int dir = 0; dir < ndirs; ++dir) {
      for (int z = 0; z < nz; ++z) {
        int zsec = P.bFastSIM ? (z * ndirs * nphases + dir * nphases) : (dir * nz * nphases + z * nphases);
        for (int phase = 0; phase < nphases; ++phase) {
          const float* src = raw + (size_t)zsec * nx * ny;
          T* dst = bands[dir][phase].data() + (size_t)z * xdim * ny;
          for (int l = 0; l < ny; ++l) {
            for (int k = 0; k < nx; ++k) {
              T v = ((T)src[l * nx + k] - background[l * nx + k] - (T)0) * slope[l * nx + k] * inscale;
              dst[l * xdim + k] = v;
            }
            for (int k = nx; k < xdim; ++k) dst[l * xdim + k] = (T)0;
          }
          ++zsec;
        }
      }
    }
  }

  // rescale (gpuFunctionsImpl.cu:110-195), called per (dir, z) by rescaleDriver 520-540.
  T plane_sum(const T* img) const {
    const int gx = (nx + RED_X - 1) / RED_X, gy = (ny + RED_Y - 1) / RED_Y;
    T total = 0;
    T buf[RED_X * RED_Y];
    for (int by = 0; by < gy; ++by)
      for (int bx = 0; bx < gx; ++bx) {
        for (int ty = 0; ty < RED_Y; ++ty)
          for (int tx = 0; tx < RED_X; ++tx) {
            int k = bx * RED_X + tx, l = by * RED_Y + ty;
            buf[ty * RED_X + tx] = (k < nx && l < ny) ? img[k + l * xdim] : (T)0;
          }
        for (int s = RED_X * RED_Y / 2; s > 0; s >>= 1)
          for (int t = 0; t < s; ++t) buf[t] += buf[t + s];
        total += buf[0];  // host loop over blocks in blockIndex order
      }
    return total;
  }

  void rescale(int it) {
    for (int dir = 0; dir < ndirs; ++dir)
      for (int z = 0; z < nz; ++z) {
        std::vector<T> sum(nphases);
        for (int phase = 0; phase < nphases; ++phase)
          sum[phase] = plane_sum(bands[dir][phase].data() + (size_t)z * xdim * ny);
        if (dir == 0 && !(P.equalizet && it != 0)) sum_dir0_phase0[z] = (double)sum[0];
        T ref = P.equalizez ? (T)sum_dir0_phase0[0] : (T)sum_dir0_phase0[z];
        for (int phase = 0; phase < nphases; ++phase) {
          T ratio = ref / sum[phase];
          T* img = bands[dir][phase].data() + (size_t)z * xdim * ny;
          for (int l = 0; l < ny; ++l)
            for (int k = 0; k < nx; ++k) img[l * xdim + k] *= ratio;
        }
      }
  }

  // apodize_x_kernel / apodize_y_kernel (gpuFunctionsImpl.cu:52-82)
  T apo_fact(int l) const {
    const int napod = P.napodize;
    if (sizeof(T) == sizeof(float)) {
      double a = ((double)l + 0.5) / (double)(float)napod * kPI * 0.5;
      return (T)(1.0 - (double)sinf((float)a));
    }
    return (T)(1.0 - std::sin(((double)l + 0.5) / (double)napod * kPI * 0.5));
  }
  void apodize_plane(T* img) {
    const int napod = P.napodize;
    for (int k = 0; k < nx; ++k) {
      T diff = (T)((double)(img[(ny - 1) * xdim + k] - img[k]) / 2.0);
      for (int l = 0; l < napod; ++l) {
        T fact = apo_fact(l);
        img[l * xdim + k] += diff * fact;
        img[(ny - 1 - l) * xdim + k] -= diff * fact;
      }
    }
    for (int l = 0; l < ny; ++l) {
      T diff = (T)((double)(img[l * xdim + nx - 1] - img[l * xdim]) / 2.0);
      for (int k = 0; k < napod; ++k) {
        T fact = apo_fact(k);
        img[l * xdim + k] += diff * fact;
        img[l * xdim + (nx - 1 - k)] -= diff * fact;
      }
    }
  }

  // transformXYSlice (cudaSirecon.cpp:542-579): in-place R2C of each (ny, nx) plane.
  void fft_xy() {
    orc::FFTPlan<T> px(nx), py(ny);
    const int nxh = nx / 2 + 1;
    for (int dir = 0; dir < ndirs; ++dir)
      for (int ph = 0; ph < nphases; ++ph) {
        T* base = bands[dir][ph].data();
#pragma omp parallel
        {
          std::vector<C2<T>> line(std::max(nx, ny)), work(std::max(nx, ny));
#pragma omp for schedule(static)
          for (long long r = 0; r < (long long)ny * nz; ++r) {
            T* row = base + r * xdim;
            for (int k = 0; k < nx; ++k) line[k] = {row[k], (T)0};
            px.exec(line.data(), work.data(), false);
            C2<T>* crow = reinterpret_cast<C2<T>*>(row);
            for (int k = 0; k < nxh; ++k) crow[k] = line[k];
          }
        }
        C2<T>* c = reinterpret_cast<C2<T>*>(base);
        const long long plane = (long long)nxh * ny;
        orc::fft_lines(py, c, (long long)nxh * nz, nxh,
                       [&](long long l) { return (l / nxh) * plane + (l % nxh); }, false);
      }
  }

  // separate + separate_kernel (gpuFunctionsImpl.cu:214-291): per padded-real float.
  void separate(int dir) {
    const int nb = 2 * norders - 1;
    std::vector<std::vector<T>> out(nb, std::vector<T>(2 * H));
    const long long n = 2 * H;
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < n; ++e)
      for (int i = 0; i < nb; ++i) {
        T result = 0;
        for (int j = 0; j < nphases; ++j) result += sepMatrix[i * nphases + j] * bands[dir][j][e];
        out[i][e] = result;
      }
    for (int i = 0; i < nphases; ++i) bands[dir][i].swap(out[i]);
  }

  // z-FFT (cudaSirecon.cpp:238-257)
  void fft_z(int dir) {
    if (nz <= 1) return;
    orc::FFTPlan<T> pz(nz0);
    const long long stride = (long long)(nx / 2 + 1) * ny;
    for (int i = 0; i < nphases; ++i)
      orc::fft_lines(pz, band(dir, i), stride, stride, [](long long l) { return l; }, false);
  }

  // A full (not rotationally averaged) OTF table [kz][ky][kx] over kx >= 0:
  // trilinear lookup, Hermitian mirror conj(OTF(-k)) for kx < 0, ky and kz
  // wrapping in FFT order, kx clamped at the edge.  The reference's
  // dev_otfinterpolate returns 0 for such tables (gpuFunctionsImpl.cu:1682);
  // this is the engine's extension of its (kr, kz) scheme, parity unpinned
  // against the reference and pinned by its exactness on multilinear tables.
  C2<T> otfinterp3d(const std::vector<C2<T>>& t, T kx, T ky, int kz, T kzscale) const {
    const bool cj = kx < 0;
    if (cj) {
      kx = -kx;
      ky = -ky;
      kz = -kz;
    }
    const int nxo = P.nxotf, nyo = P.nyotf, nzo = P.nzotf;
    const T fx = kx / (T)P.dkrotf;
    T fy = ky / (T)P.dkrotf;
    if (fy < 0) fy += (T)nyo;
    T fz = (T)kz * kzscale;
    if (fz < 0) fz += (T)nzo;
    int ix = (int)std::floor(fx), iy = (int)std::floor(fy), iz = (int)std::floor(fz);
    T ax = fx - (T)ix, ay = fy - (T)iy, az = fz - (T)iz;
    if (ix > nxo - 2) { ix = nxo - 2; ax = (T)1; }
    if (ix < 0) { ix = 0; ax = 0; }
    if (iy > nyo - 1) iy = nyo - 1;
    if (iy < 0) iy = 0;
    if (iz > nzo - 1) iz = nzo - 1;
    if (iz < 0) iz = 0;
    const int iy1 = iy == nyo - 1 ? 0 : iy + 1, iz1 = iz == nzo - 1 ? 0 : iz + 1;
    auto at = [&](int y, int z) {
      const C2<T> a = t[((size_t)z * nyo + y) * nxo + ix], b = t[((size_t)z * nyo + y) * nxo + ix + 1];
      return C2<T>{a.x * (1 - ax) + b.x * ax, a.y * (1 - ax) + b.y * ax};
    };
    const C2<T> v00 = at(iy, iz), v10 = at(iy1, iz), v01 = at(iy, iz1), v11 = at(iy1, iz1);
    const C2<T> w0{v00.x * (1 - ay) + v10.x * ay, v00.y * (1 - ay) + v10.y * ay};
    const C2<T> w1{v01.x * (1 - ay) + v11.x * ay, v01.y * (1 - ay) + v11.y * ay};
    C2<T> v{w0.x * (1 - az) + w1.x * az, w0.y * (1 - az) + w1.y * az};
    if (cj) v.y = -v.y;
    return v;
  }

  // dev_otfinterpolate (gpuFunctionsImpl.cu:1677-1715) for RA tables; full
  // tables go to otfinterp3d.
  // Indices are clamped to the table (the reference reads past it when the
  // support exceeds the table; both oracle and engine clamp).
  C2<T> otfinterp(const std::vector<C2<T>>& t, T kx, T ky, int kz, T kzscale) const {
    C2<T> v{0, 0};
    if (!P.bRadAvgOTF) return otfinterp3d(t, kx, ky, kz, kzscale);
    const int nzotf = P.nzotf, nxotf = P.nxotf;
    T krindex = Tsqrt(kx * kx + ky * ky) / (T)P.dkrotf;
    T kzindex = (T)kz * kzscale;
    if (kzindex < 0) kzindex += (T)nzotf;
    int irindex = (int)std::floor(krindex);
    int izindex = (int)std::floor(kzindex);
    T ar = krindex - (T)irindex;
    T az = kzindex - (T)izindex;
    if (irindex > nxotf - 2) { irindex = nxotf - 2; ar = (T)1; }
    if (irindex < 0) { irindex = 0; ar = 0; }
    if (izindex > nzotf - 1) izindex = nzotf - 1;
    if (izindex < 0) izindex = 0;
    int i00, i01, i10, i11;
    if (izindex == nzotf - 1) {
      i00 = irindex * nzotf + izindex;
      i01 = irindex * nzotf;
      i10 = (irindex + 1) * nzotf + izindex;
      i11 = (irindex + 1) * nzotf;
    } else {
      i00 = irindex * nzotf + izindex;
      i01 = irindex * nzotf + (izindex + 1);
      i10 = (irindex + 1) * nzotf + izindex;
      i11 = (irindex + 1) * nzotf + (izindex + 1);
    }
    v.x = (1 - ar) * (t[i00].x * (1 - az) + t[i01].x * az) + ar * (t[i10].x * (1 - az) + t[i11].x * az);
    v.y = (1 - ar) * (t[i00].y * (1 - az) + t[i01].y * az) + ar * (t[i10].y * (1 - az) + t[i11].y * az);
    return v;
  }

  struct Geo {
    T dkx, dky, dkz, kzscale, rdistcutoff, lambdaem, lambdaexc, alpha;
  };
  Geo geometry() const {
    Geo g;
    g.dkx = (T)1 / ((T)nx * dxy);
    g.dky = (T)1 / ((T)ny * dxy);
    g.dkz = dz > 0 ? (T)1 / ((T)nz * dz) : (T)P.dkzotf;
    g.kzscale = g.dkz / (T)P.dkzotf;
    g.rdistcutoff = (T)((double)(T)P.na * 2.0 / ((double)wave * 0.001));
    if ((double)g.rdistcutoff > 1. / (2. * (double)dxy)) g.rdistcutoff = (T)(1. / (2. * (double)dxy));
    g.lambdaem = (T)((double)((T)wave / (T)P.nimm) / 1.0e3);
    g.lambdaexc = (T)(0.88 * (double)g.lambdaem);
    g.alpha = Tasin((T)P.na / (T)P.nimm);
    return g;
  }

  // makeoverlaps + makeOverlaps0/1Kernel + inverse FFT (gpuFunctionsImpl.cu:497-872),
  // followed by the per-(x,y) z-sums that reductionKernel/aTimesConjBKernel form.
  void makeoverlaps(int dir, int order1, int order2, T k0x, T k0y) {
    const int dirOTF = P.bOneOTFperAngle ? dir : 0;
    Geo g = geometry();
    // 498-506: order-0 weight 5 in 3D, 4 for Bessel-sheet data
    const T order0_2_factor = nz > 1 ? (P.bBessel ? (T)4 : (T)5) : (T)1;
    T zdistcutoff;
    if (!P.bBessel) {  // 530-532
      zdistcutoff = (T)(int)std::ceil(((1.0 - (double)Tcos(g.alpha)) / (double)g.lambdaem) / (double)g.dkz);
    } else {  // 533-539: kzExMax in double then float; the sum in double
      const T kzExMax = (T)(2.0 * (double)P.BesselNA / (double)P.BesselLambdaEx);
      const T k0mag = Tsqrt(k0x * k0x + k0y * k0y);
      const T c = k0mag * (T)order2 / (T)(norders - 1) / kzExMax;
      if (!(c <= (T)1)) throw std::runtime_error("oracle: Bessel kz extent below the pattern frequency");
      const T halfangle = Tacos(c);
      zdistcutoff = (T)std::ceil(((double)(kzExMax * Tsin(halfangle)) + (1.0 - (double)Tcos(g.alpha)) /
                                  (double)g.lambdaem) / (double)g.dkz);
    }
    if (zdistcutoff > (T)(nz / 2)) zdistcutoff = (T)(((nz / 2 - 1) > 0) ? (nz / 2 - 1) : 0);
    // zdistcutoff == nz/2 would make planes +-nz/2 write the same output plane
    if (2 * (int)zdistcutoff + 1 > nz) throw std::runtime_error("oracle: overlap z support wider than the stack");
    const T kx = k0x * (T)(order2 - order1);
    const T ky = k0y * (T)(order2 - order1);
    const T otfcutoff = P.bBessel ? (T)0.01f : (T)P.otfcutoff;  // 553-555
    const T rc = g.rdistcutoff;
    const long long N = (long long)nx * ny * nz;
    std::vector<C2<T>> ov0(N, C2<T>{0, 0}), ov1(N, C2<T>{0, 0});
    const C2<T>* band1re = order1 == 0 ? band(dir, 0) : band(dir, order1 * 2 - 1);
    const C2<T>* band1im = order1 == 0 ? nullptr : band(dir, order1 * 2);
    const C2<T>* band2re = band(dir, order2 * 2 - 1);
    const C2<T>* band2im = band(dir, order2 * 2);
    const auto& otf1t = otf[dirOTF][order1];
    const auto& otf2t = otf[dirOTF][order2];
    const int zd = (int)zdistcutoff;
    const int nxh = nx / 2 + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int zb = 0; zb < 2 * zd + 1; ++zb)
      for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
          int z0 = zb - zd;
          T x1f = (T)j, y1f = (T)i;
          if (x1f > (T)(nx / 2)) x1f -= (T)nx;
          if (y1f > (T)(ny / 2)) y1f -= (T)ny;
          x1f *= g.dkx;
          y1f *= g.dky;
          T rdist1 = Tsqrt(x1f * x1f + y1f * y1f);
          T x12 = x1f - kx, y12 = y1f - ky;
          T rdist12 = Tsqrt(x12 * x12 + y12 * y12);
          T x21 = x1f + kx, y21 = y1f + ky;
          T rdist21 = Tsqrt(x21 * x21 + y21 * y21);
          int iin, jin, conj;
          if (j <= nx / 2) { iin = i; jin = j; conj = 0; }
          else { jin = nx - j; iin = (ny - i) % ny; conj = 1; }
          int zin = conj ? -z0 : z0;
          zin = (zin + nz) % nz;
          long long indin = (long long)zin * nxh * ny + (long long)iin * nxh + jin;
          long long indout = (long long)((z0 + nz) % nz) * nx * ny + (long long)i * nx + j;
          // overlap0 (makeOverlaps0Kernel 677-778)
          if (rdist1 <= rc && (rdist12 <= rc || rdist21 <= rc) && rdist12 <= rc &&
              !(z0 == 0 && P.bNoKz0)) {
            C2<T> otf1 = otfinterp(otf1t, x1f, y1f, z0, g.kzscale);
            if (Tsqrt(otf1.x * otf1.x + otf1.y * otf1.y) > otfcutoff) {
              C2<T> otf12 = otfinterp(otf2t, x12, y12, z0, g.kzscale);
              if (Tsqrt(otf12.x * otf12.x + otf12.y * otf12.y) * order0_2_factor > otfcutoff) {
                C2<T> v1re = band1re[indin];
                C2<T> v1im{0, 0};
                if (order1 > 0) v1im = band1im[indin];
                T root = Tsqrt(otf1.x * otf1.x + otf1.y * otf1.y + otf12.x * otf12.x + otf12.y * otf12.y);
                C2<T> fact{otf12.x / root, otf12.y / root};
                if (conj) { v1re.y = -v1re.y; if (order1 > 0) v1im.y = -v1im.y; }
                T temp = v1re.x * fact.x - v1re.y * fact.y;
                v1re.y = v1re.x * fact.y + v1re.y * fact.x;
                v1re.x = temp;
                if (order1 > 0) {
                  temp = v1im.x * fact.x - v1im.y * fact.y;
                  v1im.y = v1im.x * fact.y + v1im.y * fact.x;
                  v1im.x = temp;
                }
                ov0[indout] = {v1re.x - v1im.y, v1re.y + v1im.x};
              }
            }
          }
          // overlap1 (makeOverlaps1Kernel 780-872)
          if (rdist1 < rc && (rdist12 <= rc || rdist21 <= rc) && rdist21 <= rc &&
              !(z0 == 0 && P.bNoKz0)) {
            C2<T> otf2 = otfinterp(otf2t, x1f, y1f, z0, g.kzscale);
            if (Tsqrt(otf2.x * otf2.x + otf2.y * otf2.y) * order0_2_factor > otfcutoff) {
              C2<T> otf21 = otfinterp(otf1t, x21, y21, z0, g.kzscale);
              if (Tsqrt(otf21.x * otf21.x + otf21.y * otf21.y) > otfcutoff) {
                C2<T> v2re = band2re[indin];
                C2<T> v2im = band2im[indin];
                T root = Tsqrt(otf2.x * otf2.x + otf2.y * otf2.y + otf21.x * otf21.x + otf21.y * otf21.y);
                C2<T> fact{otf21.x / root, otf21.y / root};
                if (conj) { v2re.y = -v2re.y; v2im.y = -v2im.y; }
                T temp = v2re.x * fact.x - v2re.y * fact.y;
                v2re.y = v2re.x * fact.y + v2re.y * fact.x;
                v2re.x = temp;
                temp = v2im.x * fact.x - v2im.y * fact.y;
                v2im.y = v2im.x * fact.y + v2im.y * fact.x;
                v2im.x = temp;
                ov1[indout] = {v2re.x - v2im.y, v2re.y + v2im.x};
              }
            }
          }
        }
    orc::fft3d(ov0.data(), nx, ny, nz, true);
    orc::fft3d(ov1.data(), nx, ny, nz, true);
    if (keep_overlaps) { last_ov0 = ov0; last_ov1 = ov1; }
    // per-thread z sums of reductionKernel (1178-1187)
    const long long nxy = (long long)nx * ny;
    Pf.assign(nxy, {0, 0});
    SXf.assign(nxy, 0);
    SYf.assign(nxy, 0);
    Pd.assign(nxy, {0, 0});
    SXd.assign(nxy, 0);
    SYd.assign(nxy, 0);
#pragma omp parallel for schedule(static)
    for (long long e = 0; e < nxy; ++e) {
      C2<T> xsy{0, 0};
      T sxm = 0, sym = 0;
      C2<double> xsyd{0, 0};
      double sxd = 0, syd = 0;
      for (int z = 0; z < nz; ++z) {
        C2<T> X = ov0[z * nxy + e], Y = ov1[z * nxy + e];
        xsy.x += X.x * Y.x + X.y * Y.y;
        xsy.y += X.x * Y.y - X.y * Y.x;
        sxm += X.x * X.x + X.y * X.y;
        sym += Y.x * Y.x + Y.y * Y.y;
        double xr = X.x, xi = X.y, yr = Y.x, yi = Y.y;
        xsyd.x += xr * yr + xi * yi;
        xsyd.y += xr * yi - xi * yr;
        sxd += xr * xr + xi * xi;
        syd += yr * yr + yi * yi;
      }
      Pf[e] = xsy; SXf[e] = sxm; SYf[e] = sym;
      Pd[e] = xsyd; SXd[e] = sxd; SYd[e] = syd;
    }
  }
  bool keep_overlaps = false;
  std::vector<C2<T>> last_ov0, last_ov1;

  // findpeak / fitparabola (gpuFunctionsImpl.cu:443-495)
  static T fitparabola(T a1, T a2, T a3) {
    T slope = (T)(0.5 * (double)(a3 - a1));
    T curve = (a3 + a1) - 2 * a2;
    if (curve == 0) return 0;
    T peak = -slope / curve;
    if (peak > (T)1.5 || peak < (T)-1.5) return 0;
    return peak;
  }
  static C2<T> findpeak(const std::vector<T>& a, int sizex, int sizey) {
    int xcent = 0, ycent = 0;
    T big = (T)-1e11;
    for (int i = 0; i < sizey; i++)
      for (int j = 0; j < sizex; j++)
        if (a[i * sizex + j] > big) { big = a[i * sizex + j]; ycent = i; xcent = j; }
    T a1, a2, a3;
    a1 = xcent == 0 ? a[ycent * sizex + xcent - 1 + sizex] : a[ycent * sizex + xcent - 1];
    a2 = a[ycent * sizex + xcent];
    a3 = xcent == sizex - 1 ? a[ycent * sizex + xcent + 1 - sizex] : a[ycent * sizex + xcent + 1];
    C2<T> peak;
    peak.x = fitparabola(a1, a2, a3) + (T)xcent;
    a1 = ycent == 0 ? a[(ycent - 1 + sizey) * sizex + xcent] : a[(ycent - 1) * sizex + xcent];
    a2 = a[ycent * sizex + xcent];
    a3 = ycent == sizey - 1 ? a[(ycent + 1 - sizey) * sizex + xcent] : a[(ycent + 1) * sizex + xcent];
    peak.y = fitparabola(a1, a2, a3) + (T)ycent;
    return peak;
  }

  // findk0 316-323, fitk0andmodamps 880-887: Bessel-sheet data fit order 1 in 3D
  int fitorder2() const { return nz > 1 && !P.bBessel ? 2 : 1; }

  // findk0 (gpuFunctionsImpl.cu:306-372)
  void findk0(int dir) {
    makeoverlaps(dir, 0, fitorder2(), k0[dir].x, k0[dir].y);
    const long long nxy = (long long)nx * ny;
    std::vector<C2<T>> cc(nxy);
    for (long long e = 0; e < nxy; ++e) {
      if (fit64 && sizeof(T) == sizeof(float))
        cc[e] = {(T)Pd[e].x, (T)(-Pd[e].y)};
      else
        cc[e] = {Pf[e].x, -Pf[e].y};  // aTimesConjB = conj of the reduction thread sum
    }
    orc::fft3d(cc.data(), nx, ny, 1, false);
    std::vector<T> inten(nxy);
    for (long long e = 0; e < nxy; ++e) inten[e] = cc[e].x * cc[e].x + cc[e].y * cc[e].y;
    Geo g = geometry();
    C2<T> old = k0[dir];
    C2<T> pk = findpeak(inten, nx, ny);
    if (old.x / g.dkx < pk.x - (T)(nx / 2)) pk.x -= (T)nx;
    if (old.x / g.dkx > pk.x + (T)(nx / 2)) pk.x += (T)nx;
    if (old.y / g.dky < pk.y - (T)(ny / 2)) pk.y -= (T)ny;
    if (old.y / g.dky > pk.y + (T)(ny / 2)) pk.y += (T)ny;
    pk.x *= g.dkx / (T)fitorder2();
    pk.y *= g.dky / (T)fitorder2();
    k0[dir] = pk;
  }

  // Reference-exact FP32 reduction: reductionKernel (1150-1220) + cpuReduce.
  void reduce_ref(T k0x, T k0y, C2<T>& XS, T& SX, T& SY) {
    const int gx = (nx + RED_X - 1) / RED_X, gy = (ny + RED_Y - 1) / RED_Y;
    C2<T> rx{0, 0};
    T rs = 0, ry = 0;
    const int NT = RED_X * RED_Y;
    std::vector<C2<T>> bx_(NT);
    std::vector<T> bs(NT), by_(NT);
    for (int by = 0; by < gy; ++by)
      for (int bxi = 0; bxi < gx; ++bxi) {
        for (int ty = 0; ty < RED_Y; ++ty)
          for (int tx = 0; tx < RED_X; ++tx) {
            int i = bxi * RED_X + tx, j = by * RED_Y + ty;
            T inner;
            if (sizeof(T) == sizeof(float)) {
              float a = ((float)i - 0.5f * (float)nx) * (float)k0x;
              float b = ((float)j - 0.5f * (float)ny) * (float)k0y;
              inner = (T)(a + b);
            } else {
              inner = ((T)i - (T)0.5 * (T)nx) * k0x + ((T)j - (T)0.5 * (T)ny) * k0y;
            }
            T angle = (T)(2.0 * kPI * (double)inner * (double)dxy);
            C2<T> e{Tcos(angle), Tsin(angle)};
            C2<T> xsy{0, 0};
            T sxm = 0, sym = 0;
            if (i < nx && j < ny) {
              long long idx = (long long)j * nx + i;
              xsy = Pf[idx];
              sxm = SXf[idx];
              sym = SYf[idx];
            }
            T temp = xsy.x * e.x - xsy.y * e.y;
            xsy.y = xsy.x * e.y + xsy.y * e.x;
            xsy.x = temp;
            int lt = ty * RED_X + tx;
            bx_[lt] = xsy; bs[lt] = sxm; by_[lt] = sym;
          }
        for (int s = NT / 2; s > 0; s >>= 1)
          for (int t = 0; t < s; ++t) {
            bx_[t].x += bx_[t + s].x;
            bx_[t].y += bx_[t + s].y;
            bs[t] += bs[t + s];
            by_[t] += by_[t + s];
          }
        rx.x += bx_[0].x; rx.y += bx_[0].y;
        rs += bs[0];
        ry += by_[0];
      }
    XS = rx; SX = rs; SY = ry;
  }

  // FP64 reduction on the same z-sums (engine default precision plan).
  void reduce_d(double k0x, double k0y, C2<double>& XS, double& SX, double& SY) {
    C2<double> acc{0, 0};
    double sx = 0, sy = 0;
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        long long idx = (long long)j * nx + i;
        double ang = 2.0 * kPI * ((i - 0.5 * nx) * k0x + (j - 0.5 * ny) * k0y) * (double)dxy;
        double c = std::cos(ang), s = std::sin(ang);
        C2<double> p = Pd[idx];
        acc.x += p.x * c - p.y * s;
        acc.y += p.x * s + p.y * c;
        sx += SXd[idx];
        sy += SYd[idx];
      }
    XS = acc; SX = sx; SY = sy;
  }

  // findrealspacemodamp (1074-1148); returns corr coefficient; amp2 (|modamp|^2) in double.
  double realspacemodamp(int dir, int order1, int order2, C2<T> kv, int redo, C2<T>* modamp,
                         double* amp2) {
    if (redo) makeoverlaps(dir, order1, order2, kv.x, kv.y);
    stats[dir].n_red++;
    T k0x = kv.x * (T)(order2 - order1);
    T k0y = kv.y * (T)(order2 - order1);
    if (fit64 || sizeof(T) == sizeof(double)) {
      C2<double> XS;
      double SX, SY;
      reduce_d((double)k0x, (double)k0y, XS, SX, SY);
      double mx = XS.x / SX, my = XS.y / SX;
      *modamp = {(T)mx, (T)my};
      *amp2 = mx * mx + my * my;
      return std::sqrt((XS.x * XS.x + XS.y * XS.y) / (SX * SY));
    }
    C2<T> XS;
    T SX, SY;
    reduce_ref(k0x, k0y, XS, SX, SY);
    C2<T> m{XS.x / SX, XS.y / SX};
    *modamp = m;
    *amp2 = (double)(m.x * m.x + m.y * m.y);
    T cc = (XS.x * XS.x + XS.y * XS.y) / (SX * SY);
    return (double)Tsqrt(cc);
  }

  // getmodamp (1046-1072)
  double getmodamp(int dir, T kangle, T klength, int o1, int o2, C2<T>* modamp, int redo, float* cc = nullptr) {
    C2<T> k1;
    if (sizeof(T) == sizeof(float)) {
      k1.x = (T)((float)klength * fcos1((float)kangle));
      k1.y = (T)((float)klength * fsin1((float)kangle));
    } else {
      k1.x = klength * Tcos(kangle);
      k1.y = klength * Tsin(kangle);
    }
    double a2;
    double c = realspacemodamp(dir, o1, o2, k1, redo, modamp, &a2);
    if (cc) *cc = (float)c;
    // the reference returns a float amp2; in the FP32 mirror keep float rounding
    if (!fit64 && sizeof(T) == sizeof(float)) a2 = (double)(float)a2;
    return a2;
  }

  // fitxyparabola (1022-1044); evaluated in double for the FP64 fit, in T otherwise.
  template <typename U>
  static U fitxyparabola(U x1, U y1, U x2, U y2, U x3, U y3) {
    if (x1 == x2 || x2 == x3 || x3 == x1) return 0;
    U xbar1 = (U)(0.5 * (double)(x1 + x2));
    U xbar2 = (U)(0.5 * (double)(x2 + x3));
    U slope1 = (y2 - y1) / (x2 - x1);
    U slope2 = (y3 - y2) / (x3 - x2);
    U curve = (slope2 - slope1) / (xbar2 - xbar1);
    if (curve == 0) return 0;
    return xbar2 - slope2 / curve;
  }

  // fitk0andmodamps (874-1020).  The step values (angle, mag) are accumulated
  // in T exactly as the reference; the amplitudes are double (exact float values
  // in the FP32 mirror).
  void fitk0andmodamps(int dir) {
    const int fo1 = 0, fo2 = fitorder2();
    C2<T> kv = k0[dir];
    T k0mag = Tsqrt(kv.x * kv.x + kv.y * kv.y);
    T k0angle = sizeof(T) == sizeof(float) ? (T)fatan2_1((float)kv.y, (float)kv.x) : Tatan2(kv.y, kv.x);
    const bool dparab = fit64 || sizeof(T) == sizeof(double);
    auto parab = [&](T x1, double y1, T x2, double y2, T x3, double y3) -> T {
      if (dparab) return (T)fitxyparabola<double>((double)x1, y1, (double)x2, y2, (double)x3, y3);
      return fitxyparabola<T>(x1, (T)y1, x2, (T)y2, x3, (T)y3);
    };
    int redo = P.recalcarrays >= 1;
    T x2 = k0angle;
    C2<T> modamp;
    double amp2 = getmodamp(dir, k0angle, k0mag, fo1, fo2, &modamp, redo);
    stats[dir].n_angle++;
    redo = P.recalcarrays >= 3;
    const T deltaangle = (T)0.001;
    const T deltamag = (T)(0.1 / (double)((T)std::max(nx, ny) * dxy));
    T angle = k0angle + deltaangle;
    T x3 = angle;
    double amp3 = getmodamp(dir, angle, k0mag, fo1, fo2, &modamp, redo);
    stats[dir].n_angle++;
    double amp1 = 0;
    T x1 = 0;
    if (amp3 > amp2) {
      while (amp3 > amp2) {
        amp1 = amp2; x1 = x2; amp2 = amp3; x2 = x3;
        angle += deltaangle;
        x3 = angle;
        amp3 = getmodamp(dir, angle, k0mag, fo1, fo2, &modamp, redo);
        stats[dir].n_angle++;
      }
    } else {
      angle = k0angle;
      double a = amp3; amp3 = amp2; amp2 = a;
      T ax = x3; x3 = x2; x2 = ax;
      while (amp3 > amp2) {
        amp1 = amp2; x1 = x2; amp2 = amp3; x2 = x3;
        angle -= deltaangle;
        x3 = angle;
        amp3 = getmodamp(dir, angle, k0mag, fo1, fo2, &modamp, redo);
        stats[dir].n_angle++;
      }
    }
    angle = parab(x1, amp1, x2, amp2, x3, amp3);

    x2 = k0mag;
    amp2 = getmodamp(dir, angle, k0mag, fo1, fo2, &modamp, redo);
    stats[dir].n_mag++;
    T mag = k0mag + deltamag;
    x3 = mag;
    amp3 = getmodamp(dir, angle, mag, fo1, fo2, &modamp, redo);
    stats[dir].n_mag++;
    if (amp3 > amp2) {
      while (amp3 > amp2) {
        amp1 = amp2; x1 = x2; amp2 = amp3; x2 = x3;
        mag += deltamag;
        x3 = mag;
        amp3 = getmodamp(dir, angle, mag, fo1, fo2, &modamp, redo);
        stats[dir].n_mag++;
      }
    } else {
      mag = k0mag;
      double a = amp3; amp3 = amp2; amp2 = a;
      T ax = x3; x3 = x2; x2 = ax;
      while (amp3 > amp2) {
        amp1 = amp2; x1 = x2; amp2 = amp3; x2 = x3;
        mag -= deltamag;
        x3 = mag;
        amp3 = getmodamp(dir, angle, mag, fo1, fo2, &modamp, redo);
        stats[dir].n_mag++;
      }
    }
    mag = parab(x1, amp1, x2, amp2, x3, amp3);

    redo = P.recalcarrays >= 2;
    float cc;
    getmodamp(dir, angle, mag, fo1, fo2, &modamp, redo, &cc);
    corr[dir][fo2] = cc;
    if (sizeof(T) == sizeof(float)) {
      k0[dir].x = (T)((float)mag * fcos1((float)angle));
      k0[dir].y = (T)((float)mag * fsin1((float)angle));
    } else {
      k0[dir].x = mag * Tcos(angle);
      k0[dir].y = mag * Tsin(angle);
    }
    amp[dir][fo2] = modamp;
    redo = 1;
    if (nz == 1) {
      for (int order = 2; order < norders; ++order) {
        getmodamp(dir, angle, mag, order - 1, order, &modamp, redo, &cc);
        amp[dir][order] = modamp;
        corr[dir][order] = cc;
      }
    } else {
      for (int order = 1; order < norders; ++order)
        if (order != fo2) {
          getmodamp(dir, angle, mag, 0, order, &modamp, redo, &cc);
          amp[dir][order] = modamp;
          corr[dir][order] = cc;
        }
    }
  }

  static C2<T> cmul(C2<T> a, C2<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
  static T cmag(C2<T> a) { return Tsqrt(a.x * a.x + a.y * a.y); }

  // findModulationVectorsAndPhasesForAllDirections, k0 policy part (cudaSirecon.cpp:296-492)
  void modulation(int dir) {
    const int ntimes = P.ntimes;
    const bool reuse = ntimes > 1 && curTimeIdx > 0 && P.bUseTime0k0;
    if (!reuse) k0[dir] = k0guess[dir];
    if (P.bSearchforvector && !reuse) {
      findk0(dir);
      fitk0andmodamps(dir);
      if (curTimeIdx == 0) k0_time0[dir] = k0[dir];
      C2<T> d{k0[dir].x - k0guess[dir].x, k0[dir].y - k0guess[dir].y};
      T dist = Tsqrt(d.x * d.x + d.y * d.y);
      if (ntimes > 1 && curTimeIdx > 0 && dist > (T)(2 * K0_WARNING_THRESH)) {
        k0[dir] = k0_time0[dir];
        for (int order = 1; order < norders; ++order) {
          C2<T> m;
          double a2;
          int o1 = nz > 1 ? 0 : order - 1;
          corr[dir][order] = (float)realspacemodamp(dir, o1, order, k0[dir], 1, &m, &a2);
          amp[dir][order] = m;
        }
      }
    } else {
      for (int order = 1; order < norders; ++order) {
        C2<T> m;
        double a2;
        corr[dir][order] = (float)realspacemodamp(dir, 0, order, k0[dir], 1, &m, &a2);
        amp[dir][order] = m;
      }
    }
    if (nz == 1)
      for (int order = 2; order < norders; ++order) amp[dir][order] = cmul(amp[dir][order], amp[dir][order - 1]);
    if (P.nforceamp > 0 && P.forceamp[0] > 0.0f) {
      for (int order = 1; order < norders; ++order) {
        T a = cmag(amp[dir][order]);
        if (order - 1 < P.nforceamp && a < (T)P.forceamp[order - 1]) {
          T f = (T)P.forceamp[order - 1] / a;
          amp[dir][order].x *= f;
          amp[dir][order].y *= f;
        }
      }
    }
    if (!P.bFitallphases) {
      T base_phase = Tatan2(amp[dir][1].y, amp[dir][1].x);
      for (int order = 2; order < norders; ++order) {
        C2<T> amplitude{cmag(amp[dir][order]), 0};
        T phi = (T)order * base_phase;
        C2<T> e{(T)std::cos((double)phi), (T)std::sin((double)phi)};
        amp[dir][order] = cmul(amplitude, e);
      }
    }
  }

  // ---- filterbands (gpuFunctionsImpl.cu:1222-1675) ----
  static T dev_suppress(T x) {
    T x6 = x * x * x;
    x6 *= x6;
    if (sizeof(T) == sizeof(float)) {
      float den = 1.0f + 20000.0f / ((float)x6 + 20.0f);
      return (T)(1.0 / (double)den);
    }
    return (T)1 / ((T)1 + (T)20000 / (x6 + (T)20));
  }
  static T dev_order0damping(T radius, T zindex, T rlimit, int zlimit) {
    T rf = radius / rlimit;
    T zf = std::fabs(zindex / (T)zlimit);
    return rf * rf + zf * zf * zf;
  }

  struct FilterConst {
    T wiener, kzscale, rdistcutoff, apocutoff, zapocutoff, dkx, dky, min_dkr, suppRadius;
    int zd[SIMR_MAX_ORDERS];
    T ampmag2[SIMR_MAX_ORDERS];
    C2<T> conjamp[SIMR_MAX_ORDERS];
    std::vector<T> ampmag2_all;
  };

  FilterConst filter_consts(int dir) {
    FilterConst fc;
    Geo g = geometry();
    fc.wiener = (T)P.wiener * (T)P.wiener;
    fc.kzscale = g.kzscale;
    T k0mag = Tsqrt(k0[0].x * k0[0].x + k0[0].y * k0[0].y);
    // filterbands 1243-1249 (rdistcutoff expression differs in form from makeoverlaps)
    T lambdaem = (T)((double)((T)wave / (T)P.nimm) / 1000.0);
    T alpha = Tasin((T)P.na / (T)P.nimm);
    T lambdaexc = (T)(0.88 * (double)lambdaem);
    T rdistcutoff = (T)((double)((T)P.na * (T)2) / ((double)wave / 1000.0));
    if ((double)rdistcutoff > 1. / (2. * (double)dxy)) rdistcutoff = (T)(1. / (2. * (double)dxy));
    // filterbands 1254: evaluated in float (unlike makeoverlaps 531, which promotes to double)
    if (!P.bBessel) {
      T zq = ((T)1 - Tcos(alpha)) / lambdaem;
      zq = zq / g.dkz;
      int zd0 = (int)std::ceil(zq);
      fc.zd[0] = zd0;
      fc.zd[norders - 1] = (int)(1.3 * (double)zd0);
      if (norders >= 3)
        for (int o = 1; o < norders - 1; o++) fc.zd[o] = (int)(((T)1 + lambdaem / lambdaexc) * (T)zd0);
    } else {  // 1260-1271: kzExMax and the order-0 sum in the working type, the others in double
      T kzExMax = (T)2 * (T)P.BesselNA / (T)P.BesselLambdaEx;
      fc.zd[0] = (int)std::rint((kzExMax + ((T)1 - Tcos(alpha)) / lambdaem) / g.dkz);
      for (int o = 1; o < norders; o++) {
        T c = k0mag * (T)o / (T)(norders - 1) / kzExMax;
        if (!(c <= (T)1)) throw std::runtime_error("oracle: Bessel kz extent below the pattern frequency");
        T halfangle = Tacos(c);
        fc.zd[o] = (int)std::ceil(((double)(kzExMax * Tsin(halfangle)) + (1.0 - (double)Tcos(alpha)) / (double)lambdaem) /
                                  (double)g.dkz);
      }
    }
    for (int o = 0; o < norders; o++)
      if (fc.zd[o] >= nz / 2) fc.zd[o] = ((nz / 2 - 1) > 0 ? (nz / 2 - 1) : 0);
    fc.rdistcutoff = rdistcutoff;
    fc.apocutoff = rdistcutoff + k0mag * (T)(norders - 1);
    fc.zapocutoff = (norders > 1 && !P.bBessel) ? (T)fc.zd[1] : (T)fc.zd[0];  // 1292-1296
    for (int o = 0; o < norders; o++) {
      fc.ampmag2[o] = amp[dir][o].x * amp[dir][o].x + amp[dir][o].y * amp[dir][o].y;
      fc.conjamp[o] = {amp[dir][o].x, -amp[dir][o].y};
    }
    fc.ampmag2_all.resize((size_t)ndirs * norders);
    for (int d2 = 0; d2 < ndirs; d2++)
      for (int o2 = 0; o2 < norders; o2++)
        fc.ampmag2_all[d2 * norders + o2] = amp[d2][o2].x * amp[d2][o2].x + amp[d2][o2].y * amp[d2][o2].y;
    fc.dkx = (T)(1. / (double)((T)nx * dxy));
    fc.dky = (T)(1. / (double)((T)ny * dxy));
    fc.min_dkr = fc.dkx < fc.dky ? fc.dkx : fc.dky;
    fc.suppRadius = (T)P.suppression_radius * fc.min_dkr;
    return fc;
  }

  // one thread of filterbands_kernel1 (1443-1656)
  void filter_visit(int dir, int order, int x1, int y1, int z0, const FilterConst& fc, C2<T>* bp, C2<T>* bp2) {
    const int dirOTF = P.bOneOTFperAngle ? dir : 0;
    T kx = (T)order * k0[dir].x, ky = (T)order * k0[dir].y;
    int iin, jin, conj;
    if (x1 >= 0) { iin = y1; jin = x1; conj = 0; }
    else { iin = -y1; jin = -x1; conj = 1; }
    if (order == 0 && conj) return;
    if (iin < 0) iin += ny;
    const int nxh = nx / 2 + 1;
    long long xyind = (long long)iin * nxh + jin;
    T x1f = (T)x1 * fc.dkx, y1f = (T)y1 * fc.dky;
    T rdist1 = Tsqrt(x1f * x1f + y1f * y1f);
    if (rdist1 <= fc.rdistcutoff) {
      T xabs = x1f + kx, yabs = y1f + ky;
      T rdistabs = Tsqrt(xabs * xabs + yabs * yabs);
      C2<T> otf1 = otfinterp(otf[dirOTF][order], x1f, y1f, z0, fc.kzscale);
      T weight = otf1.x * otf1.x + otf1.y * otf1.y;
      if (order != 0) weight *= fc.ampmag2[order];
      T dampfact = (T)(1. / (double)noiseVarFactors[dir * norders + order]);
      if (P.bSuppress_singularities && order != 0 && rdist1 <= fc.suppRadius)
        dampfact *= dev_suppress(rdist1 / fc.min_dkr);
      else if (!P.bDampenOrder0 && P.bSuppress_singularities && order == 0 && rdist1 <= fc.suppRadius)
        dampfact *= dev_suppress(rdist1 / fc.min_dkr);
      else if (P.bDampenOrder0 && order == 0)
        dampfact *= dev_order0damping(rdist1, (T)z0, fc.rdistcutoff, fc.zd[0]);
      if (order == 0 && z0 == 0 && P.bNoKz0) dampfact = 0;
      weight *= dampfact;
      T sumweight = weight;
      for (int dir2 = 0; dir2 < ndirs; dir2++) {
        const int d2OTF = P.bOneOTFperAngle ? dir2 : 0;
        (void)d2OTF;
        for (int order2 = -(norders - 1); order2 < norders; order2++) {
          if (dir2 == dir && order2 == order) continue;
          if (!P.bFilteroverlaps && !(order2 == 0 && order == 0)) continue;
          int ao2 = order2 < 0 ? -order2 : order2;
          T amp2mag2 = fc.ampmag2_all[dir2 * norders + ao2];
          T kx2 = (T)order2 * k0[dir2].x, ky2 = (T)order2 * k0[dir2].y;
          T x2 = xabs - kx2, y2 = yabs - ky2;
          T rdist2 = Tsqrt(x2 * x2 + y2 * y2);
          if (rdist2 < fc.rdistcutoff) {
            // const_otfPtrs holds the OTFs of the direction being filtered
            C2<T> otf2 = otfinterp(otf[dirOTF][ao2], x2, y2, z0, fc.kzscale);
            T w = (otf2.x * otf2.x + otf2.y * otf2.y) / noiseVarFactors[dir2 * norders + ao2];
            if (order2 != 0) w *= amp2mag2;
            if (P.bSuppress_singularities && order2 != 0 && rdist2 <= fc.suppRadius)
              w *= dev_suppress(rdist2 / fc.min_dkr);
            else if (!P.bDampenOrder0 && P.bSuppress_singularities && order2 == 0 && rdist2 <= fc.suppRadius)
              w *= dev_suppress(rdist2 / fc.min_dkr);
            else if (P.bDampenOrder0 && order2 == 0)
              w *= dev_order0damping(rdist2, (T)z0, fc.rdistcutoff, fc.zd[0]);
            if (P.bNoKz0 && order2 == 0 && z0 == 0) w = 0;
            sumweight += w;
          }
        }
      }
      sumweight += fc.wiener;
      C2<T> scale{dampfact * otf1.x / sumweight, dampfact * (-otf1.y) / sumweight};
      if (P.apodizeoutput) {
        T rho;
        T zdistabs = (T)std::abs(z0);
        if (fc.zapocutoff > 0 && P.bBessel) {  // 1570-1577: x against rdistcutoff, y against apocutoff
          T rhox = xabs / fc.rdistcutoff, rhoy = yabs / fc.apocutoff, rhoz = zdistabs / fc.zapocutoff;
          rho = Tsqrt(rhox * rhox + rhoy * rhoy + rhoz * rhoz);
        } else if (fc.zapocutoff > 0)
          rho = Tsqrt((rdistabs / fc.apocutoff) * (rdistabs / fc.apocutoff) +
                      (zdistabs / fc.zapocutoff) * (zdistabs / fc.zapocutoff));
        else
          rho = Tsqrt((rdistabs / fc.apocutoff) * (rdistabs / fc.apocutoff));
        if (rho > (T)1) rho = (T)1;
        T apofact = 0;
        if (P.apodizeoutput == 1)
          apofact = Tcos((T)((kPI * 0.5) * (double)rho));
        else if (P.apodizeoutput == 2)
          apofact = (T)1 - rho;
        scale.x *= apofact;
        scale.y *= apofact;
      }
      int z = conj ? -z0 : z0;
      int iz = (z + nz) % nz;
      long long ind = (long long)iz * nxh * ny + xyind;
      if (order == 0) {
        bp[ind] = cmul(bp[ind], scale);
      } else {
        scale = cmul(scale, fc.conjamp[order]);
        C2<T> re = bp[ind], im = bp2[ind];
        if (conj) { re.y = -re.y; im.y = -im.y; }
        C2<T> plus{re.x - im.y, re.y + im.x};
        C2<T> minus{re.x + im.y, re.y - im.x};
        plus = cmul(plus, scale);
        re.x = (T)0.5 * (plus.x + minus.x);
        re.y = (T)0.5 * (plus.y + minus.y);
        im.x = (T)0.5 * (plus.y - minus.y);
        im.y = (T)0.5 * (-plus.x + minus.x);
        if (conj) { re.y = -re.y; im.y = -im.y; }
        bp[ind] = re;
        bp2[ind] = im;
      }
    } else {
      int iz = (z0 + nz) % nz;
      long long ind = (long long)iz * nxh * ny + xyind;
      bp[ind] = {0, 0};
      if (order != 0) bp2[ind] = {0, 0};
    }
  }

  void filterbands(int dir) {
    FilterConst fc = filter_consts(dir);
    last_zd.assign(fc.zd, fc.zd + norders);
    const int nxh = nx / 2 + 1;
    for (int order = 0; order < norders; order++) {
      C2<T>* bp = order == 0 ? band(dir, 0) : band(dir, 2 * order - 1);
      C2<T>* bp2 = order == 0 ? nullptr : band(dir, 2 * order);
      const int zd = fc.zd[order];
      for (int pass = 0; pass < 2; ++pass) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int zb = 0; zb < 2 * zd + 1; ++zb)
          for (int yb = 0; yb < ny; ++yb)
            for (int t = 0; t < nx / 2; ++t) {
              int x1 = t + 1;
              if (pass == 1) x1 -= nx / 2;
              filter_visit(dir, order, x1, yb - ny / 2, zb - zd, fc, bp, bp2);
            }
      }
      // filterbands_kernel3 (1660-1675)
      if ((nz - zd) > (zd + 1)) {
        for (int z0 = zd + 1; z0 < nz - zd; ++z0)
          for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nxh; ++x) {
              long long ind = (long long)z0 * nxh * ny + (long long)y * nxh + x;
              bp[ind] = {0, 0};
              if (order != 0) bp2[ind] = {0, 0};
            }
      }
    }
  }
  std::vector<int> last_zd;

  // ---- assemblerealspacebands (gpuFunctionsImpl.cu:1741-1954) ----
  void assemble(int dir) {
    const int xdim_o = (int)std::rint((double)P.zoomfact * nx);
    const int ydim_o = (int)std::rint((double)P.zoomfact * ny);
    const int zdim_o = P.z_zoom * nz;
    const long long nbig = (long long)xdim_o * ydim_o * zdim_o;
    const long long nxyo = (long long)xdim_o * ydim_o;
    std::vector<C2<T>> big(nbig);
    const T fact = (T)((double)P.explodefact / 0.5);
    const T dxyo = dxy / (T)P.zoomfact;
    std::vector<T> coslut(nxyo), sinlut(nxyo);
    for (int order = 0; order < norders; ++order) {
      std::fill(big.begin(), big.end(), C2<T>{0, 0});
      move(dir, order, big.data(), xdim_o, ydim_o, zdim_o);
      orc::fft3d(big.data(), xdim_o, ydim_o, zdim_o, true);
      if (order == 0) {
#pragma omp parallel for schedule(static)
        for (long long e = 0; e < nbig; ++e) outbuffer[e] += big[e].x;
      } else {
        T k0x = k0[dir].x * (T)order, k0y = k0[dir].y * (T)order;
        const int NX = xdim_o, NY = ydim_o;
        for (int i = 0; i < NY; ++i)
          for (int j = 0; j < NX; ++j) {
            T angle;
            if (sizeof(T) == sizeof(float)) {
              float a = (float)(j - NX / 2) * (float)k0x;
              float b = (float)(i - NY / 2) * (float)k0y;
              float s = a + b;
              angle = (T)((double)(float)fact * kPI * (double)s * (double)(float)dxyo);
            } else {
              angle = fact * (T)kPI * ((T)(j - NX / 2) * k0x + (T)(i - NY / 2) * k0y) * dxyo;
            }
            sinlut[(long long)i * NX + j] = Tsin(angle);
            coslut[(long long)i * NX + j] = Tcos(angle);
          }
#pragma omp parallel for schedule(static)
        for (long long e = 0; e < nbig; ++e) {
          long long xy = e % nxyo;
          if (sizeof(T) == sizeof(float)) {
            double d = (double)big[e].x * 2.0 * (double)coslut[xy] - (double)big[e].y * 2.0 * (double)sinlut[xy];
            outbuffer[e] = (T)((double)outbuffer[e] + d);
          } else {
            outbuffer[e] += big[e].x * 2 * coslut[xy] - big[e].y * 2 * sinlut[xy];
          }
        }
      }
    }
  }

  // move_kernel (1853-1913)
  void move(int dir, int order, C2<T>* out, int xdim_o, int ydim_o, int zdim_o) {
    const C2<T>* in1 = order == 0 ? band(dir, 0) : band(dir, 2 * order - 1);
    const C2<T>* in2 = order == 0 ? band(dir, 0) : band(dir, 2 * order);
    const long long nxy = (long long)(nx / 2 + 1) * ny;
    const long long nxyout = (long long)xdim_o * ydim_o;
    for (int bz = 0; bz < nz; ++bz)
      for (int by = 0; by < ny; ++by)
        for (int bx = 0; bx < nx; ++bx) {
          int x = bx - (nx / 2 - 1);
          int y = by - (ny / 2 - 1);
          int z = nz > 1 ? bz - (nz / 2 - 1) : 0;
          int xout = x < 0 ? x + xdim_o : x;
          int yout = y < 0 ? y + ydim_o : y;
          int zout = z < 0 ? z + zdim_o : z;
          long long indout = (long long)zout * nxyout + (long long)yout * xdim_o + xout;
          int conj = 0;
          if (x < 0) { x = -x; y = -y; z = -z; conj = 1; }
          if (y < 0) y += ny;
          if (z < 0) z += nz;
          long long indin = (long long)z * nxy + (long long)y * (nx / 2 + 1) + x;
          C2<T> val;
          if (order == 0) {
            val = in1[indin];
            if (conj) val.y = -val.y;
          } else {
            C2<T> re = in1[indin], im = in2[indin];
            if (conj) { re.y = -re.y; im.y = -im.y; }
            val = {re.x - im.y, re.y + im.x};
          }
          out[indout] = val;
        }
  }

  // processOneVolume (cudaSirecon.cpp:1264-1321) for the current bands.
  void process_volume() {
    for (int dir = 0; dir < ndirs; ++dir)
      for (int z = 0; z < nz; ++z)
        for (int ph = 0; ph < nphases; ++ph)
          if (P.napodize > 0) apodize_plane(bands[dir][ph].data() + (size_t)z * xdim * ny);
    fft_xy();
    for (int dir = 0; dir < ndirs; ++dir) {
      separate(dir);
      fft_z(dir);
      modulation(dir);
    }
    filter_and_assemble();
  }
  void filter_and_assemble(bool filter = true) {
    const long long nout = (long long)std::rint((double)P.zoomfact * nx) *
                           (long long)std::rint((double)P.zoomfact * ny) * P.z_zoom * nz;
    outbuffer.assign(nout, (T)0);
    for (int dir = 0; dir < ndirs; ++dir) {
      if (filter) filterbands(dir);
      assemble(dir);
    }
  }
};

thread_local std::string g_err;

struct Handle {
  int prec = 32;
  std::unique_ptr<Oracle<float>> f;
  std::unique_ptr<Oracle<double>> d;
};

template <typename Fn>
int guarded(Fn fn) {
  try {
    fn();
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

template <typename T>
void copy_fit(Oracle<T>& o, int dir, simr_fit* out) {
  std::memset(out, 0, sizeof(*out));
  out->k0x = (float)o.k0[dir].x;
  out->k0y = (float)o.k0[dir].y;
  out->k0guess_x = (float)o.k0guess[dir].x;
  out->k0guess_y = (float)o.k0guess[dir].y;
  for (int k = 0; k < o.norders; ++k) {
    out->amp_re[k] = (float)o.amp[dir][k].x;
    out->amp_im[k] = (float)o.amp[dir][k].y;
    out->corr[k] = o.corr[dir][k];
  }
  out->n_angle_evals = o.stats[dir].n_angle;
  out->n_mag_evals = o.stats[dir].n_mag;
  out->n_reductions = o.stats[dir].n_red;
}

template <typename T>
void set_fit(Oracle<T>& o, int dir, const simr_fit* in) {
  o.k0[dir] = {(T)in->k0x, (T)in->k0y};
  for (int k = 0; k < o.norders; ++k) o.amp[dir][k] = {(T)in->amp_re[k], (T)in->amp_im[k]};
}

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

// precision: 32 or 64. fit64: 1 to run the k0/modamp reductions in double (T=float only).
void* orc_create(const simr_params* p, const float* otf, int precision, int fit64) {
  Handle* h = new Handle();
  h->prec = precision;
  int rc = guarded([&] {
    if (precision == 64) {
      h->d.reset(new Oracle<double>());
      h->d->setup(*p, otf, true);
    } else {
      h->f.reset(new Oracle<float>());
      h->f->setup(*p, otf, fit64 != 0);
    }
  });
  if (rc) {
    delete h;
    return nullptr;
  }
  return h;
}

void orc_destroy(void* hp) { delete (Handle*)hp; }

#define DISPATCH(h, expr)                   \
  guarded([&] {                             \
    Handle* H_ = (Handle*)(h);              \
    if (H_->prec == 64) {                   \
      auto& o = *H_->d;                     \
      expr;                                 \
    } else {                                \
      auto& o = *H_->f;                     \
      expr;                                 \
    }                                       \
  })

// raw: [nImages][ny][nx] float
int orc_load(void* h, const float* raw) { return DISPATCH(h, o.load(raw)); }
int orc_rescale(void* h, int it) { return DISPATCH(h, if (o.P.do_rescale) o.rescale(it)); }
int orc_set_time(void* h, int it) { return DISPATCH(h, o.curTimeIdx = it); }
int orc_process_volume(void* h) { return DISPATCH(h, o.process_volume()); }

// Stage-level entry points for stage parity.
int orc_preprocess_bands(void* h) {
  return DISPATCH(h, {
    for (int dir = 0; dir < o.ndirs; ++dir)
      for (int z = 0; z < o.nz; ++z)
        for (int ph = 0; ph < o.nphases; ++ph)
          if (o.P.napodize > 0) o.apodize_plane(o.bands[dir][ph].data() + (size_t)z * o.xdim * o.ny);
    o.fft_xy();
    for (int dir = 0; dir < o.ndirs; ++dir) {
      o.separate(dir);
      o.fft_z(dir);
    }
  });
}
int orc_modulation(void* h, int dir) { return DISPATCH(h, o.modulation(dir)); }
int orc_filter_assemble(void* h) { return DISPATCH(h, o.filter_and_assemble()); }
int orc_filter_dir(void* h, int dir) { return DISPATCH(h, o.filterbands(dir)); }
// assemblerealspacebands of every direction's current bands into a zeroed output (no filtering)
int orc_assemble(void* h) { return DISPATCH(h, o.filter_and_assemble(false)); }

// band: complex float [H] (converted from T)
int orc_get_band(void* h, int dir, int b, float* out) {
  return DISPATCH(h, {
    const auto& v = o.bands[dir][b];
    for (size_t i = 0; i < v.size(); ++i) out[i] = (float)v[i];
  });
}
int orc_set_band(void* h, int dir, int b, const float* in) {
  return DISPATCH(h, {
    auto& v = o.bands[dir][b];
    for (size_t i = 0; i < v.size(); ++i) v[i] = in[i];
  });
}
int orc_get_band_d(void* h, int dir, int b, double* out) {
  return DISPATCH(h, {
    const auto& v = o.bands[dir][b];
    for (size_t i = 0; i < v.size(); ++i) out[i] = (double)v[i];
  });
}
long long orc_result_elems(void* h) {
  long long n = 0;
  DISPATCH(h, n = (long long)o.outbuffer.size());
  return n;
}
int orc_get_result(void* h, float* out) {
  return DISPATCH(h, {
    for (size_t i = 0; i < o.outbuffer.size(); ++i) out[i] = (float)o.outbuffer[i];
  });
}
int orc_get_result_d(void* h, double* out) {
  return DISPATCH(h, {
    for (size_t i = 0; i < o.outbuffer.size(); ++i) out[i] = (double)o.outbuffer[i];
  });
}
int orc_get_fit(void* h, int dir, simr_fit* out) { return DISPATCH(h, copy_fit(o, dir, out)); }
int orc_set_fit(void* h, int dir, const simr_fit* in) { return DISPATCH(h, set_fit(o, dir, in)); }
// Flat-field maps of bgAndSlope (cudaSirecon.cpp:139-161): background and
// slope per raw pixel (max(nx, nx_raw) * ny floats each); NULL restores the
// constant background and unit slope.
int orc_set_flatfield(void* h, const float* bg, const float* slope) {
  return DISPATCH(h, {
    const size_t n = o.background.size();
    for (size_t e = 0; e < n; ++e) {
      o.background[e] = bg ? (decltype(o.inscale))bg[e] : (decltype(o.inscale))o.P.constbkgd;
      o.slope[e] = slope ? (decltype(o.inscale))slope[e] : (decltype(o.inscale))1;
    }
  });
}

// OpenMP threads of the oracle's parallel loops from here on (timing runs).
int orc_set_threads(int n) {
  if (n < 1) return -1;
  omp_set_num_threads(n);
  return 0;
}
int orc_get_threads(void) { return omp_get_max_threads(); }

int orc_get_inscale(void* h, double* out) { return DISPATCH(h, *out = (double)o.inscale); }
int orc_get_zdist(void* h, int* out) {
  return DISPATCH(h, {
    for (size_t i = 0; i < o.last_zd.size(); ++i) out[i] = o.last_zd[i];
  });
}
int orc_get_sepmatrix(void* h, double* out) {
  return DISPATCH(h, {
    for (size_t i = 0; i < o.sepMatrix.size(); ++i) out[i] = (double)o.sepMatrix[i];
  });
}

// Overlaps for (order1, order2) at k0 (1/um); returns the per-(x,y) z-sums
// P = sum_z conj(X) Y (complex double [ny][nx]) and, when ov0/ov1 are non-null,
// the full inverse-transformed overlap volumes (complex float [nz][ny][nx]).
int orc_makeoverlaps(void* h, int dir, int o1, int o2, float k0x, float k0y, double* P, double* sx,
                     double* sy, float* ov0, float* ov1) {
  return DISPATCH(h, {
    o.keep_overlaps = ov0 != nullptr;
    using TT = std::remove_reference_t<decltype(o.k0[0].x)>;
    o.makeoverlaps(dir, o1, o2, (TT)k0x, (TT)k0y);
    for (size_t e = 0; e < o.Pd.size(); ++e) {
      P[2 * e] = o.Pd[e].x;
      P[2 * e + 1] = o.Pd[e].y;
      sx[e] = o.SXd[e];
      sy[e] = o.SYd[e];
    }
    if (ov0) {
      for (size_t e = 0; e < o.last_ov0.size(); ++e) {
        ov0[2 * e] = (float)o.last_ov0[e].x;
        ov0[2 * e + 1] = (float)o.last_ov0[e].y;
        ov1[2 * e] = (float)o.last_ov1[e].x;
        ov1[2 * e + 1] = (float)o.last_ov1[e].y;
      }
    }
    o.keep_overlaps = false;
  });
}

// Full one-shot reconstruction of one time point.
int orc_recon(const simr_params* p, const float* otf, const float* raw, int precision, int fit64,
              float* out, simr_fit* fits) {
  void* h = orc_create(p, otf, precision, fit64);
  if (!h) return -1;
  int rc = orc_load(h, raw);
  if (!rc) rc = orc_rescale(h, 0);
  if (!rc) rc = orc_set_time(h, 0);
  if (!rc) rc = orc_process_volume(h);
  if (!rc && out) rc = orc_get_result(h, out);
  if (!rc && fits)
    for (int d = 0; d < p->ndirs && !rc; ++d) rc = orc_get_fit(h, d, &fits[d]);
  orc_destroy(h);
  return rc;
}

// Peak/parabola helpers, exposed for pinning against closed forms.
float orc_fitparabola(float a1, float a2, float a3) { return Oracle<float>::fitparabola(a1, a2, a3); }
double orc_fitxyparabola(double x1, double y1, double x2, double y2, double x3, double y3) {
  return Oracle<double>::fitxyparabola<double>(x1, y1, x2, y2, x3, y3);
}
int orc_findpeak(const float* a, int nx, int ny, float* px, float* py) {
  std::vector<float> v(a, a + (size_t)nx * ny);
  C2<float> p = Oracle<float>::findpeak(v, nx, ny);
  *px = p.x;
  *py = p.y;
  return 0;
}

// Bare FFT access for pinning the oracle FFT against numpy / brute force.
// data: complex [n2][n1][n0] with n0 fastest; dims 1..3.
int orc_fft3d_f(float* data, int n0, int n1, int n2, int inverse) {
  return guarded([&] { orc::fft3d(reinterpret_cast<C2<float>*>(data), n0, n1, n2, inverse != 0); });
}
int orc_fft3d_d(double* data, int n0, int n1, int n2, int inverse) {
  return guarded([&] { orc::fft3d(reinterpret_cast<C2<double>*>(data), n0, n1, n2, inverse != 0); });
}

}  // extern "C"
