// TEST INFRASTRUCTURE ONLY.
//
// CPU oracle of the makeotf path: a straight restatement of the reference's
// src/otf/radialft.cpp (PSF -> rotationally averaged OTF), in its arithmetic
// order and precision (float data, the double expressions the reference
// writes with unqualified C math calls or double literals kept in double).
// The 3D R2C FFT (FFTW in the reference, radialft.cpp:339-348) is the
// oracle's own FFT (fft_oracle.hpp) applied to the real data as a complex
// transform; its rounding differs from FFTW's (parity unpinned below 1e-6).
// Used only by tests/ as the checker of simr_makeotf.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/simr.h"
#include "fft_oracle.hpp"

namespace {

using cf = std::complex<float>;
constexpr double kPI = 3.14159265358979323846;

thread_local std::string g_mk_err;

// fitparabola (radialft.cpp:677-695)
float fitparabola(float a1, float a2, float a3) {
  const float slope = 0.5 * (a3 - a1);
  const float curve = (a3 + a1) - 2 * a2;
  if (curve == 0) return 0.0f;
  const float peak = -slope / curve;
  if (peak > 1.5 || peak < -1.5) return 0.0f;
  return peak;
}

// estimate_background (radialft.cpp:487-503): mean of the pixels within
// border_size of the edges (note the '>' on the far edges).
float estimate_background(const float* img, int nx, int ny, int border) {
  unsigned total = 0;
  double sum = 0.;
  for (int y = 0; y < ny; y++)
    for (int x = 0; x < nx; x++)
      if (y < border || y > ny - border || x < border || x > nx - border) {
        sum += img[(size_t)y * nx + x];
        total++;
      }
  return (float)(sum / total);
}

// apodize (radialft.cpp:702-725) on one section of row pitch nxp
void apodize(int napodize, int nx, int nxp, int ny, float* image) {
  for (int k = 0; k < nx; k++) {
    const float diff = (image[(size_t)(ny - 1) * nxp + k] - image[k]) / 2;
    for (int l = 0; l < napodize; l++) {
      const float fact = 1 - std::sin((((float)l + 0.5) / napodize) * kPI * 0.5);
      image[(size_t)l * nxp + k] += diff * fact;
      image[(size_t)(ny - 1 - l) * nxp + k] -= diff * fact;
    }
  }
  for (int l = 0; l < ny; l++) {
    const float diff = (image[(size_t)l * nxp + nx - 1] - image[(size_t)l * nxp]) / 2;
    for (int k = 0; k < napodize; k++) {
      const float fact = 1 - std::sin((((float)k + 0.5) / napodize) * kPI * 0.5);
      image[(size_t)l * nxp + k] += diff * fact;
      image[(size_t)l * nxp + nx - 1 - k] -= diff * fact;
    }
  }
}

// cosapodize (radialft.cpp:731-744)
void cosapodize(int nx, int nxp, int ny, float* image) {
  for (int k = 0; k < nx; k++) {
    const float xfact = std::sin(kPI * ((float)k + 0.5) / nx);
    for (int l = 0; l < ny; l++) {
      const float yfact = std::sin(kPI * ((float)l + 0.5) / ny);
      image[(size_t)l * nxp + k] *= (xfact * yfact);
    }
  }
}

// sphereFFT (radialft.cpp:1155-1167)
double sphereFFT(float k, float radius) {
  double a, x;
  if (k > 0 || k < 0) {
    x = 2 * kPI * radius * k;
    a = radius / (kPI * k * k) * (std::sin(x) / x - std::cos(x));
  } else {
    a = 4 * kPI * std::pow(radius, 3) / 3;
  }
  return a;
}

// 3D forward transform of a real [nz][ny][nx] volume, kept as the half
// spectrum [nz][ny][nx/2+1] (what FFTW's r2c produces).
void rfft3d(const float* in, int nx, int ny, int nz, cf* out) {
  const size_t n = (size_t)nx * ny * nz;
  std::vector<orc::C2<float>> v(n), work(std::max(nx, std::max(ny, nz)));
  for (size_t i = 0; i < n; i++) v[i] = {in[i], 0.f};
  std::vector<orc::C2<float>> line;
  if (nx > 1) {
    orc::FFTPlan<float> px(nx);
    for (size_t r = 0; r < (size_t)ny * nz; r++) px.exec(&v[r * nx], work.data(), false);
  }
  if (ny > 1) {
    orc::FFTPlan<float> py(ny);
    line.resize(ny);
    for (int z = 0; z < nz; z++)
      for (int x = 0; x < nx; x++) {
        for (int y = 0; y < ny; y++) line[y] = v[((size_t)z * ny + y) * nx + x];
        py.exec(line.data(), work.data(), false);
        for (int y = 0; y < ny; y++) v[((size_t)z * ny + y) * nx + x] = line[y];
      }
  }
  if (nz > 1) {
    orc::FFTPlan<float> pz(nz);
    line.resize(nz);
    for (size_t p = 0; p < (size_t)nx * ny; p++) {
      for (int z = 0; z < nz; z++) line[z] = v[(size_t)z * nx * ny + p];
      pz.exec(line.data(), work.data(), false);
      for (int z = 0; z < nz; z++) v[(size_t)z * nx * ny + p] = line[z];
    }
  }
  const int nxh = nx / 2 + 1;
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nxh; x++) {
        const auto& c = v[((size_t)z * ny + y) * nx + x];
        out[((size_t)z * ny + y) * nxh + x] = cf(c.x, c.y);
      }
}

// shift_center (radialft.cpp:846-874)
void shift_center(cf* bands, int nx, int ny, int nz, float xc, float yc, float zc) {
  const int kycent = ny / 2, kzcent = nz / 2, nxy = (nx / 2 + 1) * ny;
  const float dphiz = 2 * kPI * zc / nz;
  const float dphiy = 2 * kPI * yc / ny;
  const float dphix = 2 * kPI * xc / nx;
  for (int kin = 0; kin < nz; kin++) {
    int kz = kin;
    if (kz > kzcent) kz -= nz;
    const float phi1 = dphiz * kz;
    for (int iin = 0; iin < ny; iin++) {
      int ky = iin;
      if (iin > kycent) ky -= ny;
      const float phi2 = dphiy * ky;
      for (int jin = 0; jin < nx / 2 + 1; jin++) {
        const int kx = jin;
        const size_t indin = (size_t)kin * nxy + (size_t)iin * (nx / 2 + 1) + jin;
        const float phi = phi1 + phi2 + dphix * kx;
        const cf ph((float)std::cos((double)phi), (float)std::sin((double)phi));
        const cf b = bands[indin];
        bands[indin] = cf(b.real() * ph.real() - b.imag() * ph.imag(), b.real() * ph.imag() + b.imag() * ph.real());
      }
    }
  }
}

// beadsize_compensate (radialft.cpp:784-841)
void beadsize_compensate(std::vector<cf*>& bands, float k0angle, float linespacing, float bead_diameter, int norders,
                         int nx, int ny, int nz, float dr, float dz) {
  const float dkx = 1. / (nx * dr);
  const float dky = 1. / (ny * dr);
  const float dkz = 1. / (nz * dz);
  const int kycent = ny / 2, kxcent = nx / 2, kzcent = nz / 2, nxy = (nx / 2 + 1) * ny;
  const float k0mag = 1 / linespacing;
  const float radius = bead_diameter * 0.5;
  const float limit_at_origin = 4 * kPI * radius * radius * radius / 3;
  for (int order = 0; order < norders; order++) {
    cf *bandptr = nullptr, *bandptr1 = nullptr;
    if (order == 0) {
      bandptr = bands[order];
    } else {
      bandptr = bands[2 * order - 1];
      bandptr1 = bands[2 * order];
    }
    // float ratio and magnitude, then the double cos/sin of the unqualified C call
    const float k0x = (float)((double)(((float)order) / (norders - 1) * k0mag) * std::cos((double)k0angle));
    const float k0y = (float)((double)(((float)order) / (norders - 1) * k0mag) * std::sin((double)k0angle));
    for (int kin = 0; kin < nz; kin++) {
      int kout = kin;
      if (kout > kzcent) kout -= nz;
      const float kz = kout * dkz;
      for (int iin = 0; iin < ny; iin++) {
        int iout = iin;
        if (iout > kycent) iout -= ny;
        const float ky = iout * dky + k0y;
        for (int jin = 0; jin < kxcent + 1; jin++) {
          const float kx = jin * dkx + k0x;
          const size_t indin = (size_t)kin * nxy + (size_t)iin * (nx / 2 + 1) + jin;
          float ratio;
          if (!(order == 0 && indin == 0)) {
            const float rho = (float)std::sqrt((double)(kz * kz + ky * ky + kx * kx));
            ratio = sphereFFT(rho, radius) / limit_at_origin;
          } else {
            ratio = 1;
          }
          bandptr[indin] = cf(bandptr[indin].real() / ratio, bandptr[indin].imag() / ratio);
          if (order > 0) bandptr1[indin] = cf(bandptr1[indin].real() / ratio, bandptr1[indin].imag() / ratio);
        }
      }
    }
  }
}

// radialft (radialft.cpp:876-936).  nr = min(nx, ny); the output has nr/2+1
// radial rows of nz values.
void radialft(const cf* band, int nx, int ny, int nz, float dr, cf* avg, int nr) {
  const int kycent = ny / 2, kxcent = nx / 2, kzcent = nz / 2, nxy = (nx / 2 + 1) * ny;
  const int nkr = nr / 2 + 1;
  std::vector<int> count((size_t)nkr * nz, 0);
  std::fill(avg, avg + (size_t)nkr * nz, cf(0, 0));
  const float dkx = 1.f / (nx * dr);
  const float dky = 1.f / (ny * dr);
  const float maxK = .5f / dr;
  const float dkr = 1.f / (nr * dr);
  for (int kin = 0; kin < nz; kin++) {
    for (int iin = 0; iin < ny; iin++) {
      int ky = iin;
      if (iin > kycent) ky -= ny;
      const float ky_f = ky * dky;
      for (int jin = 0; jin < kxcent + 1; jin++) {
        const float kx_f = jin * dkx;
        const float rdist_f = (float)std::sqrt((double)(kx_f * kx_f + ky_f * ky_f));
        if (rdist_f < maxK) {
          const size_t indin = (size_t)kin * nxy + (size_t)iin * (nx / 2 + 1) + jin;
          const unsigned indr = (unsigned)std::rint((double)(rdist_f / dkr));
          if (indr >= (unsigned)nkr) throw std::runtime_error("radial bin out of range");
          const size_t indout = (size_t)indr * nz + kin;
          avg[indout] = cf(avg[indout].real() + band[indin].real(), avg[indout].imag() + band[indin].imag());
          count[indout]++;
        }
      }
    }
  }
  for (size_t i = 0; i < (size_t)nkr * nz; i++)
    if (count[i] > 0) avg[i] = cf(avg[i].real() / (float)count[i], avg[i].imag() / (float)count[i]);
  for (int kx = 0; kx < nkr; kx++) {
    const size_t i0 = (size_t)kx * nz;
    avg[i0] = cf(avg[i0].real(), 0.f);
    for (int kz = 1; kz <= nz / 2; kz++) {
      const size_t i = i0 + kz, ic = i0 + (nz - kz);
      const cf a = avg[i], b = std::conj(avg[ic]);
      avg[i] = cf((a.real() + b.real()) / 2.f, (a.imag() + b.imag()) / 2.f);
      avg[ic] = std::conj(avg[i]);
    }
  }
}

// modify (radialft.cpp:1169-1186)
void modify(cf* otf, int nkr, int nz, int ifixz, int ifixr, int order) {
  if (ifixz)
    for (int kz = 0; kz < nz; kz++)
      if (kz != 0 || order > 0) otf[kz] = otf[nz + kz];
  if (ifixr)
    for (int kx = ifixr; kx < nkr; kx++) {
      const int i0 = kx * nz;
      const cf a = otf[i0 + 1], b = otf[i0 + nz - 1];
      otf[i0] = cf((a.real() + b.real()) / 2.f, (a.imag() + b.imag()) / 2.f);
    }
}

// cleanup (radialft.cpp:938-1034)
void cleanup(cf* otf, int order, int nr, int nz, float dr, float dz, float linespacing, unsigned lamdanm, int icleanup,
             int ileave1_1, int ileave1_2, int ileave2, int twolens, float NA, float NIMM) {
  const float lamda = lamdanm * 0.001;
  const float sinalpha = NA / NIMM;
  const float cosalpha = std::cos(std::asin((double)sinalpha));
  const float krmax = 2 * NA / lamda;
  const float k0mag = 1.0 / linespacing;
  const float dkr = 1. / (nr * dr);
  const float dkz = 1. / (nz * dz);
  const cf zero(0, 0);
  if (!twolens) {
    for (int ix = 0; ix < icleanup; ix++) {
      int kzstart, kzend, jotfshape;
      const float kr = ix * dkr;
      if (kr <= krmax) {
        const double beta = std::asin((double)((NA - kr * lamda) / NIMM));
        const double kzedge = (double)(NIMM / lamda) * (std::cos(beta) - (double)cosalpha);
        if (order == 0) {
          kzstart = (int)std::rint((kzedge / dkz) + 1);
          kzend = nz - kzstart;
          for (int iz = kzstart; iz <= kzend; iz++) otf[ix * nz + iz] = zero;
        } else if (order == 1) {
          jotfshape = (int)std::rint((kzedge / dkz) + 0.999);
          // kept inside the row: the reference writes past it when
          // ileave1_1 - jotfshape >= nz (undefined there)
          kzend = std::min(ileave1_1 - jotfshape, nz - 1);
          if (kzend >= 0)
            for (int iz = 0; iz <= kzend; iz++) {
              if (iz == 0) {
                otf[ix * nz] = zero;
              } else {
                otf[ix * nz + iz] = zero;
                otf[ix * nz + nz - iz] = zero;
              }
            }
          kzstart = ileave1_2 + jotfshape;
          for (int iz = kzstart + 1; iz < nz / 2 + 1; iz++) {
            otf[ix * nz + iz] = zero;
            otf[ix * nz + nz - iz] = zero;
          }
        } else {
          jotfshape = (int)std::rint((kzedge / dkz) + 0.999);
          kzstart = ileave2 + jotfshape;
          for (int iz = kzstart; iz < nz / 2 + 1; iz++) {
            otf[ix * nz + iz] = zero;
            otf[ix * nz + nz - iz] = zero;
          }
        }
      } else {
        for (int iz = 0; iz < nz; iz++) otf[ix * nz + iz] = zero;
      }
    }
  } else {
    const float lambdaem = lamda / NIMM;
    const float lambdaexc = 0.88 * lambdaem;
    const float two_over_lambdaem = 2 / lambdaem;
    const float two_over_lambdaexc = 2 / lambdaexc;
    const float beta = std::asin((double)(k0mag / two_over_lambdaexc));
    int center_of_arc;
    if (order == 0)
      center_of_arc = (int)std::ceil((double)(two_over_lambdaexc / dkz)) + 3;
    else if (order == 1)
      center_of_arc = (int)std::ceil(1. / lambdaexc * (std::cos((double)beta) + 1) / dkz) + 3;
    else
      center_of_arc = (int)std::ceil((double)two_over_lambdaexc * std::cos((double)beta) / dkz) + 3;
    for (int ix = 0; ix < icleanup; ix++) {
      const float kr = ix * dkr;
      if (kr <= krmax) {
        const int kzstart =
            (int)std::ceil(center_of_arc +
                           std::sqrt((double)(two_over_lambdaem * two_over_lambdaem - kr * kr)) / dkz);
        const int kzend = nz - kzstart;
        for (int iz = kzstart; iz <= kzend; iz++) otf[ix * nz + iz] = zero;
      } else {
        for (int iz = 0; iz < nz; iz++) otf[ix * nz + iz] = zero;
      }
    }
  }
}

// fixorigin (radialft.cpp:1188-1224)
void fixorigin(cf* otf, int nz, int kx1, int kx2) {
  const float meani = 0.5 * (kx1 + kx2);
  const int numvals = kx2 - kx1 + 1;
  std::vector<double> sum(kx2 + 1, 0.0);
  double totsum = 0, ysum = 0, sqsum = 0;
  for (int i = 0; i <= kx2; i++) {
    sum[i] = 0;
    if (i == 0)
      sum[i] = otf[0].real();
    else
      for (int j = 0; j < nz; j++) sum[i] += otf[i * nz + j].real();
    if (i >= kx1) {
      totsum += sum[i];
      ysum += sum[i] * (i - meani);
      sqsum += (i - meani) * (i - meani);
    }
  }
  const float slope = ysum / sqsum;
  const float avg = totsum / numvals;
  for (int i = 0; i < kx1; i++) {
    const float lineval = avg + (i - meani) * slope;
    otf[i * nz] = cf((float)(otf[i * nz].real() - (sum[i] - lineval)), otf[i * nz].imag());
  }
}

// rescale (radialft.cpp:1226-1242)
void rescale(cf* otf, int order, int nxz, float* scalefactor, int dorescale) {
  if (order == 0 || dorescale) *scalefactor = 1 / otf[0].real();
  for (int i = 0; i < nxz; i++) otf[i] = cf(otf[i].real() * *scalefactor, otf[i].imag() * *scalefactor);
}

// combine_reim (radialft.cpp:746-782); returns the phase of each order
void combine_reim(std::vector<cf*>& otf, int norders, int nx, int nkr, int nz, int forced_pi, float* phis) {
  const int nxz = nkr * nz;
  const int probe = std::min(nx / 4, nkr - 1) * nz;  // kr = nx/4, kz = 0
  for (int order = 1; order < norders; order++) {
    double bandre_mag = 0, bandim_mag = 0;
    for (int i = 0; i < nxz; i++) {
      bandre_mag += std::abs(otf[2 * order - 1][i]);
      bandim_mag += std::abs(otf[2 * order][i]);
    }
    double phi = std::atan(bandim_mag / bandre_mag);
    const float re = otf[2 * order - 1][probe].real(), im = otf[2 * order][probe].real();
    if (re < 0 && im > 0)
      phi += kPI;
    else if (re < 0 && im < 0)
      phi = kPI - phi;
    else if (re > 0 && im > 0)
      phi = -phi;
    if (order == 1 && forced_pi) phi = kPI - phi;
    phis[order] = (float)phi;
    const float c = (float)std::cos(phi), s = (float)std::sin(-phi);
    for (int i = 0; i < nxz; i++) {
      const cf a = otf[2 * order - 1][i], b = otf[2 * order][i];
      otf[2 * order - 1][i] = cf(a.real() * c + b.real() * s, a.imag() * c + b.imag() * s);
      otf[2 * order][i] = cf(0, 0);
    }
  }
}

}  // namespace

extern "C" {

const char* orc_makeotf_error(void) { return g_mk_err.c_str(); }

// Mirrors main() of radialft.cpp:82-434 from the loaded sections on.
int orc_makeotf(const simr_otf_params* p, const float* psf, int nx, int ny, int nsections, const float* section_bg,
                float* otf_out, float* band0, simr_otf_info* info) {
  try {
    const int nphases = p->nphases;
    if (nphases < 1 || nphases > 2 * SIMR_MAX_ORDERS - 1) throw std::runtime_error("nphases must be 1..5");
    const int nz = nsections / nphases;
    if (nz < 1 || nx < 2 || ny < 2) throw std::runtime_error("PSF too small");
    const int norders = (nphases + 1) / 2;
    const size_t nxy = (size_t)nx * ny;
    const int nr = std::min(nx, ny), nkr = nr / 2 + 1;
    const int border = 20;
    // read + camera fix + background + apodize (radialft.cpp:207-260)
    std::vector<std::vector<float>> img(nphases, std::vector<float>(nxy * nz));
    std::vector<float> background((size_t)nz * nphases, 0.f);
    std::vector<float> buf(nxy);
    int zsec = 0;
    for (int z = 0; z < nz; z++)
      for (int phase = 0; phase < nphases; phase++) {
        std::memcpy(buf.data(), psf + nxy * zsec, nxy * sizeof(float));
        if (buf[nxy - 1] < 0) buf[nxy - 1] = buf[nxy - 2];
        if (section_bg)
          background[zsec] = section_bg[zsec];
        else if (p->background >= 0)
          background[zsec] = p->background;
        else
          background[zsec] = estimate_background(buf.data(), nx, ny, border);
        float* sec = img[phase].data() + nxy * z;
        std::memcpy(sec, buf.data(), nxy * sizeof(float));
        if (p->napodize >= 0)
          apodize(p->napodize, nx, nx, ny, sec);
        else if (p->napodize == -1)
          cosapodize(nx, nx, ny, sec);
        zsec++;
      }
    // determine_center (radialft.cpp:507-665) on the phase mean; the
    // reference's stack is nx+pad wide, so x neighbours wrap with that width
    // and the padded columns read as 0.
    const int nxe = (nx % 2) ? nx + 1 : nx + 2;
    auto mean_at = [&](int x, int y, int z) -> float {
      if (x >= nx) return 0.f;
      float s = 0.f;
      for (int ph = 0; ph < nphases; ph++) s += img[ph][nxy * z + (size_t)y * nx + x];
      return s / nphases;
    };
    float maxval = 0.0f;
    int maxi = 0, maxj = 0, maxk = 0;
    for (int z = 0; z < nz; z++)
      for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
          const float v = mean_at(x, y, z);
          if (v > maxval) {
            maxval = v;
            maxi = y;
            maxj = x;
            maxk = z;
          }
        }
    int iminus = maxi - 1, iplus = maxi + 1, jminus = maxj - 1, jplus = maxj + 1, kminus = maxk - 1, kplus = maxk + 1;
    if (iminus < 0) iminus += ny;
    if (iplus >= ny) iplus -= ny;
    if (jminus < 0) jminus += nxe;
    if (jplus >= nxe) jplus -= nxe;
    if (kminus < 0) kminus += nz;
    if (kplus >= nz) kplus -= nz;
    float xc, yc, zc;
    if (nz > 1)
      zc = maxk + fitparabola(mean_at(maxj, maxi, kminus), maxval, mean_at(maxj, maxi, kplus));
    else
      zc = maxk;
    yc = maxi + fitparabola(mean_at(maxj, iminus, maxk), maxval, mean_at(maxj, iplus, maxk));
    xc = maxj + fitparabola(mean_at(jminus, maxi, maxk), maxval, mean_at(jplus, maxi, maxk));
    // background subtraction (radialft.cpp:280-286)
    for (int z = 0; z < nz; z++)
      for (int phase = 0; phase < nphases; phase++) {
        float* sec = img[phase].data() + nxy * z;
        const float b = background[(size_t)z * nphases + phase];
        for (size_t i = 0; i < nxy; i++) sec[i] -= b;
      }
    // makematrix / separate (radialft.cpp:440-485), 1/nphases normalised
    if (nphases > 1) {
      std::vector<float> M((size_t)nphases * nphases);  // M[j*nphases + i] = sepMatrix(j, i)
      const double phi = 2 * kPI / nphases;
      for (int j = 0; j < nphases; j++) {
        M[(size_t)j * nphases + 0] = 1.0 / nphases;
        for (int order = 1; order < norders; order++) {
          M[(size_t)j * nphases + 2 * order - 1] = std::cos(j * order * phi) / nphases;
          M[(size_t)j * nphases + 2 * order] = std::sin(j * order * phi) / nphases;
        }
      }
      std::vector<float> out(nphases);
      for (size_t x = 0; x < nxy * nz; x++) {
        for (int i = 0; i < nphases; i++) {
          out[i] = 0.0;
          for (int j = 0; j < nphases; j++) out[i] += img[j][x] * M[(size_t)j * nphases + i];
        }
        for (int i = 0; i < nphases; i++) img[i][x] = out[i];
      }
    }
    if (band0) std::memcpy(band0, img[0].data(), nxy * nz * sizeof(float));
    // FFT, shift_center, beadsize_compensate (radialft.cpp:339-366)
    const size_t H = (size_t)(nx / 2 + 1) * ny * nz;
    std::vector<std::vector<cf>> spec(nphases, std::vector<cf>(H));
    std::vector<cf*> bands(nphases);
    for (int ph = 0; ph < nphases; ph++) {
      rfft3d(img[ph].data(), nx, ny, nz, spec[ph].data());
      bands[ph] = spec[ph].data();
      shift_center(bands[ph], nx, ny, nz, xc, yc, zc);
    }
    if (p->do_compen)
      beadsize_compensate(bands, p->k0angle, p->linespacing, p->bead_diameter, norders, nx, ny, nz, p->dr, p->dz);
    // radial averages (radialft.cpp:369-385)
    std::vector<std::vector<cf>> avgbuf(nphases, std::vector<cf>((size_t)nkr * nz));
    std::vector<cf*> avg(nphases);
    for (int ph = 0; ph < nphases; ph++) {
      avg[ph] = avgbuf[ph].data();
      radialft(bands[ph], nx, ny, nz, p->dr, avg[ph], nr);
    }
    // modify / cleanup / fixorigin / rescale (radialft.cpp:393-412)
    float scalefactor = 1.0f;
    const int icleanup = nx / 2 + 1 < nkr ? nx / 2 + 1 : nkr;
    for (int phase = 0; phase < nphases; phase++) {
      const int order = (phase + 1) / 2;
      modify(avg[phase], nkr, nz, p->ifixz, p->ifixr, order);
      if (p->coherent_bsim) {
      } else if (p->leavekz[0] > 0) {
        cleanup(avg[phase], order, nr, nz, p->dr, p->dz, p->linespacing, (unsigned)p->wavelength, icleanup, p->leavekz[0],
                p->leavekz[1], p->leavekz[2], p->twolens, p->na, p->nimm);
      }
      if (phase == 0 && p->interpkr[0] > 0) fixorigin(avg[phase], nz, p->interpkr[0], p->interpkr[1]);
      rescale(avg[phase], order, nkr * nz, &scalefactor, p->dorescale);
    }
    float phis[SIMR_MAX_ORDERS] = {0, 0, 0};
    if (!p->five_bands && nphases > 1) combine_reim(avg, norders, nx, nkr, nz, p->forced_pi_shift, phis);
    if (p->conjugate)
      for (int ph = 0; ph < nphases; ph++)
        for (int i = 0; i < nkr * nz; i++) avg[ph][i] = std::conj(avg[ph][i]);
    // output order (outputdata, radialft.cpp:1037-1071)
    const int nbo = p->five_bands ? 2 * norders - 1 : norders;
    const size_t per = (size_t)nkr * nz;
    int slot = 0;
    for (int i = 0; i < norders; i++) {
      if (i == 0) {
        std::memcpy(otf_out + 2 * per * slot++, avg[0], per * sizeof(cf));
      } else {
        std::memcpy(otf_out + 2 * per * slot++, avg[2 * i - 1], per * sizeof(cf));
        if (p->five_bands) std::memcpy(otf_out + 2 * per * slot++, avg[2 * i], per * sizeof(cf));
      }
    }
    if (info) {
      std::memset(info, 0, sizeof(*info));
      info->nx = nx;
      info->ny = ny;
      info->nz = nz;
      info->norders = norders;
      info->nbands_out = nbo;
      info->nkr = nkr;
      info->dkr = 1 / (nx * p->dr);
      info->dkz = 1 / (nz * p->dz);
      info->xc = xc;
      info->yc = yc;
      info->zc = zc;
      for (int o = 0; o < SIMR_MAX_ORDERS; o++) info->phi[o] = phis[o];
      info->scalefactor = scalefactor;
    }
    return 0;
  } catch (const std::exception& e) {
    g_mk_err = e.what();
    return -1;
  }
}

// The I2M branch of main() (radialft.cpp:267-286 reading, determine_center's
// I2M search 624-664, 306-307 background, 345-357 FFT and centre shift,
// 383-386 radial average, 425-430 fixorigin + rescale): i2m holds the nz I2M
// sections [nz][ny][nx]; otf_out receives [nkr][nz] complex.
int orc_makeotf_i2m(const simr_otf_params* p, const float* i2m, int nx, int ny, int nz, float* otf_out,
                    simr_otf_info* info) {
  try {
    if (nz < 1 || nx < 2 || ny < 2) throw std::runtime_error("PSF too small");
    const size_t nxy = (size_t)nx * ny;
    const int nr = std::min(nx, ny), nkr = nr / 2 + 1;
    const int border = 20;
    std::vector<float> img(nxy * nz);
    float background_i2m = 0.f;
    for (int z = 0; z < nz; z++) {
      float* sec = img.data() + nxy * z;
      std::memcpy(sec, i2m + nxy * z, nxy * sizeof(float));
      if (sec[nxy - 1] == -1.0) sec[nxy - 1] = sec[nxy - 2];  // fix camera error at even binnings
      if (z == nz / 2) background_i2m = estimate_background(sec, nx, ny, border);
    }
    // the peak of the I2M volume (before the background subtraction), x
    // neighbours wrapping with the padded width, padded columns reading 0
    const int nxe = (nx % 2) ? nx + 1 : nx + 2;
    auto at = [&](int x, int y, int z) -> float { return x >= nx ? 0.f : img[nxy * z + (size_t)y * nx + x]; };
    float maxval = 0.0f;
    int maxi = 0, maxj = 0, maxk = 0;
    for (int z = 0; z < nz; z++)
      for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
          const float v = at(x, y, z);
          if (v > maxval) {
            maxval = v;
            maxi = y;
            maxj = x;
            maxk = z;
          }
        }
    int iminus = maxi - 1, iplus = maxi + 1, jminus = maxj - 1, jplus = maxj + 1, kminus = maxk - 1, kplus = maxk + 1;
    if (iminus < 0) iminus += ny;
    if (iplus >= ny) iplus -= ny;
    if (jminus < 0) jminus += nxe;
    if (jplus >= nxe) jplus -= nxe;
    if (kminus < 0) kminus += nz;
    if (kplus >= nz) kplus -= nz;
    const float zc = maxk + fitparabola(at(maxj, maxi, kminus), maxval, at(maxj, maxi, kplus));
    const float yc = maxi + fitparabola(at(maxj, iminus, maxk), maxval, at(maxj, iplus, maxk));
    const float xc = maxj + fitparabola(at(jminus, maxi, maxk), maxval, at(jplus, maxi, maxk));
    for (float& v : img) v -= background_i2m;
    const size_t H = (size_t)(nx / 2 + 1) * ny * nz;
    std::vector<cf> spec(H);
    rfft3d(img.data(), nx, ny, nz, spec.data());
    shift_center(spec.data(), nx, ny, nz, xc, yc, zc);
    std::vector<cf> avg((size_t)nkr * nz);
    radialft(spec.data(), nx, ny, nz, p->dr, avg.data(), nr);
    // unconditional in the reference; it changes nothing unless kx1 > 0
    if (p->interpkr[0] > 0) fixorigin(avg.data(), nz, p->interpkr[0], p->interpkr[1]);
    float scalefactor = 1.0f;
    rescale(avg.data(), 0, nkr * nz, &scalefactor, p->dorescale);
    std::memcpy(otf_out, avg.data(), avg.size() * sizeof(cf));
    if (info) {
      std::memset(info, 0, sizeof(*info));
      info->nx = nx;
      info->ny = ny;
      info->nz = nz;
      info->norders = 1;
      info->nbands_out = 1;
      info->nkr = nkr;
      info->dkr = 1 / (nx * p->dr);
      info->dkz = 1 / (nz * p->dz);
      info->xc = xc;
      info->yc = yc;
      info->zc = zc;
      info->scalefactor = scalefactor;
    }
    return 0;
  } catch (const std::exception& e) {
    g_mk_err = e.what();
    return -1;
  }
}

}  // extern "C"
