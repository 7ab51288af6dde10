// cos, sin and atan2 in double, by a fixed sequence of IEEE operations, so
// that the host and the device compute the same bits.  The line searches
// form their candidate vectors from these (rounded once to float): the host
// walk and the device walk then evaluate identical candidates.
//
// cos/sin: x = k pi/2 + r with pi/2 in three parts (Cody-Waite; k pi/2 is
// exact in the first part for |k| < 2^20), then Taylor polynomials in r,
// |r| <= pi/4, to r^22 / r^23 (truncation below 1e-23).  atan2: the ratio
// t = min/max of |x|, |y| in [0, 1]; atan t = pi/4 + atan((t-1)/(t+1)) for
// t > tan(pi/8), leaving |u| <= 0.4143, and the alternating series to u^61
// (truncation below 1e-24); then the octant and quadrant.  Each is within a
// few ulps of the exact value for the arguments the fit uses (|x| < 1e5).
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define SIMR_HD __host__ __device__
#else
#define SIMR_HD
#endif

namespace simr {
namespace dtrig {

// pi/2 = P1 + P2 + P3 (P1: 33 significant bits, P2 the next 33, P3 the rest)
constexpr double kP1 = 1.57079632673412561417e+00;
constexpr double kP2 = 6.07710050630396597660e-11;
constexpr double kP3 = 2.02226624879595063154e-21;
constexpr double kTwoOverPi = 6.36619772367581382433e-01;
constexpr double kPi = 3.14159265358979311600e+00;
constexpr double kPi2 = 1.57079632679489655800e+00;
constexpr double kPi4 = 7.85398163397448278999e-01;
constexpr double kTanPi8 = 4.14213562373095145475e-01;

SIMR_HD inline double reduce(double x, int* q) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  const double k = std::rint(x * kTwoOverPi);
  *q = (int)((long long)k & 3);
  return ((x - k * kP1) - k * kP2) - k * kP3;
}

// sin r and cos r, |r| <= pi/4 (Horner on r^2)
SIMR_HD inline double sin_poly(double r) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  const double r2 = r * r;
  double p = 1.0 / 51090942171709440000.0;  // 1/21!
  p = p * r2 - 1.0 / 121645100408832000.0;  // 1/19!
  p = p * r2 + 1.0 / 355687428096000.0;     // 1/17!
  p = p * r2 - 1.0 / 1307674368000.0;       // 1/15!
  p = p * r2 + 1.0 / 6227020800.0;          // 1/13!
  p = p * r2 - 1.0 / 39916800.0;            // 1/11!
  p = p * r2 + 1.0 / 362880.0;              // 1/9!
  p = p * r2 - 1.0 / 5040.0;                // 1/7!
  p = p * r2 + 1.0 / 120.0;                 // 1/5!
  p = p * r2 - 1.0 / 6.0;                   // 1/3!
  return r + (r * r2) * p;
}
SIMR_HD inline double cos_poly(double r) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  const double r2 = r * r;
  double p = -1.0 / 1124000727777607680000.0;  // -1/22!
  p = p * r2 + 1.0 / 2432902008176640000.0;    // 1/20!
  p = p * r2 - 1.0 / 6402373705728000.0;       // 1/18!
  p = p * r2 + 1.0 / 20922789888000.0;         // 1/16!
  p = p * r2 - 1.0 / 87178291200.0;            // 1/14!
  p = p * r2 + 1.0 / 479001600.0;              // 1/12!
  p = p * r2 - 1.0 / 3628800.0;                // 1/10!
  p = p * r2 + 1.0 / 40320.0;                  // 1/8!
  p = p * r2 - 1.0 / 720.0;                    // 1/6!
  p = p * r2 + 1.0 / 24.0;                     // 1/4!
  p = p * r2 - 0.5;                            // 1/2!
  return 1.0 + r2 * p;
}

SIMR_HD inline double dcos(double x) {
  if (!(std::fabs(x) < 1e5)) return std::cos(x);  // not reached by the fit (NaN and Inf pass through)
  int q;
  const double r = reduce(x, &q);
  switch (q) {
    case 0: return cos_poly(r);
    case 1: return -sin_poly(r);
    case 2: return -cos_poly(r);
    default: return sin_poly(r);
  }
}
SIMR_HD inline double dsin(double x) {
  if (!(std::fabs(x) < 1e5)) return std::sin(x);
  int q;
  const double r = reduce(x, &q);
  switch (q) {
    case 0: return sin_poly(r);
    case 1: return cos_poly(r);
    case 2: return -sin_poly(r);
    default: return -cos_poly(r);
  }
}

// atan t, 0 <= t <= 1
SIMR_HD inline double atan01(double t) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  double base = 0.0, u = t;
  if (t > kTanPi8) {
    base = kPi4;
    u = (t - 1.0) / (t + 1.0);
  }
  const double u2 = u * u;
  double p = 1.0 / 61.0;
  for (int k = 59; k >= 1; k -= 2) p = (k % 4 == 1 ? 1.0 : -1.0) / (double)k + u2 * p;
  return base + u * p;
}

SIMR_HD inline double datan2(double y, double x) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  if (std::isnan(x) || std::isnan(y)) return x + y;
  const double ax = std::fabs(x), ay = std::fabs(y);
  if (ax == 0.0 && ay == 0.0) return std::signbit(x) ? std::copysign(kPi, y) : std::copysign(0.0, y);
  if (std::isinf(ax) || std::isinf(ay)) return std::atan2(y, x);
  double a = ay <= ax ? atan01(ay / ax) : kPi2 - atan01(ax / ay);
  if (std::signbit(x)) a = kPi - a;
  return std::signbit(y) ? -a : a;
}

// float results, rounded once
SIMR_HD inline float fcos(float x) { return (float)dcos((double)x); }
SIMR_HD inline float fsin(float x) { return (float)dsin((double)x); }
SIMR_HD inline float fatan2(float y, float x) { return (float)datan2((double)y, (double)x); }

}  // namespace dtrig
}  // namespace simr
