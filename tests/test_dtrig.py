"""The line searches' cos/sin/atan2 (cudasirecon_amd/csrc/dtrig.hpp): a fixed
sequence of IEEE double operations, compiled here for the host, checked
against the C library's double functions - within a few ulps, and the same
float after one rounding for a million arguments of the fit's range, plus
the special values atan2 must keep."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = r'''
#include "dtrig.hpp"
#include <cmath>
#include <cstdio>
#include <random>
using namespace simr::dtrig;
static double ulps(double a, double b) {
  if (a == b) return 0;
  return std::fabs(a - b) / std::fabs(std::nextafter(b, INFINITY) - b);
}
int main() {
  std::mt19937_64 g(7);
  std::uniform_real_distribution<double> A(-8, 8), V(-5, 5);
  double mc = 0, ms = 0, ma = 0;
  long bad = 0;
  for (long i = 0; i < 1000000; ++i) {
    const float x = (float)A(g), y = (float)V(g), z = (float)V(g);
    const double c = dcos(x), s = dsin(x), a = datan2(y, z);
    mc = std::fmax(mc, ulps(c, std::cos((double)x)));
    ms = std::fmax(ms, ulps(s, std::sin((double)x)));
    ma = std::fmax(ma, ulps(a, std::atan2((double)y, (double)z)));
    bad += (float)c != (float)std::cos((double)x);
    bad += (float)s != (float)std::sin((double)x);
    bad += (float)a != (float)std::atan2((double)y, (double)z);
  }
  const bool sp = datan2(0.0, 0.0) == 0.0 && datan2(0.0, -0.0) == M_PI && datan2(-0.0, -1.0) == -M_PI &&
                  datan2(1.0, 0.0) == M_PI / 2 && datan2(-1.0, 0.0) == -M_PI / 2 && std::isnan(datan2(NAN, 1.0)) &&
                  dcos(0.0) == 1.0 && dsin(0.0) == 0.0;
  std::printf("%.3f %.3f %.3f %ld %d\n", mc, ms, ma, bad, sp ? 1 : 0);
  return 0;
}
'''


def test_dtrig_against_libm(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(SRC)
    exe = tmp_path / "t"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "cudasirecon_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True)
    mc, ms, ma, bad, sp = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert float(mc) <= 2 and float(ms) <= 2 and float(ma) <= 4, (mc, ms, ma)
    assert int(bad) == 0
    assert sp == "1"
