// Host check of beat::sincos_phase against the C library (compiled by the CPU test).
#include <cmath>
#include <cstdio>
#include <random>

#include "../../beat_engine_amd/csrc/bem_math.h"

int main() {
  std::mt19937_64 rng(7);
  double worst_s = 0, worst_c = 0;
  auto check = [&](double x) {
    double s, c;
    beat::sincos_phase(x, &s, &c);
    const double es = std::fabs(s - std::sin(x)), ec = std::fabs(c - std::cos(x));
    worst_s = std::fmax(worst_s, es);
    worst_c = std::fmax(worst_c, ec);
  };
  std::uniform_real_distribution<double> u(0.0, 1.0);
  for (int i = 0; i < 2000000; ++i) check(u(rng) * 2000.0);
  for (int i = 0; i < 200000; ++i) check(-u(rng) * 50.0);
  for (int i = 0; i < 200000; ++i) check(u(rng) * 2e6);
  for (int k = -64; k <= 64; ++k) {  // quadrant boundaries
    const double b = k * 1.5707963267948966;
    check(b);
    check(std::nextafter(b, 1e9));
    check(std::nextafter(b, -1e9));
  }
  check(0.0);
  std::printf("%.3e %.3e\n", worst_s, worst_c);
  return (worst_s < 4e-16 && worst_c < 4e-16) ? 0 : 1;
}
