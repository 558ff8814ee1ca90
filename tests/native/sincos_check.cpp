// Host check of beat::sincos_phase and beat::sincos_table against the C library
// (compiled by the CPU test).
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
  // table form, over its range
  double trig[2 * beat::TRIG_N];
  beat::fill_trig_table(trig);
  double wt_s = 0, wt_c = 0;
  auto check_t = [&](double x) {
    double s, c;
    beat::sincos_table(x, trig, &s, &c);
    wt_s = std::fmax(wt_s, std::fabs(s - std::sin(x)));
    wt_c = std::fmax(wt_c, std::fabs(c - std::cos(x)));
  };
  for (int i = 0; i < 2000000; ++i) check_t(u(rng) * 2000.0);
  for (int i = 0; i < 200000; ++i) check_t(-u(rng) * 50.0);
  for (int i = 0; i < 200000; ++i) check_t((2.0 * u(rng) - 1.0) * beat::SINCOS_TABLE_LIMIT);
  for (int k = -400; k <= 400; ++k) {  // across the table's pi/64 steps and half-steps
    const double b = k * (3.141592653589793 / 64.0);
    check_t(b);
    check_t(std::nextafter(b, 1e9));
    check_t(std::nextafter(b, -1e9));
    check_t(b + 3.141592653589793 / 128.0);
  }
  check_t(0.0);
  std::printf("%.3e %.3e %.3e %.3e\n", worst_s, worst_c, wt_s, wt_c);
  return (worst_s < 4e-16 && worst_c < 4e-16 && wt_s < 4e-16 && wt_c < 4e-16) ? 0 : 1;
}
