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
  // step form: the phase given in table steps t = x 64/pi, checked against the
  // exact phase t pi/64 of that t, evaluated in long double after removing whole
  // periods exactly (t = n + f with f = t - rint(t) exact; n pi/64 is a whole
  // number of periods plus (n mod 128) pi/64)
  const long double pi_ld = 3.141592653589793238462643383279502884L;
  double ws_s = 0, ws_c = 0;
  auto check_s = [&](double t) {
    double s, c;
    beat::sincos_steps(t, trig, &s, &c);
    const double n = std::rint(t);
    const long double m = (long double)(((long long)n % 128 + 128) % 128);
    const long double x = (m + (long double)(t - n)) * (pi_ld / 64.0L);
    ws_s = std::fmax(ws_s, std::fabs(s - (double)sinl(x)));
    ws_c = std::fmax(ws_c, std::fabs(c - (double)cosl(x)));
  };
  const double lim_t = beat::SINCOS_TABLE_LIMIT * 64.0 / 3.141592653589793;
  for (int i = 0; i < 2000000; ++i) check_s(u(rng) * 2000.0 * 64.0 / 3.141592653589793);
  for (int i = 0; i < 200000; ++i) check_s(-u(rng) * 1000.0);
  for (int i = 0; i < 200000; ++i) check_s((2.0 * u(rng) - 1.0) * lim_t);
  for (int k = -400; k <= 400; ++k) {  // integer steps, their neighbours and half-steps
    check_s(k);
    check_s(std::nextafter((double)k, 1e9));
    check_s(std::nextafter((double)k, -1e9));
    check_s(k + 0.5);
    check_s(std::nextafter(k + 0.5, 1e9));
  }
  check_s(0.0);
  // the wavenumber scale is k 64/pi rounded once
  const double ks = beat::phase_steps_per_unit(91.6);
  const long double ks_exact = 91.6L * 64.0L / pi_ld;
  const double ks_rel = (double)std::fabs(((long double)ks - ks_exact) / ks_exact);
  std::printf("%.3e %.3e %.3e %.3e %.3e %.3e %.3e\n", worst_s, worst_c, wt_s, wt_c, ws_s, ws_c, ks_rel);
  return (worst_s < 4e-16 && worst_c < 4e-16 && wt_s < 4e-16 && wt_c < 4e-16 && ws_s < 4e-16 && ws_c < 4e-16 &&
          ks_rel <= 1.2e-16)
             ? 0
             : 1;
}
