// FP64 sine and cosine for the Green's function phase k r.
//
// Cody-Waite reduction by pi/2 with a three-part constant (the first two parts
// hold 30 significant bits, so n * part is exact for |n| < 2^22), followed by
// the Taylor polynomials of sin to degree 15 and cos to degree 16 on
// |y| <= pi/4, where the first omitted terms are below 1e-16 relative. The
// constants are pi/2 split from a 60-digit value of pi (Machin's formula) and
// the reciprocal factorials 1/n!. Phases beyond 2^20 pi/2 go to the library
// routine.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define BEAT_HD __host__ __device__ __forceinline__
#else
#define BEAT_HD inline
#endif

namespace beat {

BEAT_HD void sincos_reduced(double x, double* s, double* c) {
  const double n = rint(x * 0.6366197723675814);  // 2/pi
  double y = fma(-n, 1.5707963276654482, x);
  y = fma(-n, -8.705515692000731e-10, y);
  y = fma(-n, -3.50343439808993e-19, y);
  const double y2 = y * y;
  double ps = -7.647163731819816e-13;        // -1/15!
  ps = fma(ps, y2, 1.6059043836821613e-10);  // 1/13!
  ps = fma(ps, y2, -2.505210838544172e-08);  // -1/11!
  ps = fma(ps, y2, 2.7557319223985893e-06);  // 1/9!
  ps = fma(ps, y2, -0.0001984126984126984);  // -1/7!
  ps = fma(ps, y2, 0.008333333333333333);    // 1/5!
  ps = fma(ps, y2, -0.16666666666666666);    // -1/3!
  const double sy = fma(ps * y2, y, y);
  double pc = 4.779477332387385e-14;         // 1/16!
  pc = fma(pc, y2, -1.1470745597729725e-11); // -1/14!
  pc = fma(pc, y2, 2.08767569878681e-09);    // 1/12!
  pc = fma(pc, y2, -2.755731922398589e-07);  // -1/10!
  pc = fma(pc, y2, 2.48015873015873e-05);    // 1/8!
  pc = fma(pc, y2, -0.001388888888888889);   // -1/6!
  pc = fma(pc, y2, 0.041666666666666664);    // 1/4!
  pc = fma(pc, y2, -0.5);
  const double cy = fma(pc, y2, 1.0);
  // x = n pi/2 + y
  const int q = (int)(long long)n & 3;
  double ss = (q & 1) ? cy : sy;
  double cc = (q & 1) ? sy : cy;
  if (q & 2) ss = -ss;
  if ((q + 1) & 2) cc = -cc;
  *s = ss;
  *c = cc;
}

// Table form for |x| < SINCOS_TABLE_LIMIT (below 2^23 pi/64, where n * part stays
// exact): n = rint(x 64/pi), y = x - n pi/64 with the pi/2 split above scaled by
// 1/32 (exact), |y| <= pi/128; Taylor polynomials of sin to y^7 and cos to y^6
// (first omitted terms below 1e-17 absolute), then the rotation by the table
// entry (sin, cos)(m pi/64), m = n mod 128: sin x = S cos y + C sin y,
// cos x = C cos y - S sin y. The reduction and the rotation are separate
// functions, so that sincos_steps below shares the rotation.
constexpr double SINCOS_TABLE_LIMIT = 4.0e5;
constexpr int TRIG_N = 128;  // table entries (sin, cos) over one period

// sin y, cos y for |y| <= pi/128, rotated by the table entry (S, C).
BEAT_HD void sincos_rotate(double y, double S, double C, double* s, double* c) {
  const double y2 = y * y;
  double ps = -0.0001984126984126984;        // -1/7!
  ps = fma(ps, y2, 0.008333333333333333);    // 1/5!
  ps = fma(ps, y2, -0.16666666666666666);    // -1/3!
  const double sy = fma(ps * y2, y, y);
  double pc = -0.001388888888888889;         // -1/6!
  pc = fma(pc, y2, 0.041666666666666664);    // 1/4!
  pc = fma(pc, y2, -0.5);
  const double cy = fma(pc, y2, 1.0);
  *s = fma(S, cy, C * sy);
  *c = fma(C, cy, -(S * sy));
}

// The same on the remainder in table steps, u = t - n (|u| <= 1/2, y = u pi/64):
// the powers of pi/64 sit in the coefficients, sin y = u (h - h^3/3! u^2 + h^5/5!
// u^4 - h^7/7! u^6), cos y = 1 - h^2/2! u^2 + h^4/4! u^4 - h^6/6! u^6 with h = pi/64
// (coefficients rounded once from a 50-digit evaluation).
BEAT_HD void sincos_rotate_steps(double u, double S, double C, double* s, double* c) {
  const double u2 = u * u;
  double ps = -1.36256978594427006e-13;
  ps = fma(ps, u2, 2.37502533931037925e-09);
  ps = fma(ps, u2, -1.97132598115919886e-05);
  ps = fma(ps, u2, 4.90873852123405174e-02);
  const double sy = ps * u;
  double pc = -1.94306306199663886e-11;
  pc = fma(pc, u2, 2.41918094540641814e-07);
  pc = fma(pc, u2, -1.20478569349235327e-03);
  const double cy = fma(pc, u2, 1.0);
  *s = fma(S, cy, C * sy);
  *c = fma(C, cy, -(S * sy));
}

BEAT_HD void sincos_table(double x, const double* trig, double* s, double* c) {
  const double n = rint(x * 20.371832715762604);  // 64/pi
  double y = fma(-n, 1.5707963276654482 / 32, x);
  y = fma(-n, -8.705515692000731e-10 / 32, y);
  y = fma(-n, -3.50343439808993e-19 / 32, y);
  const int m = (int)n & (TRIG_N - 1);
  sincos_rotate(y, trig[2 * m], trig[2 * m + 1], s, c);
}

// The table form on a phase given in table steps, t = x 64/pi with |t| below
// SINCOS_TABLE_LIMIT 64/pi: n = rint(t), y = (t - n) pi/64, where t - n is exact
// and the product adds a relative rounding of y alone (|y| <= pi/128). Callers
// fold 64/pi into the wavenumber, so forming t is one product: the phase then
// carries the rounding of t (|x| 2^-53) and of the folded wavenumber (another
// |x| 2^-53), against |x| 2^-53 for rounding k r in sincos_table.
BEAT_HD void sincos_steps(double t, const double* trig, double* s, double* c) {
  const double n = rint(t);
  const double y = (t - n) * 0.04908738521234052;  // pi/64
  const int m = (int)n & (TRIG_N - 1);
  sincos_rotate(y, trig[2 * m], trig[2 * m + 1], s, c);
}

// Wavenumber scale for sincos_steps: k 64/pi, rounded once (host code).
inline double phase_steps_per_unit(double k) {
  return (double)((long double)k * 64.0L / 3.141592653589793238462643383279502884L);
}

// (sin, cos)(m pi/64) for m = 0..TRIG_N-1, interleaved, evaluated in long double
// and rounded once (host code).
inline void fill_trig_table(double* t) {
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int m = 0; m < TRIG_N; ++m) {
    const long double a = (long double)m * (pi / 64.0L);
    t[2 * m] = (double)sinl(a);
    t[2 * m + 1] = (double)cosl(a);
  }
}

BEAT_HD void sincos_phase(double x, double* s, double* c) {
  if (fabs(x) < 1647099.3291652855) {  // 2^20 pi/2
    sincos_reduced(x, s, c);
  } else {
    *s = sin(x);
    *c = cos(x);
  }
}

}  // namespace beat
