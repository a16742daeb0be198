// FP64 elementary functions of the kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>

namespace h2hip {

// sin and cos of x: Cody and Waite's reduction by pi/2 in three parts
// (n = rint(2x / pi), r = x - n pi/2 with fused products) to |r| <= pi/4,
// then the Taylor polynomials of sin r and cos r to degree 17 and 16 in
// Horner form (coefficients (-1)^k / k!; the first terms left out are
// 8.4e-20 and 2.0e-18 at most, at |r| = pi/4).  Absolute error below
// 0.59 * 2^-52 for |x| < 2^45 (the same operations in C with fma, against
// long-double sinl / cosl on 4e6 arguments, a quarter of them next to
// multiples of pi/2); beyond, where k r of a coupling never comes, the
// error grows with the rounding of 2x / pi; NaN and infinities give NaN.
// No branches, so the evaluations of the coupling kernels stay branch-free
// (the library's sincos carries a large-argument path).
__device__ __forceinline__ void sincos_cw(double x, double &s, double &c) {
  constexpr double TWO_OVER_PI = 0.6366197723675814;
  constexpr double P1 = 1.5707963267948966, P2 = 6.123233995736766e-17,
                   P3 = -1.4973849048591698e-33;
  const double n = rint(x * TWO_OVER_PI);
  double r = fma(-n, P1, x);
  r = fma(-n, P2, r);
  r = fma(-n, P3, r);
  const double r2 = r * r;
  double ps = 1.0 / 355687428096000.0;  // 1 / 17!
  ps = fma(ps, r2, -1.0 / 1307674368000.0);
  ps = fma(ps, r2, 1.0 / 6227020800.0);
  ps = fma(ps, r2, -1.0 / 39916800.0);
  ps = fma(ps, r2, 1.0 / 362880.0);
  ps = fma(ps, r2, -1.0 / 5040.0);
  ps = fma(ps, r2, 1.0 / 120.0);
  ps = fma(ps, r2, -1.0 / 6.0);
  const double sr = fma(r * r2, ps, r);
  double pc = 1.0 / 20922789888000.0;  // 1 / 16!
  pc = fma(pc, r2, -1.0 / 87178291200.0);
  pc = fma(pc, r2, 1.0 / 479001600.0);
  pc = fma(pc, r2, -1.0 / 3628800.0);
  pc = fma(pc, r2, 1.0 / 40320.0);
  pc = fma(pc, r2, -1.0 / 720.0);
  pc = fma(pc, r2, 1.0 / 24.0);
  pc = fma(pc, r2, -0.5);
  const double cr = fma(r2, pc, 1.0);
  // quadrant n mod 4: (sin, cos) = (sr, cr), (cr, -sr), (-sr, -cr), (-cr, sr)
  const int q = (int)(long long)n & 3;
  const double s0 = (q & 1) ? cr : sr, c0 = (q & 1) ? sr : cr;
  s = (q & 2) ? -s0 : s0;
  c = ((q + 1) & 2) ? -c0 : c0;
}

}  // namespace h2hip
