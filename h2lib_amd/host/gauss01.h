/* Gauss-Legendre rule of m points on [0, 1] by Newton's method on the
 * Legendre three-term recurrence, started at the Tricomi-type guesses
 * cos(pi (4 i + 3) / (4 m + 2)).  The reference builds its rules from the
 * Golub-Welsch eigenproblem (Library/gaussquad.c:12-37); both give the
 * same nodes and weights to rounding.  Shared by the device rule tables
 * (csrc/rules.cpp) and the host integrators (h2interp.c, functionals.c);
 * each caller compiles it with its own floating-point flags. */
#ifndef H2LIB_AMD_GAUSS01_H
#define H2LIB_AMD_GAUSS01_H

#include <math.h>

static inline void
h2_gauss01(unsigned m, double *x, double *w)
{
  for (unsigned i = 0; i < m; i++) {
    double z = -cos(M_PI * (4.0 * i + 3.0) / (4.0 * m + 2.0)), dp = 1.0;
    for (int it = 0; it < 64; it++) {
      double p0 = 0.0, p1 = 1.0; /* P_{n-1}, P_n */
      for (unsigned n = 0; n < m; n++) {
        double p2 = ((2.0 * n + 1.0) * z * p1 - n * p0) / (n + 1.0);
        p0 = p1;
        p1 = p2;
      }
      dp = m * (p0 - z * p1) / (1.0 - z * z);
      double step = p1 / dp;
      z -= step;
      if (fabs(step) <= 1e-16 * fabs(z) + 1e-300)
        break;
    }
    double p0 = 0.0, p1 = 1.0;
    for (unsigned n = 0; n < m; n++) {
      double p2 = ((2.0 * n + 1.0) * z * p1 - n * p0) / (n + 1.0);
      p0 = p1;
      p1 = p2;
    }
    dp = m * (p0 - z * p1) / (1.0 - z * z);
    /* [-1, 1] -> [0, 1] */
    x[i] = 0.5 + 0.5 * z;
    w[i] = 0.5 * (2.0 / ((1.0 - z * z) * dp * dp));
  }
}

#endif
