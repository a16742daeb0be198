// Gauss-Legendre rules for the Galerkin integrals over pairs of flat
// triangles.
//
// Regular pairs use the tensor product of two single-triangle rules (a
// Duffy-collapsed q x q Gauss rule, Library/singquad2d.c:25-78).  The
// Sauter-Schwab points of pairs with common vertices are formed inside the
// singular kernel from the same Gauss nodes (Library/singquad2d.c:80-281).
// Gauss-Legendre nodes are computed by Newton's method on the Legendre
// recurrence (the reference uses the Golub-Welsch eigenproblem,
// Library/gaussquad.c:12-37; both give the same nodes to rounding).
#include <cmath>

#include "h2hip_internal.h"

namespace h2hip {

void gauss_legendre01(unsigned m, std::vector<double> &x,
                      std::vector<double> &w) {
  x.assign(m, 0.0);
  w.assign(m, 0.0);
  for (unsigned i = 0; i < m; i++) {
    // i-th node from the left on [-1, 1]
    double z = -std::cos(M_PI * (4.0 * i + 3.0) / (4.0 * m + 2.0));
    double dpm = 1.0;
    for (int iter = 0; iter < 64; iter++) {
      double pm1 = 0.0, pm = 1.0;  // P_{n-1}, P_n
      for (unsigned n = 0; n < m; n++) {
        double pn1 = ((2.0 * n + 1.0) * z * pm - n * pm1) / (n + 1.0);
        pm1 = pm;
        pm = pn1;
      }
      dpm = m * (pm1 - z * pm) / (1.0 - z * z);
      double step = pm / dpm;
      z -= step;
      if (std::fabs(step) <= 1e-16 * std::fabs(z) + 1e-300) break;
    }
    double pm1 = 0.0, pm = 1.0;
    for (unsigned n = 0; n < m; n++) {
      double pn1 = ((2.0 * n + 1.0) * z * pm - n * pm1) / (n + 1.0);
      pm1 = pm;
      pm = pn1;
    }
    dpm = m * (pm1 - z * pm) / (1.0 - z * z);
    double wz = 2.0 / ((1.0 - z * z) * dpm * dpm);
    // map [-1, 1] -> [0, 1]
    x[i] = 0.5 + 0.5 * z;
    w[i] = 0.5 * wz;
  }
}

TriRule make_triangle_rule(unsigned q) {
  std::vector<double> g, gw;
  gauss_legendre01(q, g, gw);
  TriRule r;
  for (unsigned i = 0; i < q; i++)
    for (unsigned j = 0; j < q; j++) {
      r.x0.push_back(g[i]);
      r.x1.push_back(g[i] * g[j]);
      r.w.push_back(gw[i] * gw[j] * g[i]);
    }
  return r;
}

}  // namespace h2hip
