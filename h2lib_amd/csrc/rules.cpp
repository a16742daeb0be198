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
#include "../host/gauss01.h"

namespace h2hip {

void gauss_legendre01(unsigned m, std::vector<double> &x,
                      std::vector<double> &w) {
  x.assign(m, 0.0);
  w.assign(m, 0.0);
  h2_gauss01(m, x.data(), w.data());
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
