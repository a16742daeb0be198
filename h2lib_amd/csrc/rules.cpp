// Quadrature rules for Galerkin integrals over pairs of flat triangles.
//
// Regular pairs use the tensor product of two single-triangle rules
// (a Duffy-collapsed q x q Gauss rule).  Pairs with a common vertex, a common
// edge or identical triangles use the Sauter-Schwab subdivision into 2, 6 and
// 6 subregions with the Duffy maps of the reference
// (Library/singquad2d.c:25-281); the coordinates (x0, x1) describe the point
// A (1 - x0) + B (x0 - x1) + C x1 of a triangle ABC whose shared vertices
// come first.  Gauss-Legendre nodes are computed by Newton's method on the
// Legendre recurrence (the reference uses the Golub-Welsch eigenproblem,
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

namespace {

struct Emit {
  PairRule &r;
  void operator()(double w, double a0, double a1, double b0, double b1) {
    r.w.push_back(w);
    r.x0.push_back(a0);
    r.x1.push_back(a1);
    r.y0.push_back(b0);
    r.y1.push_back(b1);
  }
};

}  // namespace

// Common vertex A: two subregions (Library/singquad2d.c:80-125).
PairRule make_vertex_rule(unsigned q) {
  std::vector<double> g, gw;
  gauss_legendre01(q, g, gw);
  PairRule r;
  Emit emit{r};
  for (unsigned i = 0; i < q; i++)
    for (unsigned j = 0; j < q; j++)
      for (unsigned k = 0; k < q; k++)
        for (unsigned l = 0; l < q; l++) {
          const double e1 = g[i], e2 = g[j], e3 = g[k], xi = g[l];
          const double wt = gw[i] * gw[j] * gw[k] * gw[l];
          const double w = wt * xi * xi * xi * e1;
          emit(w, xi, e2 * xi, e1 * xi, e1 * e3 * xi);
          emit(w, e1 * xi, e1 * e3 * xi, xi, e2 * xi);
        }
  return r;
}

// Common edge AB: six subregions (Library/singquad2d.c:127-201).
PairRule make_edge_rule(unsigned q) {
  std::vector<double> g, gw;
  gauss_legendre01(q, g, gw);
  PairRule r;
  Emit emit{r};
  for (unsigned i = 0; i < q; i++)
    for (unsigned j = 0; j < q; j++)
      for (unsigned k = 0; k < q; k++)
        for (unsigned l = 0; l < q; l++) {
          const double e1 = g[i], e2 = g[j], x1 = g[k];
          const double x2 = g[l] * (1.0 - x1);
          const double wt = gw[i] * gw[j] * gw[k] * (gw[l] * (1.0 - x1));
          const double w = wt * x1 * x1 * e1;
          const double s = x1 + x2;
          emit(w, s, x1 * e1 * e2, x1 * (1.0 - e1) + x2, x1 * (1.0 - e1));
          emit(w, s, x1 * e1, x1 * (1.0 - e1 * e2) + x2,
               x1 * (1.0 - e1 * e2));
          emit(w, s, x1, x1 * e1 + x2, x1 * e1 * e2);
          emit(w, x1 * (1.0 - e1 + e1 * e2) + x2, x1 * e1 * e2, s, x1);
          emit(w, x1 * e1 * e2 + x2, x1 * e1 * e2, s, x1 * e1);
          emit(w, x1 * e1 + x2, x1 * e1, s, x1 * e1 * e2);
        }
  return r;
}

// Identical triangles: six subregions (Library/singquad2d.c:203-281).
PairRule make_identical_rule(unsigned q) {
  std::vector<double> g, gw;
  gauss_legendre01(q, g, gw);
  PairRule r;
  Emit emit{r};
  for (unsigned i = 0; i < q; i++)
    for (unsigned j = 0; j < q; j++)
      for (unsigned k = 0; k < q; k++)
        for (unsigned l = 0; l < q; l++) {
          const double eta = g[i], x1 = g[j];
          const double x2 = g[k] * (1.0 - x1);
          const double x3 = g[l] * (1.0 - x1 - x2);
          const double wt = gw[i] * gw[j] * (gw[k] * (1.0 - x1)) *
                            (gw[l] * (1.0 - x1 - x2));
          const double w = wt * x1;
          const double a = x1 + x2 + x3;       // full
          const double b = x1 * (1.0 - eta);   // shortened
          emit(w, a, x1 + x2, b + x2 + x3, x2);
          emit(w, b + x2 + x3, b + x2, a, x2);
          emit(w, a, x1 * eta + x2, x2 + x3, x2);
          emit(w, b + x2 + x3, x2, a, x1 + x2);
          emit(w, a, x2, b + x2 + x3, b + x2);
          emit(w, x2 + x3, x2, a, x1 * eta + x2);
        }
  return r;
}

}  // namespace h2hip
