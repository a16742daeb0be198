/* Boundary data, right-hand sides and error functionals of the Laplace
 * Dirichlet problem with the piecewise linear basis.
 *
 * Restates integrate_bem3d_l_avector (Library/bem3d.c:8631-8702),
 * projectL2_bem3d_l_avector (Library/bem3d.c:9009-9088, CG on the mass
 * matrix), addeval_mass_linear_bem3d (Library/bem3d.c:8878-8908),
 * normL2_bem3d (Library/bem3d.c:8704-8751) and normL2diff_l_bem3d
 * (Library/bem3d.c:8954-9007), with the boundary data of
 * eval_dirichlet_fundamental_laplacebem3d and
 * eval_neumann_fundamental_laplacebem3d (Library/laplacebem3d.c:1390-1430)
 * for a point source at x0 (the reference fixes x0 = (1.2, 1.2, 1.2)).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "h2bem.h"
#include "gauss01.h"

/* f(x, n): 0 Dirichlet data 1/|x - x0|, 1 Neumann data
 * -<x - x0, n>/|x - x0|^3 */
static double
field_value(int kind, const double *x0, const double *x, const double *n)
{
  double d[3] = {x[0] - x0[0], x[1] - x0[1], x[2] - x0[2]};
  double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  if (kind == 0)
    return 1.0 / sqrt(r2);
  return -(d[0] * n[0] + d[1] * n[1] + d[2] * n[2]) / (r2 * sqrt(r2));
}

static unsigned
tri_rule(unsigned q, double *x0, double *x1, double *w)
{
  /* Gauss-Legendre on [0, 1], then the Duffy-collapsed tensor rule on the
   * reference triangle */
  double g[16], gw[16];
  h2_gauss01(q, g, gw);
  unsigned k = 0;
  for (unsigned i = 0; i < q; i++)
    for (unsigned j = 0; j < q; j++, k++) {
      x0[k] = g[i];
      x1[k] = g[i] * g[j];
      w[k] = gw[i] * gw[j] * g[i];
    }
  return k;
}

void
h2b_integrate_l(const h2b_surface *s, unsigned q, int kind, const double *x0,
                double *out)
{
  double rx0[256], rx1[256], rw[256];
  const unsigned nq = tri_rule(q, rx0, rx1, rw);
  memset(out, 0, sizeof(double) * s->vertices);
  for (uint32_t t = 0; t < s->triangles; t++) {
    const uint32_t *tv = s->t + 3 * t;
    const double *A = s->x + 3 * tv[0], *B = s->x + 3 * tv[1],
      *C = s->x + 3 * tv[2];
    double sum[3] = {0.0, 0.0, 0.0};
    for (unsigned k = 0; k < nq; k++) {
      double u = rx0[k], v = rx1[k], x[3];
      for (int d = 0; d < 3; d++)
        x[d] = A[d] * (1.0 - u) + B[d] * (u - v) + C[d] * v;
      double f = rw[k] * field_value(kind, x0, x, s->n + 3 * t);
      sum[0] += f * (1.0 - u);
      sum[1] += f * (u - v);
      sum[2] += f * v;
    }
    for (int l = 0; l < 3; l++)
      out[tv[l]] += sum[l] * s->g[t];
  }
}

void
h2b_mass_l_addeval(const h2b_surface *s, double alpha, const double *x,
                   double *y)
{
  for (uint32_t t = 0; t < s->triangles; t++) {
    const uint32_t *tv = s->t + 3 * t;
    const double gt = s->g[t];
    for (int l = 0; l < 3; l++) {
      double acc = 0.0;
      for (int k = 0; k < 3; k++)
        acc += x[tv[k]] * gt * (k == l ? 1.0 / 12.0 : 1.0 / 24.0);
      y[tv[l]] += alpha * acc;
    }
  }
}

int
h2b_projectL2_l(const h2b_surface *s, unsigned q, int kind, const double *x0,
                double *out)
{
  const uint32_t n = s->vertices;
  double *b = (double *) malloc(sizeof(double) * n);
  double *r = (double *) malloc(sizeof(double) * n);
  double *p = (double *) malloc(sizeof(double) * n);
  double *ap = (double *) malloc(sizeof(double) * n);
  h2b_integrate_l(s, q, kind, x0, b);
  /* CG from x = 0 on the (symmetric positive definite) mass matrix to
   * |r| <= 1e-15 |b|, at most 500 steps as the reference */
  memset(out, 0, sizeof(double) * n);
  memcpy(r, b, sizeof(double) * n);
  memcpy(p, b, sizeof(double) * n);
  double rr = 0.0, bb;
  for (uint32_t i = 0; i < n; i++)
    rr += r[i] * r[i];
  bb = rr;
  int it;
  for (it = 0; it < 500 && rr > 1e-30 * bb; it++) {
    memset(ap, 0, sizeof(double) * n);
    h2b_mass_l_addeval(s, 1.0, p, ap);
    double pap = 0.0;
    for (uint32_t i = 0; i < n; i++)
      pap += p[i] * ap[i];
    double a = rr / pap, rr2 = 0.0;
    for (uint32_t i = 0; i < n; i++) {
      out[i] += a * p[i];
      r[i] -= a * ap[i];
      rr2 += r[i] * r[i];
    }
    double beta = rr2 / rr;
    rr = rr2;
    for (uint32_t i = 0; i < n; i++)
      p[i] = r[i] + beta * p[i];
  }
  free(b);
  free(r);
  free(p);
  free(ap);
  return it;
}

double
h2b_normL2diff_l(const h2b_surface *s, unsigned q, const double *xv,
                 int kind, const double *x0)
{
  double rx0[256], rx1[256], rw[256];
  const unsigned nq = tri_rule(q, rx0, rx1, rw);
  double norm = 0.0;
  for (uint32_t t = 0; t < s->triangles; t++) {
    const uint32_t *tv = s->t + 3 * t;
    const double *A = s->x + 3 * tv[0], *B = s->x + 3 * tv[1],
      *C = s->x + 3 * tv[2];
    double sum = 0.0;
    for (unsigned k = 0; k < nq; k++) {
      double u = rx0[k], v = rx1[k], x[3];
      for (int d = 0; d < 3; d++)
        x[d] = A[d] * (1.0 - u) + B[d] * (u - v) + C[d] * v;
      double bf = xv[tv[0]] * (1.0 - u) + xv[tv[1]] * (u - v) + xv[tv[2]] * v;
      double e = field_value(kind, x0, x, s->n + 3 * t) - bf;
      sum += rw[k] * e * e;
    }
    norm += s->g[t] * sum;
  }
  return sqrt(norm);
}

double
h2b_normL2(const h2b_surface *s, unsigned q, int kind, const double *x0)
{
  double rx0[256], rx1[256], rw[256];
  const unsigned nq = tri_rule(q, rx0, rx1, rw);
  double norm = 0.0;
  for (uint32_t t = 0; t < s->triangles; t++) {
    const uint32_t *tv = s->t + 3 * t;
    const double *A = s->x + 3 * tv[0], *B = s->x + 3 * tv[1],
      *C = s->x + 3 * tv[2];
    double sum = 0.0;
    for (unsigned k = 0; k < nq; k++) {
      double u = rx0[k], v = rx1[k], x[3];
      for (int d = 0; d < 3; d++)
        x[d] = A[d] * (1.0 - u) + B[d] * (u - v) + C[d] * v;
      double f = field_value(kind, x0, x, s->n + 3 * t);
      sum += rw[k] * f * f;
    }
    norm += s->g[t] * sum;
  }
  return sqrt(norm);
}

/* ---- Helmholtz point source, piecewise constant basis (complex) ----------
 * g(x, n) = e^{ik r} / r (kind 0, rhs_dirichlet_point_helmholtzbem3d) or
 * e^{ik r} / r^3 (ik r - 1) <n, x - x0> (kind 1,
 * rhs_neumann_point_helmholtzbem3d), r = |x - x0|
 * (Library/helmholtzbem3d.c:1521-1578), real k. */
static void
helm_value(int kind, double k, const double *x0, const double *x,
           const double *n, double *re, double *im)
{
  const double d[3] = {x[0] - x0[0], x[1] - x0[1], x[2] - x0[2]};
  const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  const double rs = 1.0 / sqrt(r2), r = r2 * rs;
  double a = cos(k * r) * rs, b = sin(k * r) * rs;
  if (kind == 1) {
    /* (a + ib) (ikr - 1) <n, d> / r^2 */
    const double f = (n[0] * d[0] + n[1] * d[1] + n[2] * d[2]) * rs * rs;
    const double kr = k * r;
    const double na = -a - b * kr, nb = a * kr - b;
    a = na * f;
    b = nb * f;
  }
  *re = a;
  *im = b;
}

void
h2b_projectL2_c_helmholtz(const h2b_surface *s, unsigned q, int kind,
                          double k, const double *x0, double *out)
{
  double rx0[256], rx1[256], rw[256];
  const unsigned nq = tri_rule(q, rx0, rx1, rw);
  for (uint32_t t = 0; t < s->triangles; t++) {
    const uint32_t *tv = s->t + 3 * t;
    const double *A = s->x + 3 * tv[0], *B = s->x + 3 * tv[1],
      *C = s->x + 3 * tv[2];
    double sre = 0.0, sim = 0.0;
    for (unsigned j = 0; j < nq; j++) {
      const double u = rx0[j], v = rx1[j];
      double x[3], re, im;
      for (int d = 0; d < 3; d++)
        x[d] = A[d] * (1.0 - u) + B[d] * (u - v) + C[d] * v;
      helm_value(kind, k, x0, x, s->n + 3 * t, &re, &im);
      sre += rw[j] * re;
      sim += rw[j] * im;
    }
    /* the rule's weights sum to 1/2: twice the sum is the mean value */
    out[2 * t] = 2.0 * sre;
    out[2 * t + 1] = 2.0 * sim;
  }
}

void
h2b_normL2_c_helmholtz(const h2b_surface *s, unsigned q, const double *xv,
                       int kind, double k, const double *x0, double *out)
{
  double rx0[256], rx1[256], rw[256];
  const unsigned nq = tri_rule(q, rx0, rx1, rw);
  double diff = 0.0, norm = 0.0;
  for (uint32_t t = 0; t < s->triangles; t++) {
    const uint32_t *tv = s->t + 3 * t;
    const double *A = s->x + 3 * tv[0], *B = s->x + 3 * tv[1],
      *C = s->x + 3 * tv[2];
    double sd = 0.0, sn = 0.0;
    for (unsigned j = 0; j < nq; j++) {
      const double u = rx0[j], v = rx1[j];
      double x[3], re, im;
      for (int d = 0; d < 3; d++)
        x[d] = A[d] * (1.0 - u) + B[d] * (u - v) + C[d] * v;
      helm_value(kind, k, x0, x, s->n + 3 * t, &re, &im);
      const double er = re - (xv ? xv[2 * t] : 0.0);
      const double ei = im - (xv ? xv[2 * t + 1] : 0.0);
      sd += rw[j] * (er * er + ei * ei);
      sn += rw[j] * (re * re + im * im);
    }
    diff += s->g[t] * sd;
    norm += s->g[t] * sn;
  }
  out[0] = sqrt(diff);
  out[1] = sqrt(norm);
}

void
h2b_integrate_c_helmholtz(const h2b_surface *s, unsigned q, int kind,
                          double k, const double *x0, double *out)
{
  double rx0[256], rx1[256], rw[256];
  const unsigned nq = tri_rule(q, rx0, rx1, rw);
  for (uint32_t t = 0; t < s->triangles; t++) {
    const uint32_t *tv = s->t + 3 * t;
    const double *A = s->x + 3 * tv[0], *B = s->x + 3 * tv[1],
      *C = s->x + 3 * tv[2];
    double sre = 0.0, sim = 0.0;
    for (unsigned j = 0; j < nq; j++) {
      const double u = rx0[j], v = rx1[j];
      double x[3], re, im;
      for (int d = 0; d < 3; d++)
        x[d] = A[d] * (1.0 - u) + B[d] * (u - v) + C[d] * v;
      helm_value(kind, k, x0, x, s->n + 3 * t, &re, &im);
      sre += rw[j] * re;
      sim += rw[j] * im;
    }
    out[2 * t] = sre * s->g[t];
    out[2 * t + 1] = sim * s->g[t];
  }
}

void
h2b_integrate_l_helmholtz(const h2b_surface *s, unsigned q, int kind,
                          double k, const double *x0, double *out)
{
  double rx0[256], rx1[256], rw[256];
  const unsigned nq = tri_rule(q, rx0, rx1, rw);
  memset(out, 0, sizeof(double) * 2 * s->vertices);
  for (uint32_t t = 0; t < s->triangles; t++) {
    const uint32_t *tv = s->t + 3 * t;
    const double *A = s->x + 3 * tv[0], *B = s->x + 3 * tv[1],
      *C = s->x + 3 * tv[2];
    double sre[3] = {0.0, 0.0, 0.0}, sim[3] = {0.0, 0.0, 0.0};
    for (unsigned j = 0; j < nq; j++) {
      const double u = rx0[j], v = rx1[j];
      const double phi[3] = {1.0 - u, u - v, v};
      double x[3], re, im;
      for (int d = 0; d < 3; d++)
        x[d] = A[d] * (1.0 - u) + B[d] * (u - v) + C[d] * v;
      helm_value(kind, k, x0, x, s->n + 3 * t, &re, &im);
      for (int l = 0; l < 3; l++) {
        sre[l] += rw[j] * phi[l] * re;
        sim[l] += rw[j] * phi[l] * im;
      }
    }
    for (int l = 0; l < 3; l++) {
      out[2 * tv[l]] += sre[l] * s->g[t];
      out[2 * tv[l] + 1] += sim[l] * s->g[t];
    }
  }
}
