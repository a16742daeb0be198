/* Oracle (test use only): Gauss-Legendre rules and the triangle-pair rules
 * of the reference's singular quadrature.  See oracle.h. */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Gauss-Legendre on [-1, 1].  The reference computes the nodes as
 * eigenvalues of the Jacobi matrix (Library/gaussquad.c:12-37); here the
 * same nodes come from Newton's method on the Legendre recurrence
 * (n+1) P_{n+1} = (2n+1) x P_n - n P_{n-1}; both agree to rounding. */
void
orc_gauss(unsigned m, double *x, double *w)
{
  for (unsigned i = 0; i < m; i++) {
    /* i-th root counted from the left; Chebyshev-like initial guess */
    double z = -cos(M_PI * (i + 0.75) / (m + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1.0, p1 = z;
      for (unsigned n = 1; n < m; n++) {
        double p2 = ((2.0 * n + 1.0) * z * p1 - n * p0) / (n + 1.0);
        p0 = p1;
        p1 = p2;
      }
      if (m == 1) {
        p1 = z;
        p0 = 1.0;
      }
      /* P_m'(z) = m (z P_m - P_{m-1}) / (z^2 - 1) */
      dp = m * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (fabs(dz) < 1e-17)
        break;
    }
    /* final evaluation at the converged node */
    double p0 = 1.0, p1 = z;
    for (unsigned n = 1; n < m; n++) {
      double p2 = ((2.0 * n + 1.0) * z * p1 - n * p0) / (n + 1.0);
      p0 = p1;
      p1 = p2;
    }
    dp = m * (z * p1 - p0) / (z * z - 1.0);
    x[i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

static void
rule_alloc(orc_rule *r, unsigned n, unsigned wcols, unsigned xcols)
{
  r->n = n;
  r->x = (double *) calloc((size_t) xcols * n, sizeof(double));
  r->y = (double *) calloc((size_t) xcols * n, sizeof(double));
  r->w = (double *) calloc((size_t) wcols * n, sizeof(double));
  r->base = 0.0;
}

static void
rule_free(orc_rule *r)
{
  free(r->x);
  free(r->y);
  free(r->w);
}

/* Single triangle: Library/singquad2d.c:25-47 (Duffy map of the square). */
static void
build_single(orc_rule *r, unsigned q, const double *xq, const double *wq)
{
  unsigned p = 0, n = r->n;
  for (unsigned i = 0; i < q; i++)
    for (unsigned j = 0; j < q; j++) {
      r->w[p + 3 * n] = wq[i] * wq[j] * xq[i];
      r->x[p] = xq[i];
      r->y[p] = xq[i] * xq[j];
      p++;
    }
}

/* Regular pair: tensor product of two single-triangle rules,
 * Library/singquad2d.c:49-78. */
static void
build_dist(orc_rule *r, unsigned q, const double *xq, const double *wq)
{
  unsigned p = 0, n = r->n;
  for (unsigned i = 0; i < q; i++)
    for (unsigned j = 0; j < q; j++)
      for (unsigned k = 0; k < q; k++)
        for (unsigned l = 0; l < q; l++) {
          r->w[p + 9 * n] = wq[i] * wq[j] * wq[k] * wq[l] * xq[i] * xq[k];
          r->x[p] = xq[i];
          r->x[p + n] = xq[i] * xq[j];
          r->y[p] = xq[k];
          r->y[p + n] = xq[k] * xq[l];
          p++;
        }
}

/* Common vertex: two subregions, Library/singquad2d.c:80-125. */
static void
build_vert(orc_rule *r, unsigned q, const double *xq, const double *wq)
{
  unsigned p = 0, n = r->n;
  for (unsigned i = 0; i < q; i++) {
    double e1 = xq[i], wi = wq[i];
    for (unsigned j = 0; j < q; j++) {
      double e2 = xq[j], wj = wq[j] * wi;
      for (unsigned k = 0; k < q; k++) {
        double e3 = xq[k], wk = wq[k] * wj;
        for (unsigned l = 0; l < q; l++) {
          double xi = xq[l], wl = wq[l] * wk;
          double ww = wl * xi * xi * xi * e1;

          r->w[p + 9 * n] = ww;
          r->x[p] = xi;
          r->x[p + n] = e2 * xi;
          r->y[p] = e1 * xi;
          r->y[p + n] = e1 * e3 * xi;
          p++;

          r->w[p + 9 * n] = ww;
          r->x[p] = e1 * xi;
          r->x[p + n] = e1 * e3 * xi;
          r->y[p] = xi;
          r->y[p + n] = e2 * xi;
          p++;
        }
      }
    }
  }
}

/* Common edge: six subregions, Library/singquad2d.c:127-201. */
static void
build_edge(orc_rule *r, unsigned q, const double *xq, const double *wq)
{
  unsigned p = 0, n = r->n;
  for (unsigned i = 0; i < q; i++) {
    double e1 = xq[i], wi = wq[i];
    for (unsigned j = 0; j < q; j++) {
      double e2 = xq[j], wj = wq[j] * wi;
      for (unsigned k = 0; k < q; k++) {
        double x1 = xq[k], wk = wq[k] * wj;
        for (unsigned l = 0; l < q; l++) {
          double x2 = xq[l] * (1.0 - x1);
          double wl = wq[l] * wk * (1.0 - x1);
          double ww = wl * x1 * x1 * e1;
          /* (x0, x1, y0, y1) of the six subregions */
          const double pts[6][4] = {
            {x1 + x2, x1 * e1 * e2, x1 * (1.0 - e1) + x2, -(x1 * (e1 - 1.0))},
            {x1 + x2, x1 * e1, x1 * (1.0 - e1 * e2) + x2,
             -(x1 * (e1 * e2 - 1.0))},
            {x1 + x2, x1, x1 * e1 + x2, x1 * e1 * e2},
            {x1 * (1.0 - e1 + e1 * e2) + x2, x1 * e1 * e2, x1 + x2, x1},
            {x1 * e1 * e2 + x2, x1 * e1 * e2, x1 + x2, x1 * e1},
            {x1 * e1 + x2, x1 * e1, x1 + x2, x1 * e1 * e2},
          };
          for (int s = 0; s < 6; s++) {
            r->w[p + 9 * n] = ww;
            r->x[p] = pts[s][0];
            r->x[p + n] = pts[s][1];
            r->y[p] = pts[s][2];
            r->y[p + n] = pts[s][3];
            p++;
          }
        }
      }
    }
  }
}

/* Identical triangles: six subregions, Library/singquad2d.c:203-281. */
static void
build_id(orc_rule *r, unsigned q, const double *xq, const double *wq)
{
  unsigned p = 0, n = r->n;
  for (unsigned i = 0; i < q; i++) {
    double eta = xq[i], wi = wq[i];
    for (unsigned j = 0; j < q; j++) {
      double x1 = xq[j], wj = wq[j] * wi;
      for (unsigned k = 0; k < q; k++) {
        double x2 = xq[k] * (1.0 - x1);
        double wk = wq[k] * wj * (1.0 - x1);
        for (unsigned l = 0; l < q; l++) {
          double x3 = xq[l] * (1.0 - x1 - x2);
          double wl = wq[l] * wk * (1.0 - x1 - x2);
          double ww = wl * x1;
          const double pts[6][4] = {
            {x1 + x2 + x3, x1 + x2, x1 * (1.0 - eta) + x2 + x3, x2},
            {x1 * (1.0 - eta) + x2 + x3, x1 * (1.0 - eta) + x2, x1 + x2 + x3,
             x2},
            {x1 + x2 + x3, x1 * eta + x2, x2 + x3, x2},
            {x1 * (1.0 - eta) + x2 + x3, x2, x1 + x2 + x3, x1 + x2},
            {x1 + x2 + x3, x2, x1 * (1.0 - eta) + x2 + x3,
             x1 * (1.0 - eta) + x2},
            {x2 + x3, x2, x1 + x2 + x3, x1 * eta + x2},
          };
          for (int s = 0; s < 6; s++) {
            r->w[p + 9 * n] = ww;
            r->x[p] = pts[s][0];
            r->x[p + n] = pts[s][1];
            r->y[p] = pts[s][2];
            r->y[p + n] = pts[s][3];
            p++;
          }
        }
      }
    }
  }
}

orc_sq *
orc_sq_new(unsigned q, unsigned q2)
{
  orc_sq *sq = (orc_sq *) calloc(1, sizeof(orc_sq));
  double *x = (double *) malloc(sizeof(double) * q);
  double *w = (double *) malloc(sizeof(double) * q);
  double *x2 = (double *) malloc(sizeof(double) * q2);
  double *w2 = (double *) malloc(sizeof(double) * q2);

  /* map [-1, 1] to [0, 1], Library/singquad2d.c:327-339 */
  orc_gauss(q, x, w);
  for (unsigned i = 0; i < q; i++) {
    x[i] = 0.5 + 0.5 * x[i];
    w[i] = w[i] * 0.5;
  }
  orc_gauss(q2, x2, w2);
  for (unsigned i = 0; i < q2; i++) {
    x2[i] = 0.5 + 0.5 * x2[i];
    w2[i] = w2[i] * 0.5;
  }
  unsigned nq = q * q * q * q, nq2 = q2 * q2 * q2 * q2;

  sq->q = q;
  sq->q2 = q2;
  rule_alloc(&sq->id, 6 * nq2, 10, 2);
  build_id(&sq->id, q2, x2, w2);
  rule_alloc(&sq->edge, 6 * nq2, 10, 2);
  build_edge(&sq->edge, q2, x2, w2);
  rule_alloc(&sq->vert, 2 * nq2, 10, 2);
  build_vert(&sq->vert, q2, x2, w2);
  rule_alloc(&sq->dist, nq, 10, 2);
  build_dist(&sq->dist, q, x, w);
  rule_alloc(&sq->single, q * q, 4, 1);
  build_single(&sq->single, q, x, w);

  free(x);
  free(w);
  free(x2);
  free(w2);
  return sq;
}

void
orc_sq_del(orc_sq *sq)
{
  if (!sq)
    return;
  rule_free(&sq->id);
  rule_free(&sq->edge);
  rule_free(&sq->vert);
  rule_free(&sq->dist);
  rule_free(&sq->single);
  free(sq);
}

/* Linear basis functions on the reference triangle in the (x0, x1)
 * coordinates: phi_0 = 1 - x0, phi_1 = x0 - x1, phi_2 = x1. */
static double
lin_basis(int i, double a0, double a1)
{
  return i == 0 ? 1.0 - a0 : (i == 1 ? a0 - a1 : a1);
}

/* Library/singquad2d.c:456-477 (ll), 479-496 (cl), 498-514 (lc). */
static void
weight_pair(orc_rule *r, int basis_r, int basis_c)
{
  unsigned n = r->n;
  if (basis_r && basis_c) {
    int idx = 0;
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++, idx++)
        for (unsigned k = 0; k < n; k++)
          r->w[k + idx * n] = r->w[k + 9 * n]
            * lin_basis(i, r->x[k], r->x[k + n])
            * lin_basis(j, r->y[k], r->y[k + n]);
  }
  else if (basis_c) {
    for (int i = 0; i < 3; i++)
      for (unsigned k = 0; k < n; k++)
        r->w[k + i * n] = r->w[k + 9 * n]
          * lin_basis(i, r->y[k], r->y[k + n]);
  }
  else if (basis_r) {
    for (int i = 0; i < 3; i++)
      for (unsigned k = 0; k < n; k++)
        r->w[k + i * n] = r->w[k + 9 * n]
          * lin_basis(i, r->x[k], r->x[k + n]);
  }
}

void
orc_sq_weight(orc_sq *sq, int basis_r, int basis_c)
{
  if (!basis_r && !basis_c)
    return;
  weight_pair(&sq->id, basis_r, basis_c);
  weight_pair(&sq->edge, basis_r, basis_c);
  weight_pair(&sq->vert, basis_r, basis_c);
  weight_pair(&sq->dist, basis_r, basis_c);
  /* Library/singquad2d.c:516-530 */
  unsigned n = sq->single.n;
  for (int i = 0; i < 3; i++)
    for (unsigned k = 0; k < n; k++)
      sq->single.w[k + i * n] = sq->single.w[k + 3 * n]
        * lin_basis(i, sq->single.x[k], sq->single.y[k]);
}

unsigned
orc_select(const orc_sq *sq, const unsigned *tv, const unsigned *sv,
           unsigned *tp, unsigned *sp, const orc_rule **rule)
{
  unsigned p = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      p += (tv[i] == sv[j]);

  for (int i = 0; i < 3; i++)
    tp[i] = sp[i] = (unsigned) i;

  switch (p) {
  case 0:
    *rule = &sq->dist;
    return 0;
  case 1:
    *rule = &sq->vert;
    break;
  case 2:
    *rule = &sq->edge;
    break;
  case 3:
    *rule = &sq->id;
    return 3;
  default:
    abort();
  }

  /* shared vertices first, in the order of the first triangle, then the
   * remaining ones in their original order (Library/singquad2d.c:599-627) */
  unsigned c = 0;
  for (unsigned i = 0; i < 3; i++)
    for (unsigned j = 0; j < 3; j++)
      if (tv[i] == sv[j]) {
        tp[c] = i;
        sp[c] = j;
        c++;
        break;
      }
  unsigned ct = c, cs = c;
  for (unsigned i = 0; i < 3; i++) {
    unsigned j;
    for (j = 0; j < c && tv[i] != tv[tp[j]]; j++)
      ;
    if (j == c)
      tp[ct++] = i;
    for (j = 0; j < c && sv[i] != sv[sp[j]]; j++)
      ;
    if (j == c)
      sp[cs++] = i;
  }
  return p;
}

int
orc_sq_export(const orc_sq *sq, int cls, double *x, double *y, double *w,
              unsigned *n, double *base)
{
  const orc_rule *r;
  switch (cls) {
  case 0: r = &sq->dist; break;
  case 1: r = &sq->vert; break;
  case 2: r = &sq->edge; break;
  case 3: r = &sq->id; break;
  case 4: r = &sq->single; break;
  default: return -1;
  }
  *n = r->n;
  *base = r->base;
  if (x) {
    unsigned xc = (cls == 4) ? 1 : 2;
    unsigned wc = (cls == 4) ? 3 : 9;
    memcpy(x, r->x, sizeof(double) * xc * r->n);
    memcpy(y, r->y, sizeof(double) * xc * r->n);
    memcpy(w, r->w + (size_t) wc * r->n, sizeof(double) * r->n);
  }
  return 0;
}
