/* Oracle (test use only): near-field Galerkin quadrature of the reference,
 * restated loop by loop.  See oracle.h. */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* 1 / (4 pi), Library/laplacebem3d.c:33 and Library/helmholtzbem3d.c:24 */
#define ORC_KERNEL_CONST 0.0795774715459476679

struct orc_bem {
  int kernel, op, basis_r, basis_c;
  unsigned nv, nt;
  double *x;        /* 3 nv */
  unsigned *t;      /* 3 nt */
  double *n;        /* 3 nt, unit normals */
  double *g;        /* nt, twice the area */
  unsigned *v2t_ptr, *v2t; /* triangles of each vertex, decreasing order */
  orc_sq *sq;
  double complex alpha, k;
  double mass[9];
};

/* Library/surface3d.c:51-108 (normals and Gram determinants only) */
static void
prepare_surface(orc_bem *b)
{
  for (unsigned i = 0; i < b->nt; i++) {
    const double *p0 = b->x + 3 * b->t[3 * i];
    const double *p1 = b->x + 3 * b->t[3 * i + 1];
    const double *p2 = b->x + 3 * b->t[3 * i + 2];
    double dx[3], dy[3], nn[3];
    for (int j = 0; j < 3; j++) {
      dx[j] = p1[j] - p0[j];
      dy[j] = p2[j] - p0[j];
    }
    nn[0] = dx[1] * dy[2] - dx[2] * dy[1];
    nn[1] = dx[2] * dy[0] - dx[0] * dy[2];
    nn[2] = dx[0] * dy[1] - dx[1] * dy[0];
    double norm = sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
    b->g[i] = norm;
    for (int j = 0; j < 3; j++)
      b->n[3 * i + j] = nn[j] / norm;
  }
}

/* Vertex-to-triangle map, Library/bem3d.c:779-804: the reference prepends
 * to a list per vertex, so each list runs in decreasing triangle order. */
static void
build_v2t(orc_bem *b)
{
  unsigned *cnt = (unsigned *) calloc(b->nv + 1, sizeof(unsigned));
  for (unsigned i = 0; i < 3 * b->nt; i++)
    cnt[b->t[i] + 1]++;
  for (unsigned v = 0; v < b->nv; v++)
    cnt[v + 1] += cnt[v];
  b->v2t_ptr = cnt;
  b->v2t = (unsigned *) malloc(sizeof(unsigned) * 3 * b->nt + 1);
  unsigned *fill = (unsigned *) calloc(b->nv, sizeof(unsigned));
  for (unsigned i = b->nt; i-- > 0;)
    for (int j = 0; j < 3; j++) {
      unsigned v = b->t[3 * i + j];
      b->v2t[cnt[v] + fill[v]++] = i;
    }
  free(fill);
}

orc_bem *
orc_bem_new(int kernel, int op, int basis_r, int basis_c, unsigned nv,
            const double *x, unsigned nt, const unsigned *t, unsigned q,
            unsigned q2, const double *alpha, const double *k)
{
  orc_bem *b = (orc_bem *) calloc(1, sizeof(orc_bem));
  b->kernel = kernel;
  b->op = op;
  b->basis_r = basis_r;
  b->basis_c = basis_c;
  b->nv = nv;
  b->nt = nt;
  b->x = (double *) malloc(sizeof(double) * 3 * nv);
  memcpy(b->x, x, sizeof(double) * 3 * nv);
  b->t = (unsigned *) malloc(sizeof(unsigned) * 3 * nt);
  memcpy(b->t, t, sizeof(unsigned) * 3 * nt);
  b->n = (double *) malloc(sizeof(double) * 3 * nt);
  b->g = (double *) malloc(sizeof(double) * nt);
  prepare_surface(b);
  build_v2t(b);
  b->sq = orc_sq_new(q, q2);
  orc_sq_weight(b->sq, basis_r, basis_c);
  /* SLP has alpha = 0 (Library/laplacebem3d.c:857) */
  b->alpha = (op == 0) ? 0.0 : (alpha[0] + alpha[1] * I);
  b->k = (k != NULL) ? (k[0] + k[1] * I) : 0.0;
  /* mass matrices of the basis pairs, Library/laplacebem3d.c:928-994 */
  if (basis_r && basis_c) {
    for (int i = 0; i < 9; i++)
      b->mass[i] = (i % 4 == 0) ? 1.0 / 12.0 : 1.0 / 24.0;
  }
  else if (basis_r || basis_c) {
    b->mass[0] = b->mass[1] = b->mass[2] = 1.0 / 6.0;
  }
  return b;
}

void
orc_bem_del(orc_bem *b)
{
  if (!b)
    return;
  free(b->x);
  free(b->t);
  free(b->n);
  free(b->g);
  free(b->v2t_ptr);
  free(b->v2t);
  orc_sq_del(b->sq);
  free(b);
}

int
orc_bem_is_complex(const orc_bem *b)
{
  return b->kernel == 1;
}

const double *
orc_bem_normals(const orc_bem *b)
{
  return b->n;
}

const double *
orc_bem_gram(const orc_bem *b)
{
  return b->g;
}

/* ---- kernel functions -------------------------------------------------- */

/* Laplace SLP 1/|x-y|, DLP <x-y, n_y>/|x-y|^3, ADLP <y-x, n_x>/|x-y|^3
 * (Library/laplacebem3d.c:35-160) and Helmholtz SLP e^{ik|x-y|}/|x-y|, DLP
 * e^{ikr}(1-ikr)<x-y,n_y>/r^3, ADLP e^{ikr}(1-ikr)<y-x,n_x>/r^3
 * (Library/helmholtzbem3d.c:26-240).  nx, ny: the normals the integrators
 * pass (those of the row and the column triangle).  The 1/(4 pi) is
 * applied by the integrators. */
static double complex
kernel_eval(const orc_bem *b, const double *x, const double *y,
            const double *nx, const double *ny)
{
  double d[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
  double norm2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double rnorm = 1.0 / sqrt(norm2);
  double nrm[3];

  if (b->op == 2) {
    /* the adjoint double layer: <y - x, n_x> = <x - y, -n_x> */
    for (int j = 0; j < 3; j++)
      nrm[j] = -nx[j];
    ny = nrm;
  }
  if (b->kernel == 0) {
    if (b->op == 0)
      return rnorm;
    double r3 = rnorm * rnorm * rnorm;
    return (d[0] * ny[0] + d[1] * ny[1] + d[2] * ny[2]) * r3;
  }
  double kr = creal(b->k), ki = -cimag(b->k);
  if (b->op == 0) {
    double norm = norm2 * rnorm;
    if (ki != 0.0)
      rnorm *= exp(ki * norm);
    norm = kr * norm;
    return rnorm * cos(norm) + (sin(norm) * rnorm) * I;
  }
  double re = (rnorm * rnorm * rnorm)
    * (d[0] * ny[0] + d[1] * ny[1] + d[2] * ny[2]);
  double norm = norm2 * rnorm;
  if (ki != 0.0)
    re *= exp(ki * norm);
  norm = kr * norm;
  double s = sin(norm), c = cos(norm);
  return (c + norm * s) * re + ((s - c * norm) * re) * I;
}

/* Kernel values at all points of the rule for the pair (tt, ss).  With
 * far != 0 the regular rule and the identity permutation are used.  Writes
 * the permuted vertex numbers to ttp / ssp (the loop bodies of
 * Library/bem3d.c:1186-1226 and 3126-3167). */
static const orc_rule *
pair_values_swap(const orc_bem *b, int far, unsigned tt, unsigned ss,
                 double complex *quad, unsigned *ttp, unsigned *ssp,
                 int swap)
{
  const unsigned *tri_t = b->t + 3 * tt, *tri_s = b->t + 3 * ss;
  unsigned tp[3] = {0, 1, 2}, sp[3] = {0, 1, 2};
  const orc_rule *r;

  if (far)
    r = &b->sq->dist;
  else
    orc_select(b->sq, tri_t, tri_s, tp, sp, &r);
  for (int i = 0; i < 3; i++) {
    ttp[i] = tri_t[tp[i]];
    ssp[i] = tri_s[sp[i]];
  }
  const unsigned *xv = swap ? ssp : ttp, *yv = swap ? ttp : ssp;
  const double *At = b->x + 3 * xv[0], *Bt = b->x + 3 * xv[1],
    *Ct = b->x + 3 * xv[2];
  const double *As = b->x + 3 * yv[0], *Bs = b->x + 3 * yv[1],
    *Cs = b->x + 3 * yv[2];
  const double *ny = b->n + 3 * ss, *nx = b->n + 3 * tt;
  unsigned n = r->n;
  for (unsigned q = 0; q < n; q++) {
    double tx = r->x[q], sx = r->x[q + n], ty = r->y[q], sy = r->y[q + n];
    double Ax = 1.0 - tx, Bx = tx - sx, Cx = sx;
    double Ay = 1.0 - ty, By = ty - sy, Cy = sy;
    double xx[3], yy[3];
    for (int j = 0; j < 3; j++) {
      xx[j] = At[j] * Ax + Bt[j] * Bx + Ct[j] * Cx;
      yy[j] = As[j] * Ay + Bs[j] * By + Cs[j] * Cy;
    }
    quad[q] = kernel_eval(b, xx, yy, nx, ny);
  }
  return r;
}

static const orc_rule *
pair_values(const orc_bem *b, int far, unsigned tt, unsigned ss,
            double complex *quad, unsigned *ttp, unsigned *ssp)
{
  return pair_values_swap(b, far, tt, ss, quad, ttp, ssp, 0);
}

static double complex
wsum(const orc_rule *r, int col, const double complex *quad)
{
  double complex res = r->base;
  const double *w = r->w + (size_t) col * r->n;
  for (unsigned q = 0; q < r->n; q++)
    res += w[q] * quad[q];
  return res;
}

/* ---- support-triangle lists of linear basis functions ------------------- */

/* The reference collects the support triangles of a set of vertices into a
 * list, prepending each triangle when first met, and records for every
 * triangle the block-local numbers of its vertices (Library/bem3d.c:3071-
 * 3116).  Here: tri[] in final list order, and for each list entry up to
 * three block-local vertex numbers (local[3*k+...], count nloc[k]). */
typedef struct {
  unsigned n;
  unsigned *tri;
  unsigned *loc;   /* 3 per triangle: block-local vertex numbers */
  unsigned *nloc;
} supp_list;

static void
supp_build(const orc_bem *b, const unsigned *idx, unsigned cnt,
           supp_list *L)
{
  int *pos = (int *) malloc(sizeof(int) * b->nt);
  for (unsigned i = 0; i < b->nt; i++)
    pos[i] = -1;
  unsigned cap = 6 * cnt + 16, n = 0;
  unsigned *tri = (unsigned *) malloc(sizeof(unsigned) * cap);
  unsigned *loc = (unsigned *) malloc(sizeof(unsigned) * 3 * cap);
  unsigned *nloc = (unsigned *) calloc(cap, sizeof(unsigned));
  for (unsigned i = 0; i < cnt; i++) {
    unsigned ii = idx ? idx[i] : i;
    for (unsigned p = b->v2t_ptr[ii]; p < b->v2t_ptr[ii + 1]; p++) {
      unsigned vv = b->v2t[p];
      if (pos[vv] < 0) {
        if (n == cap) {
          cap *= 2;
          tri = (unsigned *) realloc(tri, sizeof(unsigned) * cap);
          loc = (unsigned *) realloc(loc, sizeof(unsigned) * 3 * cap);
          nloc = (unsigned *) realloc(nloc, sizeof(unsigned) * cap);
        }
        pos[vv] = (int) n;
        tri[n] = vv;
        nloc[n] = 0;
        n++;
      }
      unsigned k = (unsigned) pos[vv];
      if (nloc[k] < 3)
        loc[3 * k + nloc[k]++] = i;
    }
  }
  /* prepend order: the last triangle met comes first */
  L->n = n;
  L->tri = (unsigned *) malloc(sizeof(unsigned) * (n + 1));
  L->loc = (unsigned *) malloc(sizeof(unsigned) * 3 * (n + 1));
  L->nloc = (unsigned *) malloc(sizeof(unsigned) * (n + 1));
  for (unsigned k = 0; k < n; k++) {
    unsigned src = n - 1 - k;
    L->tri[k] = tri[src];
    L->nloc[k] = nloc[src];
    for (unsigned j = 0; j < nloc[src]; j++)
      L->loc[3 * k + j] = loc[3 * src + j];
  }
  free(pos);
  free(tri);
  free(loc);
  free(nloc);
}

static void
supp_free(supp_list *L)
{
  free(L->tri);
  free(L->loc);
  free(L->nloc);
}

/* ---- the integrators ----------------------------------------------------- */

static void
store(const orc_bem *b, void *N, size_t pos, double complex v, int add)
{
  if (b->kernel == 0) {
    double *a = (double *) N;
    a[pos] = add ? a[pos] + creal(v) : creal(v);
  }
  else {
    double complex *a = (double complex *) N;
    a[pos] = add ? a[pos] + v : v;
  }
}

static void
clear_window(const orc_bem *b, void *N, size_t ld, unsigned r, unsigned c)
{
  size_t es = b->kernel == 0 ? sizeof(double) : sizeof(double complex);
  for (unsigned j = 0; j < c; j++)
    memset((char *) N + es * ld * j, 0, es * r);
}

/* Library/bem3d.c:1150-1247 (near) and 1438-1528 (far) */
static void
cc(const orc_bem *b, int far, const unsigned *ridx, unsigned rows,
   const unsigned *cidx, unsigned cols, int ntrans, void *N, size_t ld)
{
  for (unsigned s = 0; s < cols; s++) {
    unsigned ttp[3], ssp[3];
    unsigned ss = cidx ? cidx[s] : s;
    double factor = b->g[ss] * ORC_KERNEL_CONST;
    double complex *qd =
      (double complex *) malloc(sizeof(double complex) * b->sq->id.n);
    for (unsigned t = 0; t < rows; t++) {
      unsigned tt = ridx ? ridx[t] : t;
      double factor2 = factor * b->g[tt];
      const orc_rule *r = pair_values(b, far, tt, ss, qd, ttp, ssp);
      double complex sum = wsum(r, 9, qd);
      if (ntrans)
        store(b, N, s + t * ld, conj(sum) * factor2, 0);
      else
        store(b, N, t + s * ld, sum * factor2, 0);
      if (!far && b->alpha != 0.0 && tt == ss && t < cols) {
        /* the reference adds the identity term at position (t, t); for
         * t >= cols that position lies outside the window and is skipped
         * here (the reference writes outside the window there) */
        double complex a = ntrans ? conj(b->alpha) : b->alpha;
        store(b, N, t + t * ld, 0.5 * a * b->g[tt], 1);
      }
    }
    free(qd);
  }
}

/* Library/bem3d.c:3038-3250 (near) and 3264-3436 (far) */
static void
ll(const orc_bem *b, int far, const unsigned *ridx, unsigned rows,
   const unsigned *cidx, unsigned cols, int ntrans, void *N, size_t ld)
{
  supp_list R, C;
  double complex *quad =
    (double complex *) malloc(sizeof(double complex) * b->sq->id.n);
  clear_window(b, N, ld, ntrans ? cols : rows, ntrans ? rows : cols);
  supp_build(b, ridx, rows, &R);
  supp_build(b, cidx, cols, &C);
  for (unsigned s = 0; s < C.n; s++) {
    unsigned ss = C.tri[s];
    double factor = b->g[ss] * ORC_KERNEL_CONST;
    for (unsigned t = 0; t < R.n; t++) {
      unsigned tt = R.tri[t], ttp[3], ssp[3];
      double factor2 = factor * b->g[tt];
      const orc_rule *r = pair_values(b, far, tt, ss, quad, ttp, ssp);
      for (unsigned a = 0; a < C.nloc[s]; a++) {
        unsigned j = C.loc[3 * s + a];
        unsigned jj = cidx ? cidx[j] : j;
        for (unsigned k = 0; k < 3; k++) {
          if (jj != ssp[k])
            continue;
          for (unsigned c = 0; c < R.nloc[t]; c++) {
            unsigned i = R.loc[3 * t + c];
            unsigned ii = ridx ? ridx[i] : i;
            for (unsigned l = 0; l < 3; l++) {
              if (ii != ttp[l])
                continue;
              double complex res = wsum(r, l + 3 * k, quad);
              if (ntrans)
                store(b, N, j + i * ld, conj(res * factor2), 1);
              else
                store(b, N, i + j * ld, res * factor2, 1);
            }
          }
        }
      }
      if (!far && b->alpha != 0.0 && tt == ss) {
        const unsigned *tri = b->t + 3 * tt;
        /* factor2 is a real variable in the reference's ll integrator
         * (Library/bem3d.c:3061-3062): only Re(alpha) reaches the mass
         * term */
        double complex f2 = creal(b->alpha) * b->g[tt];
        for (unsigned a = 0; a < C.nloc[s]; a++) {
          unsigned j = C.loc[3 * s + a];
          unsigned jj = cidx ? cidx[j] : j;
          for (unsigned k = 0; k < 3; k++) {
            if (jj != tri[k])
              continue;
            for (unsigned c = 0; c < R.nloc[t]; c++) {
              unsigned i = R.loc[3 * t + c];
              unsigned ii = ridx ? ridx[i] : i;
              for (unsigned l = 0; l < 3; l++) {
                if (ii != tri[l])
                  continue;
                double complex v = b->mass[l + 3 * k] * f2;
                if (ntrans)
                  store(b, N, j + i * ld, conj(v), 1);
                else
                  store(b, N, i + j * ld, v, 1);
              }
            }
          }
        }
      }
    }
  }
  supp_free(&R);
  supp_free(&C);
  free(quad);
}

/* Library/bem3d.c:1798-1971 (near) and 1985-2127 (far).  As in the
 * reference, the transposed store is not conjugated. */
static void
cl(const orc_bem *b, int far, const unsigned *ridx, unsigned rows,
   const unsigned *cidx, unsigned cols, int ntrans, void *N, size_t ld)
{
  supp_list C;
  double complex *quad =
    (double complex *) malloc(sizeof(double complex) * b->sq->id.n);
  clear_window(b, N, ld, ntrans ? cols : rows, ntrans ? rows : cols);
  supp_build(b, cidx, cols, &C);
  for (unsigned s = 0; s < C.n; s++) {
    unsigned ss = C.tri[s];
    double factor = b->g[ss] * ORC_KERNEL_CONST;
    for (unsigned t = 0; t < rows; t++) {
      unsigned tt = ridx ? ridx[t] : t, ttp[3], ssp[3];
      double factor2 = factor * b->g[tt];
      const orc_rule *r = pair_values(b, far, tt, ss, quad, ttp, ssp);
      for (unsigned a = 0; a < C.nloc[s]; a++) {
        unsigned j = C.loc[3 * s + a];
        unsigned jj = cidx ? cidx[j] : j;
        for (unsigned k = 0; k < 3; k++) {
          if (jj != ssp[k])
            continue;
          double complex res = wsum(r, k, quad);
          if (ntrans)
            store(b, N, j + t * ld, res * factor2, 1);
          else
            store(b, N, t + j * ld, res * factor2, 1);
        }
      }
      if (!far && b->alpha != 0.0 && tt == ss) {
        const unsigned *tri = b->t + 3 * ss;
        /* real factor2 in the reference's cl integrator
         * (Library/bem3d.c:1822-1823): Re(alpha) only */
        double complex f2 = creal(b->alpha) * b->g[tt];
        for (unsigned a = 0; a < C.nloc[s]; a++) {
          unsigned j = C.loc[3 * s + a];
          unsigned jj = cidx ? cidx[j] : j;
          for (unsigned k = 0; k < 3; k++) {
            if (jj != tri[k])
              continue;
            if (ntrans)
              store(b, N, j + t * ld, f2 * b->mass[k], 1);
            else
              store(b, N, t + j * ld, f2 * b->mass[k], 1);
          }
        }
      }
    }
  }
  supp_free(&C);
  free(quad);
}

/* Library/bem3d.c:2394-2571 (near) and 2581-2723 (far).  As in the
 * reference, the transposed store is not conjugated, the identity term
 * is only added when the block has more than three columns (the reference
 * tests a loop variable that equals 3 at that point), and the far variant
 * maps the first rule coordinates onto the column triangle and the second
 * onto the row triangle (Library/bem3d.c:2637-2649 assign the row
 * triangle's vertices to A_s, B_s, C_s). */
static void
lc(const orc_bem *b, int far, const unsigned *ridx, unsigned rows,
   const unsigned *cidx, unsigned cols, int ntrans, void *N, size_t ld)
{
  supp_list R;
  double complex *quad =
    (double complex *) malloc(sizeof(double complex) * b->sq->id.n);
  clear_window(b, N, ld, ntrans ? cols : rows, ntrans ? rows : cols);
  supp_build(b, ridx, rows, &R);
  for (unsigned t = 0; t < R.n; t++) {
    unsigned tt = R.tri[t];
    double factor = b->g[tt] * ORC_KERNEL_CONST;
    for (unsigned s = 0; s < cols; s++) {
      unsigned ss = cidx ? cidx[s] : s, ttp[3], ssp[3];
      double factor2 = factor * b->g[ss];
      const orc_rule *r = pair_values_swap(b, far, tt, ss, quad, ttp, ssp,
                                           far);
      for (unsigned a = 0; a < R.nloc[t]; a++) {
        unsigned i = R.loc[3 * t + a];
        unsigned ii = ridx ? ridx[i] : i;
        for (unsigned l = 0; l < 3; l++) {
          if (ii != ttp[l])
            continue;
          double complex res = wsum(r, l, quad);
          if (ntrans)
            store(b, N, s + i * ld, res * factor2, 1);
          else
            store(b, N, i + s * ld, res * factor2, 1);
        }
      }
      if (!far && b->alpha != 0.0 && tt == ss && cols > 3) {
        const unsigned *tri = b->t + 3 * tt;
        /* real factor2 in the reference's lc integrator
         * (Library/bem3d.c:2418-2419): Re(alpha) only */
        double complex f2 = creal(b->alpha) * b->g[tt];
        for (unsigned a = 0; a < R.nloc[t]; a++) {
          unsigned i = R.loc[3 * t + a];
          unsigned ii = ridx ? ridx[i] : i;
          for (unsigned l = 0; l < 3; l++) {
            if (ii != tri[l])
              continue;
            if (ntrans)
              store(b, N, s + i * ld, f2 * b->mass[l], 1);
            else
              store(b, N, i + s * ld, f2 * b->mass[l], 1);
          }
        }
      }
    }
  }
  supp_free(&R);
  free(quad);
}

void
orc_nearfield(const orc_bem *b, int far, const unsigned *ridx, unsigned rows,
              const unsigned *cidx, unsigned cols, int ntrans, void *N,
              size_t ld)
{
  if (!b->basis_r && !b->basis_c)
    cc(b, far, ridx, rows, cidx, cols, ntrans, N, ld);
  else if (b->basis_r && b->basis_c)
    ll(b, far, ridx, rows, cidx, cols, ntrans, N, ld);
  else if (b->basis_c)
    cl(b, far, ridx, rows, cidx, cols, ntrans, N, ld);
  else
    lc(b, far, ridx, rows, cidx, cols, ntrans, N, ld);
}

static double
pair_points(const orc_bem *b, int far, unsigned tt, unsigned ss)
{
  if (far)
    return b->sq->dist.n;
  const unsigned *a = b->t + 3 * tt, *c = b->t + 3 * ss;
  unsigned p = 0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      p += (a[i] == c[j]);
  return p == 0 ? b->sq->dist.n
    : (p == 1 ? b->sq->vert.n : (p == 2 ? b->sq->edge.n : b->sq->id.n));
}

double
orc_nearfield_evals(const orc_bem *b, int far, const unsigned *ridx,
                    unsigned rows, const unsigned *cidx, unsigned cols)
{
  double sum = 0.0;
  supp_list R = {0}, C = {0};
  unsigned nr = rows, nc = cols;
  const unsigned *rt = NULL, *ct = NULL;
  if (b->basis_r) {
    supp_build(b, ridx, rows, &R);
    nr = R.n;
    rt = R.tri;
  }
  if (b->basis_c) {
    supp_build(b, cidx, cols, &C);
    nc = C.n;
    ct = C.tri;
  }
  for (unsigned s = 0; s < nc; s++) {
    unsigned ss = ct ? ct[s] : (cidx ? cidx[s] : s);
    for (unsigned t = 0; t < nr; t++) {
      unsigned tt = rt ? rt[t] : (ridx ? ridx[t] : t);
      sum += pair_points(b, far, tt, ss);
    }
  }
  if (b->basis_r)
    supp_free(&R);
  if (b->basis_c)
    supp_free(&C);
  return sum;
}
