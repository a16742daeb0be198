/* H2-matrix approximation by tensor Chebyshev interpolation.
 *
 * Restates the reference's interpolation scheme
 * (setup_h2matrix_aprx_inter_bem3d, Library/bem3d.c:9658-9686):
 *  - m Chebyshev points cos(pi (2i+1) / (2m)) per direction, mapped to the
 *    cluster's bounding box, widened by 5e-3 on each side of a direction
 *    thinner than 5e-3 (Library/bem3d.c:4978-4996, 5179-5283);
 *  - leaf matrices V_t[i, j] = int phi_i L_j (or int phi_i dL_j/dn for the
 *    column basis of the double layer) with the single-triangle Gauss rule
 *    (Library/bem3d.c:7555-7681, 7758-7884, 8024-8336);
 *  - transfer matrices E_t'[i, j] = L_j^t(xi_i^t') for a son t' of t
 *    (Library/bem3d.c:5968-6005, 8497-8531);
 *  - coupling matrices S_ts[i, j] = g(xi_i^t, xi_j^s) / (4 pi) for
 *    admissible blocks (Library/bem3d.c:6079-6108, 3439-3486).
 * Lagrange points are numbered x-direction slowest, z fastest.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "h2bem.h"
#include "gauss01.h"

#define INTERP_EPS 5.0e-3
#define KCONST 0.0795774715459476679

/* Interpolation points of a box: px[0..m), py, pz. */
static void
interp_points(const double *bmin, const double *bmax, const double *cheb,
              unsigned m, double *p)
{
  for (int d = 0; d < 3; d++) {
    double a = bmin[d], b = bmax[d];
    if (b - a < INTERP_EPS) {
      b += INTERP_EPS;
      a -= INTERP_EPS;
    }
    double c = (b + a) * 0.5, h = (b - a) * 0.5;
    for (unsigned i = 0; i < m; i++)
      p[d * m + i] = c + h * cheb[i];
  }
}

/* Lagrange polynomial mu of the points xi[0..m) at x. */
static double
lagr1d(double x, const double *xi, unsigned mu, unsigned m)
{
  double v = 1.0;
  for (unsigned l = 0; l < mu; l++)
    v *= (x - xi[l]) / (xi[mu] - xi[l]);
  for (unsigned l = mu + 1; l < m; l++)
    v *= (x - xi[l]) / (xi[mu] - xi[l]);
  return v;
}

/* Derivative of lagr1d with respect to x (product rule). */
static double
dlagr1d(double x, const double *xi, unsigned mu, unsigned m)
{
  double sum = 0.0;
  for (unsigned l = 0; l < m; l++) {
    if (l == mu)
      continue;
    double v = 1.0;
    for (unsigned l2 = 0; l2 < m; l2++) {
      if (l2 == mu || l2 == l)
        continue;
      v *= (x - xi[l2]) / (xi[mu] - xi[l2]);
    }
    sum += v / (xi[mu] - xi[l]);
  }
  return sum;
}

typedef struct {
  unsigned nq;
  double x0[64], x1[64], w[64];
} trirule;

static void
make_trirule(unsigned q, trirule *r)
{
  double g[8], gw[8];
  h2_gauss01(q, g, gw);
  r->nq = 0;
  for (unsigned i = 0; i < q; i++)
    for (unsigned j = 0; j < q; j++) {
      r->x0[r->nq] = g[i];
      r->x1[r->nq] = g[i] * g[j];
      r->w[r->nq] = gw[i] * gw[j] * g[i];
      r->nq++;
    }
}

/* Values (or normal derivatives) of all k = m^3 Lagrange polynomials at
 * point x; out[j], j = (jx m + jy) m + jz. */
static void
lagr3d(const double *x, const double *p, unsigned m, const double *n,
       double *out)
{
  double lx[16], ly[16], lz[16], dx[16], dy[16], dz[16];
  for (unsigned j = 0; j < m; j++) {
    lx[j] = lagr1d(x[0], p, j, m);
    ly[j] = lagr1d(x[1], p + m, j, m);
    lz[j] = lagr1d(x[2], p + 2 * m, j, m);
    if (n) {
      dx[j] = dlagr1d(x[0], p, j, m);
      dy[j] = dlagr1d(x[1], p + m, j, m);
      dz[j] = dlagr1d(x[2], p + 2 * m, j, m);
    }
  }
  unsigned idx = 0;
  for (unsigned a = 0; a < m; a++)
    for (unsigned b = 0; b < m; b++)
      for (unsigned c = 0; c < m; c++, idx++) {
        if (!n)
          out[idx] = lx[a] * ly[b] * lz[c];
        else
          out[idx] = n[0] * (dx[a] * ly[b] * lz[c])
            + n[1] * (lx[a] * dy[b] * lz[c]) + n[2] * (lx[a] * ly[b] * dz[c]);
      }
}

/* Leaf matrix of node t (|t| x k, column-major). */
static void
leaf_matrix(const h2b_surface *s, const h2b_ctree *T, uint32_t node,
            int basis, int dn, const double *p, unsigned m,
            const trirule *tr, const uint32_t *v2t_ptr, const uint32_t *v2t,
            double *V)
{
  const h2b_cnode *c = &T->nodes[node];
  const uint32_t rows = c->size, k = m * m * m;
  const uint32_t *idx = T->idx + c->off;
  double L[4096], x[3];
  memset(V, 0, sizeof(double) * rows * k);
  if (!basis) {
    for (uint32_t i = 0; i < rows; i++) {
      const uint32_t tri = idx[i];
      const double *A = s->x + 3 * s->t[3 * tri];
      const double *B = s->x + 3 * s->t[3 * tri + 1];
      const double *C = s->x + 3 * s->t[3 * tri + 2];
      for (unsigned q = 0; q < tr->nq; q++) {
        double u = tr->x0[q], v = tr->x1[q];
        for (int d = 0; d < 3; d++)
          x[d] = A[d] * (1.0 - u) + B[d] * (u - v) + C[d] * v;
        lagr3d(x, p, m, dn ? s->n + 3 * tri : NULL, L);
        for (uint32_t j = 0; j < k; j++)
          V[i + (size_t) j * rows] += tr->w[q] * L[j];
      }
      for (uint32_t j = 0; j < k; j++)
        V[i + (size_t) j * rows] *= s->g[tri];
    }
    return;
  }
  /* linear basis: integrate over the support triangles of the vertices,
   * phi_l = 1 - x0, x0 - x1, x1 on the reference triangle */
  for (uint32_t i = 0; i < rows; i++) {
    const uint32_t v = idx[i];
    for (uint32_t pp = v2t_ptr[v]; pp < v2t_ptr[v + 1]; pp++) {
      const uint32_t tri = v2t[pp];
      const uint32_t *tv = s->t + 3 * tri;
      int l = tv[0] == v ? 0 : (tv[1] == v ? 1 : 2);
      const double *A = s->x + 3 * tv[0];
      const double *B = s->x + 3 * tv[1];
      const double *C = s->x + 3 * tv[2];
      for (unsigned q = 0; q < tr->nq; q++) {
        double u = tr->x0[q], w = tr->x1[q];
        for (int d = 0; d < 3; d++)
          x[d] = A[d] * (1.0 - u) + B[d] * (u - w) + C[d] * w;
        double phi = l == 0 ? 1.0 - u : (l == 1 ? u - w : w);
        double wq = tr->w[q] * phi * s->g[tri];
        lagr3d(x, p, m, dn ? s->n + 3 * tri : NULL, L);
        for (uint32_t j = 0; j < k; j++)
          V[i + (size_t) j * rows] += wq * L[j];
      }
    }
  }
}

static void
build_v2t(const h2b_surface *s, uint32_t **ptr_out, uint32_t **v2t_out)
{
  uint32_t *ptr = (uint32_t *) calloc(s->vertices + 1, sizeof(uint32_t));
  for (size_t i = 0; i < 3 * (size_t) s->triangles; i++)
    ptr[s->t[i] + 1]++;
  for (uint32_t v = 0; v < s->vertices; v++)
    ptr[v + 1] += ptr[v];
  uint32_t *v2t = (uint32_t *) malloc(sizeof(uint32_t) * 3 * s->triangles);
  uint32_t *fill = (uint32_t *) calloc(s->vertices, sizeof(uint32_t));
  for (uint32_t i = 0; i < s->triangles; i++)
    for (int l = 0; l < 3; l++) {
      uint32_t v = s->t[3 * i + l];
      v2t[ptr[v] + fill[v]++] = i;
    }
  free(fill);
  *ptr_out = ptr;
  *v2t_out = v2t;
}

static void
side(const h2b_surface *s, const h2b_ctree *T, int basis, int dn,
     unsigned m, const double *cheb, const trirule *tr, double **pts,
     uint64_t **voff, double **V, uint64_t *vlen, double **E)
{
  const uint32_t nn = T->nnodes, k = m * m * m;
  double *P = (double *) malloc(sizeof(double) * 3 * m * nn);
  for (uint32_t i = 0; i < nn; i++)
    interp_points(T->nodes[i].bmin, T->nodes[i].bmax, cheb, m, P + 3 * m * i);
  uint64_t *off = (uint64_t *) malloc(sizeof(uint64_t) * nn);
  uint64_t tot = 0;
  for (uint32_t i = 0; i < nn; i++) {
    if (T->nodes[i].sons == 0) {
      off[i] = tot;
      tot += (uint64_t) T->nodes[i].size * k;
    }
    else
      off[i] = UINT64_MAX;
  }
  double *VV = (double *) malloc(sizeof(double) * (tot ? tot : 1));
  uint32_t *v2t_ptr = NULL, *v2t = NULL;
  if (basis)
    build_v2t(s, &v2t_ptr, &v2t);
#pragma omp parallel for schedule(dynamic, 4)
  for (uint32_t i = 0; i < nn; i++)
    if (T->nodes[i].sons == 0)
      leaf_matrix(s, T, i, basis, dn, P + 3 * m * i, m, tr, v2t_ptr, v2t,
                  VV + off[i]);
  /* transfer matrices, stored per son node: k_son x k_father */
  double *EE = (double *) calloc((size_t) nn * k * k, sizeof(double));
#pragma omp parallel for schedule(dynamic, 4)
  for (uint32_t f = 0; f < nn; f++) {
    const h2b_cnode *c = &T->nodes[f];
    for (uint32_t so = 0; so < c->sons; so++) {
      const uint32_t son = c->son[so];
      const double *ps = P + 3 * m * son, *pf = P + 3 * m * f;
      double *Es = EE + (size_t) son * k * k;
      double X[3];
      uint32_t i = 0;
      for (unsigned a = 0; a < m; a++)
        for (unsigned b = 0; b < m; b++)
          for (unsigned cc = 0; cc < m; cc++, i++) {
            X[0] = ps[a];
            X[1] = ps[m + b];
            X[2] = ps[2 * m + cc];
            uint32_t j = 0;
            for (unsigned ja = 0; ja < m; ja++)
              for (unsigned jb = 0; jb < m; jb++)
                for (unsigned jc = 0; jc < m; jc++, j++) {
                  double v = lagr1d(X[0], pf, ja, m);
                  v *= lagr1d(X[1], pf + m, jb, m);
                  v *= lagr1d(X[2], pf + 2 * m, jc, m);
                  Es[i + (size_t) j * k] = v;
                }
          }
    }
  }
  free(v2t_ptr);
  free(v2t);
  *pts = P;
  *voff = off;
  *V = VV;
  *vlen = tot;
  *E = EE;
}

void
h2b_interp_points(const h2b_ctree *T, unsigned m, double *out)
{
  double cheb[16];
  /* the same nodes as h2b_build_h2interp, bit for bit */
  for (unsigned i = 0; i < m; i++) {
    double e = 1.0 / (2.0 * m);
    cheb[i] = cos(M_PI * (2.0 * i * e + e));
  }
  for (uint32_t i = 0; i < T->nnodes; i++)
    interp_points(T->nodes[i].bmin, T->nodes[i].bmax, cheb, m,
                  out + 3 * (size_t) m * i);
}

h2b_h2interp *
h2b_build_h2interp(const h2b_surface *s, const h2b_ctree *rt,
                   const h2b_ctree *ct, const h2b_blocks *bl, int row_basis,
                   int col_basis, int dn_side, unsigned m, unsigned q)
{
  if (m < 1 || m > 8 || q < 1 || q > 8)
    return NULL;
  h2b_h2interp *H = (h2b_h2interp *) calloc(1, sizeof(h2b_h2interp));
  H->m = m;
  H->k = m * m * m;
  double cheb[16];
  for (unsigned i = 0; i < m; i++) {
    double e = 1.0 / (2.0 * m);
    cheb[i] = cos(M_PI * (2.0 * i * e + e));
  }
  trirule tr;
  make_trirule(q, &tr);
  side(s, rt, row_basis, dn_side == 2, m, cheb, &tr, &H->rpts, &H->rvoff,
       &H->rV, &H->rvlen, &H->rE);
  side(s, ct, col_basis, dn_side == 1, m, cheb, &tr, &H->cpts, &H->cvoff,
       &H->cV, &H->cvlen, &H->cE);
  H->nfar = bl->nfar;
  H->S = NULL;
  return H;
}

void
h2b_h2interp_coupling_laplace(h2b_h2interp *H, const h2b_blocks *bl)
{
  const uint32_t m = H->m, k = H->k;
  free(H->S);
  H->S = (double *) malloc(sizeof(double) * (size_t) k * k *
                           (bl->nfar ? bl->nfar : 1));
#pragma omp parallel for schedule(dynamic, 16)
  for (uint32_t b = 0; b < bl->nfar; b++) {
    const double *pr = H->rpts + 3 * m * bl->far[2 * b];
    const double *pc = H->cpts + 3 * m * bl->far[2 * b + 1];
    double *S = H->S + (size_t) b * k * k;
    double X[512][3], Y[512][3];
    uint32_t i = 0;
    for (unsigned a = 0; a < m; a++)
      for (unsigned bb = 0; bb < m; bb++)
        for (unsigned c = 0; c < m; c++, i++) {
          X[i][0] = pr[a];
          X[i][1] = pr[m + bb];
          X[i][2] = pr[2 * m + c];
          Y[i][0] = pc[a];
          Y[i][1] = pc[m + bb];
          Y[i][2] = pc[2 * m + c];
        }
    for (uint32_t j = 0; j < k; j++)
      for (uint32_t ii = 0; ii < k; ii++) {
        double d0 = X[ii][0] - Y[j][0], d1 = X[ii][1] - Y[j][1],
          d2 = X[ii][2] - Y[j][2];
        S[ii + (size_t) j * k] = KCONST * (1.0 / sqrt(d0 * d0 + d1 * d1
                                                      + d2 * d2));
      }
  }
}

/* Couplings of the Helmholtz kernel, complex (re, im interleaved) into
 * S_out (2 k^2 per admissible block): kernel_const e^{i kr r} e^{-ki r} / r
 * at the interpolation points of both boxes, k = kr + i ki
 * (fill_kernel_helmholtzbem3d with slp_kernel_helmholtzbem3d,
 * Library/helmholtzbem3d.c:26-55, 828-832, through
 * assemble_bem3d_inter_uniform, Library/bem3d.c:6079-6108). */
void
h2b_h2interp_coupling_helmholtz(const h2b_h2interp *H, const h2b_blocks *bl,
                                double kr, double ki, double *S_out)
{
  const uint32_t m = H->m, k = H->k;
  const double kdamp = -ki;
#pragma omp parallel for schedule(dynamic, 16)
  for (uint32_t b = 0; b < bl->nfar; b++) {
    const double *pr = H->rpts + 3 * m * bl->far[2 * b];
    const double *pc = H->cpts + 3 * m * bl->far[2 * b + 1];
    double *S = S_out + (size_t) 2 * b * k * k;
    double X[512][3], Y[512][3];
    uint32_t i = 0;
    for (unsigned a = 0; a < m; a++)
      for (unsigned bb = 0; bb < m; bb++)
        for (unsigned c = 0; c < m; c++, i++) {
          X[i][0] = pr[a];
          X[i][1] = pr[m + bb];
          X[i][2] = pr[2 * m + c];
          Y[i][0] = pc[a];
          Y[i][1] = pc[m + bb];
          Y[i][2] = pc[2 * m + c];
        }
    for (uint32_t j = 0; j < k; j++)
      for (uint32_t ii = 0; ii < k; ii++) {
        const double d0 = X[ii][0] - Y[j][0], d1 = X[ii][1] - Y[j][1],
          d2 = X[ii][2] - Y[j][2];
        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        double rs = 1.0 / sqrt(r2);
        const double r = r2 * rs;
        if (kdamp != 0.0)
          rs *= exp(kdamp * r);
        const double ph = kr * r;
        const size_t o = 2 * (ii + (size_t) j * k);
        S[o] = KCONST * (rs * cos(ph));
        S[o + 1] = KCONST * (sin(ph) * rs);
      }
  }
}

void
h2b_h2interp_free(h2b_h2interp *H)
{
  if (!H)
    return;
  free(H->rpts);
  free(H->cpts);
  free(H->rvoff);
  free(H->cvoff);
  free(H->rV);
  free(H->cV);
  free(H->rE);
  free(H->cE);
  free(H->S);
  free(H);
}
