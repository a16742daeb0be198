/* Cluster trees and block trees.
 *
 * Cluster trees follow build_bem3d_cluster (Library/bem3d.c:810-954) and
 * build_adaptive_cluster (Library/cluster.c:80-168): a cluster with more
 * than clf DOFs is split at the midpoint of the longest side of the
 * bounding box of its characteristic points; the partition is done in place
 * with the same swap sequence, so the DOF permutation equals the
 * reference's.  Leaves take the bounding box of the supports of their DOFs,
 * inner clusters the union of their sons' boxes.
 *
 * Block trees follow build_strict_block (Library/block.c:258-322) with
 * admissible_max_cluster / admissible_2_cluster (Library/block.c:29-63).
 */
#include <stdlib.h>
#include <string.h>

#include "h2bem.h"

typedef struct {
  const double *px;           /* characteristic points, 3 per DOF */
  const double *smin, *smax;  /* support boxes, 3 per DOF */
  uint32_t *idx;
  uint32_t clf;
  h2b_cnode *nodes;
  uint32_t nnodes, cap;
} builder;

static uint32_t
new_node(builder *B)
{
  if (B->nnodes == B->cap) {
    B->cap = B->cap ? 2 * B->cap : 64;
    B->nodes = (h2b_cnode *) realloc(B->nodes, sizeof(h2b_cnode) * B->cap);
  }
  memset(&B->nodes[B->nnodes], 0, sizeof(h2b_cnode));
  return B->nnodes++;
}

static void
support_box(builder *B, h2b_cnode *c)
{
  const uint32_t *idx = B->idx + c->off;
  for (int d = 0; d < 3; d++) {
    c->bmin[d] = B->smin[3 * idx[0] + d];
    c->bmax[d] = B->smax[3 * idx[0] + d];
  }
  for (uint32_t i = 1; i < c->size; i++)
    for (int d = 0; d < 3; d++) {
      double lo = B->smin[3 * idx[i] + d], hi = B->smax[3 * idx[i] + d];
      if (lo < c->bmin[d])
        c->bmin[d] = lo;
      if (hi > c->bmax[d])
        c->bmax[d] = hi;
    }
}

static uint32_t
build(builder *B, uint32_t off, uint32_t size, uint32_t level)
{
  uint32_t me = new_node(B);
  B->nodes[me].off = off;
  B->nodes[me].size = size;
  B->nodes[me].level = level;
  if (size > B->clf) {
    uint32_t *idx = B->idx + off;
    double lo[3], hi[3];
    for (int d = 0; d < 3; d++)
      lo[d] = hi[d] = B->px[3 * idx[0] + d];
    for (uint32_t i = 1; i < size; i++)
      for (int d = 0; d < 3; d++) {
        double v = B->px[3 * idx[i] + d];
        if (v < lo[d])
          lo[d] = v;
        if (v > hi[d])
          hi[d] = v;
      }
    int dir = 0;
    double ext = hi[0] - lo[0];
    for (int d = 1; d < 3; d++)
      if (ext < hi[d] - lo[d]) {
        ext = hi[d] - lo[d];
        dir = d;
      }
    if (ext > 0.0) {
      double mid = (hi[dir] + lo[dir]) / 2.0;
      uint32_t n0 = 0;
      for (uint32_t i = 0; i < size; i++)
        if (B->px[3 * idx[i] + dir] < mid) {
          uint32_t tmp = idx[i];
          idx[i] = idx[n0];
          idx[n0] = tmp;
          n0++;
        }
      if (n0 > 0 && n0 < size) {
        uint32_t s0 = build(B, off, n0, level + 1);
        uint32_t s1 = build(B, off + n0, size - n0, level + 1);
        h2b_cnode *c = &B->nodes[me];
        c->sons = 2;
        c->son[0] = s0;
        c->son[1] = s1;
        for (int d = 0; d < 3; d++) {
          double a0 = B->nodes[s0].bmin[d], a1 = B->nodes[s1].bmin[d];
          double b0 = B->nodes[s0].bmax[d], b1 = B->nodes[s1].bmax[d];
          c->bmin[d] = a0 < a1 ? a0 : a1;
          c->bmax[d] = b0 > b1 ? b0 : b1;
        }
        c->desc = 1 + B->nodes[s0].desc + B->nodes[s1].desc;
        return me;
      }
    }
  }
  support_box(B, &B->nodes[me]);
  B->nodes[me].desc = 1;
  return me;
}

h2b_ctree *
h2b_build_cluster(const h2b_surface *s, int basis, uint32_t clf)
{
  const uint32_t n = basis ? s->vertices : s->triangles;
  double *px = (double *) malloc(sizeof(double) * 3 * n);
  double *smin = (double *) malloc(sizeof(double) * 3 * n);
  double *smax = (double *) malloc(sizeof(double) * 3 * n);
  if (!basis) {
    /* centroids and triangle boxes (Library/bem3d.c:810-849) */
    for (uint32_t i = 0; i < n; i++) {
      const double *a = s->x + 3 * s->t[3 * i];
      const double *b = s->x + 3 * s->t[3 * i + 1];
      const double *c = s->x + 3 * s->t[3 * i + 2];
      for (int d = 0; d < 3; d++) {
        px[3 * i + d] = (a[d] + b[d] + c[d]) / 3.0;
        double lo = a[d] < b[d] ? a[d] : b[d], hi = a[d] > b[d] ? a[d] : b[d];
        smin[3 * i + d] = lo < c[d] ? lo : c[d];
        smax[3 * i + d] = hi > c[d] ? hi : c[d];
      }
    }
  }
  else {
    /* vertices; supports are the unions of the boxes of the adjacent
     * triangles (Library/bem3d.c:851-912) */
    memcpy(px, s->x, sizeof(double) * 3 * n);
    memcpy(smin, s->x, sizeof(double) * 3 * n);
    memcpy(smax, s->x, sizeof(double) * 3 * n);
    for (uint32_t i = 0; i < s->triangles; i++) {
      const uint32_t *tv = s->t + 3 * i;
      double lo[3], hi[3];
      for (int d = 0; d < 3; d++) {
        double a = s->x[3 * tv[0] + d], b = s->x[3 * tv[1] + d],
          c = s->x[3 * tv[2] + d];
        lo[d] = a < b ? a : b;
        lo[d] = lo[d] < c ? lo[d] : c;
        hi[d] = a > b ? a : b;
        hi[d] = hi[d] > c ? hi[d] : c;
      }
      for (int j = 0; j < 3; j++)
        for (int d = 0; d < 3; d++) {
          double *mn = smin + 3 * tv[j] + d, *mx = smax + 3 * tv[j] + d;
          if (lo[d] < *mn)
            *mn = lo[d];
          if (hi[d] > *mx)
            *mx = hi[d];
        }
    }
  }
  builder B = {px, smin, smax, NULL, clf, NULL, 0, 0};
  h2b_ctree *T = (h2b_ctree *) calloc(1, sizeof(h2b_ctree));
  T->n = n;
  T->idx = (uint32_t *) malloc(sizeof(uint32_t) * n);
  for (uint32_t i = 0; i < n; i++)
    T->idx[i] = i;
  B.idx = T->idx;
  build(&B, 0, n, 0);
  T->nodes = B.nodes;
  T->nnodes = B.nnodes;
  free(px);
  free(smin);
  free(smax);
  return T;
}

void
h2b_ctree_free(h2b_ctree *t)
{
  if (!t)
    return;
  free(t->idx);
  free(t->nodes);
  free(t);
}

/* ---- block trees ------------------------------------------------------------ */

static int
admissible(const h2b_cnode *r, const h2b_cnode *c, double eta, int admis)
{
  if (admis == 0) {
    double dr = 0.0, dc = 0.0, dist = 0.0;
    for (int d = 0; d < 3; d++) {
      double a = r->bmax[d] - r->bmin[d], b = c->bmax[d] - c->bmin[d];
      if (a > dr)
        dr = a;
      if (b > dc)
        dc = b;
      double gap = 0.0;
      if (r->bmax[d] < c->bmin[d])
        gap = c->bmin[d] - r->bmax[d];
      else if (c->bmax[d] < r->bmin[d])
        gap = r->bmin[d] - c->bmax[d];
      if (gap > dist)
        dist = gap;
    }
    return dr <= eta * dist && dc <= eta * dist;
  }
  double dr = 0.0, dc = 0.0, dist = 0.0;
  for (int d = 0; d < 3; d++) {
    double a = r->bmax[d] - r->bmin[d];
    dr += a * a;
    a = c->bmax[d] - c->bmin[d];
    dc += a * a;
    a = r->bmin[d] - c->bmax[d];
    double b = c->bmin[d] - r->bmax[d];
    double g = 0.0;
    if (a > g)
      g = a;
    if (b > g)
      g = b;
    dist += g * g;
  }
  return dr <= eta * eta * dist && dc <= eta * eta * dist;
}

typedef struct {
  uint32_t *near, *far;
  uint32_t nnear, nfar, capn, capf;
} leaves;

static void
push(uint32_t **a, uint32_t *n, uint32_t *cap, uint32_t r, uint32_t c)
{
  if (*n == *cap) {
    *cap = *cap ? 2 * *cap : 256;
    *a = (uint32_t *) realloc(*a, sizeof(uint32_t) * 2 * *cap);
  }
  (*a)[2 * *n] = r;
  (*a)[2 * *n + 1] = c;
  (*n)++;
}

static void
strict(const h2b_ctree *rt, const h2b_ctree *ct, uint32_t r, uint32_t c,
       double eta, int admis, leaves *L)
{
  const h2b_cnode *rn = &rt->nodes[r], *cn = &ct->nodes[c];
  if (admissible(rn, cn, eta, admis)) {
    push(&L->far, &L->nfar, &L->capf, r, c);
    return;
  }
  if (rn->sons == 0 && cn->sons == 0) {
    push(&L->near, &L->nnear, &L->capn, r, c);
    return;
  }
  /* sons in storage order i + j * rsons, row index fastest */
  uint32_t rs = rn->sons ? rn->sons : 1, cs = cn->sons ? cn->sons : 1;
  for (uint32_t j = 0; j < cs; j++)
    for (uint32_t i = 0; i < rs; i++)
      strict(rt, ct, rn->sons ? rn->son[i] : r, cn->sons ? cn->son[j] : c,
             eta, admis, L);
}

h2b_blocks *
h2b_build_strict_block(const h2b_ctree *rt, const h2b_ctree *ct, double eta,
                       int admis)
{
  leaves L = {NULL, NULL, 0, 0, 0, 0};
  strict(rt, ct, 0, 0, eta, admis, &L);
  h2b_blocks *b = (h2b_blocks *) calloc(1, sizeof(h2b_blocks));
  b->nnear = L.nnear;
  b->nfar = L.nfar;
  b->near = L.near;
  b->far = L.far;
  return b;
}

void
h2b_blocks_free(h2b_blocks *b)
{
  if (!b)
    return;
  free(b->near);
  free(b->far);
  free(b);
}

void
h2b_near_block_list(const h2b_blocks *b, const h2b_ctree *rt,
                    const h2b_ctree *ct, uint32_t *out)
{
  for (uint32_t k = 0; k < b->nnear; k++) {
    const h2b_cnode *r = &rt->nodes[b->near[2 * k]];
    const h2b_cnode *c = &ct->nodes[b->near[2 * k + 1]];
    out[4 * k] = r->off;
    out[4 * k + 1] = r->size;
    out[4 * k + 2] = c->off;
    out[4 * k + 3] = c->size;
  }
}
