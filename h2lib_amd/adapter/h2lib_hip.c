/* H2Lib-side binding of libh2hip (see h2lib_hip.h).
 *
 * Built inside an H2Lib tree: gcc -I<H2Lib>/Library -I<this repo>/include
 * -c h2lib_hip.c, linked with -lh2hip.  Compile with -DUSE_COMPLEX for the
 * complex H2Lib build, exactly like the Library sources.
 */
#include "h2lib_hip.h"

#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "h2hip.h"
#include "laplacebem3d.h"
#ifdef USE_COMPLEX
#include "helmholtzbem3d.h"
#endif

#ifdef USE_COMPLEX
#define FIELD_DOUBLES 2
#else
#define FIELD_DOUBLES 1
#endif

static void
hip_fail(const char *what)
{
  fprintf(stderr, "h2lib_hip: %s: %s\n", what, h2hip_last_error());
  abort();
}

static void
to_pair(field z, double *p)
{
#ifdef USE_COMPLEX
  p[0] = creal(z);
  p[1] = cimag(z);
#else
  p[0] = z;
  p[1] = 0.0;
#endif
}

/* ------------------------------------------------------------------
 * bem3d -> device operator registry
 * ------------------------------------------------------------------ */

typedef struct hipbem {
  pcbem3d   bem;
  h2hip_surface *s;
  h2hip_op *op;
  struct hipbem *next;
} hipbem;

static hipbem *registry = NULL;
static pthread_mutex_t registry_lock = PTHREAD_MUTEX_INITIALIZER;

static h2hip_op *
lookup_op(pcbem3d bem)
{
  hipbem   *e;
  h2hip_op *op = NULL;

  pthread_mutex_lock(&registry_lock);
  for (e = registry; e; e = e->next)
    if (e->bem == bem) {
      op = e->op;
      break;
    }
  pthread_mutex_unlock(&registry_lock);
  if (op == NULL) {
    fprintf(stderr, "h2lib_hip: bem3d object without device operator\n");
    abort();
  }
  return op;
}

/* The callback signature of Library/bem3d.h:380-381: rows and cols of the
 * window are read from N as assemble_cc_near_bem3d does
 * (Library/bem3d.c:1160-1161). */
static void
hip_fill(const uint * ridx, const uint * cidx, pcbem3d bem, bool ntrans,
         pamatrix N, int far)
{
  uint      rows = ntrans ? N->cols : N->rows;
  uint      cols = ntrans ? N->rows : N->cols;

  if (h2hip_nearfield(lookup_op(bem), far, ridx, rows, cidx, cols,
                      ntrans ? 1 : 0, N->a, N->ld) != H2HIP_OK)
    hip_fail("nearfield");
}

static void
hip_nearfield(const uint * ridx, const uint * cidx, pcbem3d bem,
              bool ntrans, pamatrix N)
{
  hip_fill(ridx, cidx, bem, ntrans, N, 0);
}

static void
hip_nearfield_far(const uint * ridx, const uint * cidx, pcbem3d bem,
                  bool ntrans, pamatrix N)
{
  hip_fill(ridx, cidx, bem, ntrans, N, 1);
}

/* Deferred callbacks: the block is recorded and N written by the flush, as
 * the reference's OpenCL callbacks enqueue their blocks
 * (Library/laplaceoclbem3d.c:248-279). */
static void
hip_defer(const uint * ridx, const uint * cidx, pcbem3d bem, bool ntrans,
          pamatrix N, int far)
{
  uint      rows = ntrans ? N->cols : N->rows;
  uint      cols = ntrans ? N->rows : N->cols;

  if (h2hip_nearfield_defer(lookup_op(bem), far, ridx, rows, cidx, cols,
                            ntrans ? 1 : 0, N->a, N->ld) != H2HIP_OK)
    hip_fail("nearfield (deferred)");
}

static void
hip_nearfield_deferred(const uint * ridx, const uint * cidx, pcbem3d bem,
                       bool ntrans, pamatrix N)
{
  hip_defer(ridx, cidx, bem, ntrans, N, 0);
}

static void
hip_nearfield_far_deferred(const uint * ridx, const uint * cidx,
                           pcbem3d bem, bool ntrans, pamatrix N)
{
  hip_defer(ridx, cidx, bem, ntrans, N, 1);
}

static void setup_hip(pbem3d bem, int helmholtz, int layer);

void
setup_hip_bem3d(pbem3d bem, int helmholtz, int layer)
{
  setup_hip(bem, helmholtz, layer);
  bem->nearfield = hip_nearfield;
  bem->nearfield_far = hip_nearfield_far;
}

void
setup_hip_deferred_bem3d(pbem3d bem, int helmholtz, int layer)
{
  setup_hip(bem, helmholtz, layer);
  bem->nearfield = hip_nearfield_deferred;
  bem->nearfield_far = hip_nearfield_far_deferred;
}

void
flush_hip_bem3d(pcbem3d bem)
{
  if (h2hip_nearfield_flush(lookup_op(bem)) != H2HIP_OK)
    hip_fail("nearfield flush");
}

static void
setup_hip(pbem3d bem, int helmholtz, int layer)
{
  pcsurface3d gr = bem->gr;
  hipbem   *e;
  double    alpha[2], k[2];
  int       rb = bem->row_basis == BASIS_LINEAR_BEM3D ? H2HIP_BASIS_LINEAR
    : H2HIP_BASIS_CONSTANT;
  int       cb = bem->col_basis == BASIS_LINEAR_BEM3D ? H2HIP_BASIS_LINEAR
    : H2HIP_BASIS_CONSTANT;

  e = (hipbem *) calloc(1, sizeof(hipbem));
  if (e == NULL)
    abort();
  e->bem = bem;
  if (h2hip_surface_create(&e->s, gr->vertices, (const double *) gr->x,
                           gr->triangles, (const uint32_t *) gr->t,
                           (const double *) gr->n, gr->g) != H2HIP_OK)
    hip_fail("surface");
  to_pair(bem->alpha, alpha);
  to_pair(bem->k, k);
  if (h2hip_op_create(&e->op, e->s,
                      helmholtz ? H2HIP_HELMHOLTZ : H2HIP_LAPLACE,
                      layer, rb, cb, bem->sq->q,
                      bem->sq->q2, alpha, k) != H2HIP_OK)
    hip_fail("operator");
  pthread_mutex_lock(&registry_lock);
  e->next = registry;
  registry = e;
  pthread_mutex_unlock(&registry_lock);
}

void
release_hip_bem3d(pcbem3d bem)
{
  hipbem  **p, *e = NULL;

  pthread_mutex_lock(&registry_lock);
  for (p = &registry; *p; p = &(*p)->next)
    if ((*p)->bem == bem) {
      e = *p;
      *p = e->next;
      break;
    }
  pthread_mutex_unlock(&registry_lock);
  if (e) {
    h2hip_op_destroy(e->op);
    h2hip_surface_destroy(e->s);
    free(e);
  }
}

void
set_precision_hip_bem3d(pcbem3d bem, int mode)
{
  if (h2hip_op_set_precision(lookup_op(bem), mode) != H2HIP_OK)
    hip_fail("h2hip_op_set_precision");
}

pbem3d
new_slp_laplace_hip_bem3d(pcsurface3d gr, uint q_regular, uint q_singular,
                          basisfunctionbem3d row_basis,
                          basisfunctionbem3d col_basis)
{
  pbem3d    bem = new_slp_laplace_bem3d(gr, q_regular, q_singular,
                                        row_basis, col_basis);
  setup_hip_bem3d(bem, 0, 0);
  return bem;
}

pbem3d
new_dlp_laplace_hip_bem3d(pcsurface3d gr, uint q_regular, uint q_singular,
                          basisfunctionbem3d row_basis,
                          basisfunctionbem3d col_basis, field alpha)
{
  pbem3d    bem = new_dlp_laplace_bem3d(gr, q_regular, q_singular,
                                        row_basis, col_basis, alpha);
  setup_hip_bem3d(bem, 0, 1);
  return bem;
}

pbem3d
new_adlp_laplace_hip_bem3d(pcsurface3d gr, uint q_regular, uint q_singular,
                           basisfunctionbem3d row_basis,
                           basisfunctionbem3d col_basis, field alpha)
{
  pbem3d    bem = new_adlp_laplace_bem3d(gr, q_regular, q_singular,
                                         row_basis, col_basis, alpha);
  setup_hip_bem3d(bem, 0, H2HIP_ADLP);
  return bem;
}

#ifdef USE_COMPLEX
pbem3d
new_slp_helmholtz_hip_bem3d(field k, pcsurface3d gr, uint q_regular,
                            uint q_singular, basisfunctionbem3d row_basis,
                            basisfunctionbem3d col_basis)
{
  pbem3d    bem = new_slp_helmholtz_bem3d(k, gr, q_regular, q_singular,
                                          row_basis, col_basis);
  setup_hip_bem3d(bem, 1, 0);
  return bem;
}

pbem3d
new_dlp_helmholtz_hip_bem3d(field k, pcsurface3d gr, uint q_regular,
                            uint q_singular, basisfunctionbem3d row_basis,
                            basisfunctionbem3d col_basis, field alpha)
{
  pbem3d    bem = new_dlp_helmholtz_bem3d(k, gr, q_regular, q_singular,
                                          row_basis, col_basis, alpha);
  setup_hip_bem3d(bem, 1, 1);
  return bem;
}

pbem3d
new_adlp_helmholtz_hip_bem3d(field k, pcsurface3d gr, uint q_regular,
                             uint q_singular, basisfunctionbem3d row_basis,
                             basisfunctionbem3d col_basis, field alpha)
{
  pbem3d    bem = new_adlp_helmholtz_bem3d(k, gr, q_regular, q_singular,
                                           row_basis, col_basis, alpha);
  setup_hip_bem3d(bem, 1, H2HIP_ADLP);
  return bem;
}
#endif

/* ------------------------------------------------------------------
 * h2matrix -> packed device H2-matrix
 * ------------------------------------------------------------------ */

struct _hiph2matrix {
  h2hip_h2matrix *H;
  uint      nrows, ncols;
  void     *d_x, *d_y;
};

typedef struct {
  pcclusterbasis cb;
  uint32_t  id;
} cbkey;

typedef struct {
  pcclusterbasis *cb;           /* node -> cluster basis (preorder) */
  uint32_t *tree, *rank;
  uint32_t  n, cap;
  const uint *idx0;
  uint      maxk;
  cbkey    *keys;               /* sorted by address, for side_find */
} side_t;

static int
cbkey_cmp(const void *a, const void *b)
{
  uintptr_t x = (uintptr_t) ((const cbkey *) a)->cb;
  uintptr_t y = (uintptr_t) ((const cbkey *) b)->cb;
  return x < y ? -1 : x > y;
}

static void
side_index(side_t * s)
{
  uint32_t  i;

  s->keys = (cbkey *) malloc(sizeof(cbkey) * (s->n ? s->n : 1));
  if (s->keys == NULL)
    abort();
  for (i = 0; i < s->n; i++) {
    s->keys[i].cb = s->cb[i];
    s->keys[i].id = i;
  }
  qsort(s->keys, s->n, sizeof(cbkey), cbkey_cmp);
}

static void
side_add(side_t * s, pcclusterbasis cb, uint32_t father)
{
  uint      i, id;

  if (s->n == s->cap) {
    s->cap = s->cap ? 2 * s->cap : 256;
    s->cb = (pcclusterbasis *) realloc(s->cb, sizeof(*s->cb) * s->cap);
    s->tree = (uint32_t *) realloc(s->tree, sizeof(uint32_t) * 3 * s->cap);
    s->rank = (uint32_t *) realloc(s->rank, sizeof(uint32_t) * s->cap);
    if (!s->cb || !s->tree || !s->rank)
      abort();
  }
  id = s->n++;
  s->cb[id] = cb;
  s->tree[3 * id] = (uint32_t) (cb->t->idx - s->idx0);
  s->tree[3 * id + 1] = cb->t->size;
  s->tree[3 * id + 2] = father;
  s->rank[id] = cb->k;
  if (cb->k > s->maxk)
    s->maxk = cb->k;
  for (i = 0; i < cb->sons; i++)
    side_add(s, cb->son[i], id);
}

static uint32_t
side_find(const side_t * s, pcclusterbasis cb)
{
  cbkey     key, *hit;

  key.cb = cb;
  key.id = 0;
  hit = (cbkey *) bsearch(&key, s->keys, s->n, sizeof(cbkey), cbkey_cmp);
  if (hit == NULL) {
    fprintf(stderr, "h2lib_hip: block with unknown cluster basis\n");
    abort();
  }
  return hit->id;
}

static double
real_part(field z)
{
#ifdef USE_COMPLEX
  if (cimag(z) != 0.0) {
    fprintf(stderr, "h2lib_hip: cluster bases must be real\n");
    abort();
  }
  return creal(z);
#else
  return z;
#endif
}

/* leaf matrices V (size x k) and transfer matrices E (k x k_father) of one
 * side, in the layout of h2hip_h2matrix_set_basis */
static void
side_basis(const side_t * s, double **V, uint64_t ** voff, double **E,
           uint64_t * eld)
{
  uint64_t  len = 0, ld = (uint64_t) s->maxk * s->maxk;
  uint32_t  i, r, c;

  *voff = (uint64_t *) calloc(s->n, sizeof(uint64_t));
  for (i = 0; i < s->n; i++)
    if (s->cb[i]->sons == 0) {
      (*voff)[i] = len;
      len += (uint64_t) s->cb[i]->t->size * s->cb[i]->k;
    }
  *V = (double *) calloc(len + 1, sizeof(double));
  *E = (double *) calloc(ld * s->n + 1, sizeof(double));
  if (!*voff || !*V || !*E)
    abort();
  for (i = 0; i < s->n; i++) {
    pcclusterbasis cb = s->cb[i];
    if (cb->sons == 0)
      for (c = 0; c < cb->k; c++)
        for (r = 0; r < cb->t->size; r++)
          (*V)[(*voff)[i] + r + (uint64_t) c * cb->t->size] =
            real_part(cb->V.a[r + (uint64_t) c * cb->V.ld]);
    if (s->tree[3 * i + 2] != 0xffffffffu) {
      uint32_t  kf = s->rank[s->tree[3 * i + 2]];
      for (c = 0; c < kf; c++)
        for (r = 0; r < cb->k; r++)
          (*E)[(uint64_t) i * ld + r + (uint64_t) c * cb->k] =
            real_part(cb->E.a[r + (uint64_t) c * cb->E.ld]);
    }
  }
  *eld = ld;
}

typedef struct {
  uint32_t *near, *far;
  pcamatrix *f;
  pcamatrix *S;
  uint32_t  nnear, nfar, capn, capf;
} blocks_t;

static void
collect_blocks(pch2matrix G, const side_t * r, const side_t * c,
               blocks_t * b)
{
  uint      i;

  if (G->u) {
    if (b->nfar == b->capf) {
      b->capf = b->capf ? 2 * b->capf : 256;
      b->far = (uint32_t *) realloc(b->far, sizeof(uint32_t) * 2 * b->capf);
      b->S = (pcamatrix *) realloc(b->S, sizeof(pcamatrix) * b->capf);
      if (!b->far || !b->S)
        abort();
    }
    b->far[2 * b->nfar] = side_find(r, G->rb);
    b->far[2 * b->nfar + 1] = side_find(c, G->cb);
    b->S[b->nfar++] = &G->u->S;
  }
  else if (G->f) {
    if (b->nnear == b->capn) {
      b->capn = b->capn ? 2 * b->capn : 256;
      b->near = (uint32_t *) realloc(b->near,
                                     sizeof(uint32_t) * 2 * b->capn);
      b->f = (pcamatrix *) realloc(b->f, sizeof(pcamatrix) * b->capn);
      if (!b->near || !b->f)
        abort();
    }
    b->near[2 * b->nnear] = side_find(r, G->rb);
    b->near[2 * b->nnear + 1] = side_find(c, G->cb);
    b->f[b->nnear++] = G->f;
  }
  else if (G->son) {
    for (i = 0; i < G->rsons * G->csons; i++)
      collect_blocks(G->son[i], r, c, b);
  }
}

phiph2matrix
new_hip_h2matrix(pch2matrix h2)
{
  side_t    rs, cs;
  blocks_t  bl;
  phiph2matrix A;
  double   *V, *E;
  uint64_t *voff, eld, sld, len, off;
  field    *S, *N;
  uint32_t  b, i, j;

  memset(&rs, 0, sizeof rs);
  memset(&cs, 0, sizeof cs);
  memset(&bl, 0, sizeof bl);
  rs.idx0 = h2->rb->t->idx;
  cs.idx0 = h2->cb->t->idx;
  side_add(&rs, h2->rb, 0xffffffffu);
  side_add(&cs, h2->cb, 0xffffffffu);
  side_index(&rs);
  side_index(&cs);
  collect_blocks(h2, &rs, &cs, &bl);

  A = (phiph2matrix) calloc(1, sizeof(*A));
  if (A == NULL)
    abort();
  A->nrows = h2->rb->t->size;
  A->ncols = h2->cb->t->size;
  if (h2hip_h2matrix_create(&A->H, FIELD_DOUBLES == 2, rs.n, rs.tree,
                            rs.rank, A->nrows, h2->rb->t->idx, cs.n, cs.tree,
                            cs.rank, A->ncols, h2->cb->t->idx, bl.nnear,
                            bl.near, bl.nfar, bl.far) != H2HIP_OK)
    hip_fail("h2matrix");
  side_basis(&rs, &V, &voff, &E, &eld);
  if (h2hip_h2matrix_set_basis(A->H, 0, V, voff, E, eld) != H2HIP_OK)
    hip_fail("row basis");
  free(V);
  free(voff);
  free(E);
  side_basis(&cs, &V, &voff, &E, &eld);
  if (h2hip_h2matrix_set_basis(A->H, 1, V, voff, E, eld) != H2HIP_OK)
    hip_fail("column basis");
  free(V);
  free(voff);
  free(E);

  /* couplings, block b at b * sld */
  sld = (uint64_t) rs.maxk * cs.maxk;
  S = (field *) calloc(sld * bl.nfar + 1, sizeof(field));
  if (S == NULL)
    abort();
  for (b = 0; b < bl.nfar; b++) {
    pcamatrix M = bl.S[b];
    for (j = 0; j < M->cols; j++)
      for (i = 0; i < M->rows; i++)
        S[b * sld + i + (uint64_t) j * M->rows] = M->a[i + (uint64_t) j * M->ld];
  }
  if (h2hip_h2matrix_set_couplings(A->H, S, sld) != H2HIP_OK)
    hip_fail("couplings");
  free(S);

  /* near field, packed in block order */
  len = 0;
  for (b = 0; b < bl.nnear; b++)
    len += (uint64_t) bl.f[b]->rows * bl.f[b]->cols;
  N = (field *) calloc(len + 1, sizeof(field));
  if (N == NULL)
    abort();
  off = 0;
  for (b = 0; b < bl.nnear; b++) {
    pcamatrix M = bl.f[b];
    for (j = 0; j < M->cols; j++)
      for (i = 0; i < M->rows; i++)
        N[off + i + (uint64_t) j * M->rows] = M->a[i + (uint64_t) j * M->ld];
    off += (uint64_t) M->rows * M->cols;
  }
  if (h2hip_h2matrix_set_nearfield(A->H, N) != H2HIP_OK)
    hip_fail("near field");
  free(N);
  /* flat basis transforms where the tree allows them (leaves that
   * partition the index positions, as H2Lib's cluster trees do), else the
   * level-synchronous sweep */
  if (h2hip_h2matrix_set_schedule(A->H, 2) != H2HIP_OK
      && h2hip_h2matrix_set_schedule(A->H, 1) != H2HIP_OK)
    hip_fail("schedule");

  if (h2hip_malloc(&A->d_x, sizeof(field) * (A->ncols > A->nrows ?
                                             A->ncols : A->nrows) + 16)
      != H2HIP_OK
      || h2hip_malloc(&A->d_y, sizeof(field) * (A->ncols > A->nrows ?
                                                A->ncols : A->nrows) + 16)
      != H2HIP_OK)
    hip_fail("vectors");

  free(rs.cb);
  free(rs.tree);
  free(rs.rank);
  free(rs.keys);
  free(cs.keys);
  free(cs.cb);
  free(cs.tree);
  free(cs.rank);
  free(bl.near);
  free(bl.far);
  free(bl.f);
  free(bl.S);
  return A;
}

void
del_hip_h2matrix(phiph2matrix A)
{
  if (A == NULL)
    return;
  h2hip_free(A->d_x);
  h2hip_free(A->d_y);
  h2hip_h2matrix_destroy(A->H);
  free(A);
}

static void
hip_addeval(field alpha, phiph2matrix A, pcavector x, pavector y, int trans)
{
  double    a[2];
  uint      nin = trans ? A->nrows : A->ncols;
  uint      nout = trans ? A->ncols : A->nrows;
  int       rc;

  to_pair(alpha, a);
  if (x->dim < nin || y->dim < nout) {
    fprintf(stderr, "h2lib_hip: vector dimensions do not match\n");
    abort();
  }
  if (h2hip_memcpy(A->d_x, x->v, sizeof(field) * nin, 0) != H2HIP_OK ||
      h2hip_memcpy(A->d_y, y->v, sizeof(field) * nout, 0) != H2HIP_OK)
    hip_fail("upload");
  rc = trans ? h2hip_addevaltrans_h2matrix(a, A->H, A->d_x, A->d_y, NULL)
    : h2hip_addeval_h2matrix(a, A->H, A->d_x, A->d_y, NULL);
  if (rc != H2HIP_OK)
    hip_fail("addeval");
  if (h2hip_memcpy(y->v, A->d_y, sizeof(field) * nout, 1) != H2HIP_OK)
    hip_fail("download");
}

void
addeval_hip_h2matrix_avector(field alpha, phiph2matrix A, pcavector x,
                             pavector y)
{
  hip_addeval(alpha, A, x, y, 0);
}

void
addevaltrans_hip_h2matrix_avector(field alpha, phiph2matrix A, pcavector x,
                                  pavector y)
{
  hip_addeval(alpha, A, x, y, 1);
}

void
addeval_hip_h2matrix_callback(field alpha, void *A, pcavector x, pavector y)
{
  hip_addeval(alpha, (phiph2matrix) A, x, y, 0);
}

void
mvm_hip_h2matrix_callback(field alpha, bool trans, void *A, pcavector x,
                          pavector y)
{
  hip_addeval(alpha, (phiph2matrix) A, x, y, trans ? 1 : 0);
}

static int
addeval_device(const double *alpha, void *H, const void *x, void *y,
               void *stream)
{
  return h2hip_addeval_h2matrix(alpha, (const h2hip_h2matrix *) H, x, y,
                                stream);
}

uint
solve_gmres_hip_h2matrix_avector(phiph2matrix A, pcavector b, pavector x,
                                 real eps, uint maxiter, uint kmax)
{
  uint32_t  it = 0;
  double    res = 0.0;
  uint      n = A->nrows;

  if (h2hip_memcpy(A->d_x, b->v, sizeof(field) * n, 0) != H2HIP_OK ||
      h2hip_memcpy(A->d_y, x->v, sizeof(field) * n, 0) != H2HIP_OK)
    hip_fail("upload");
  if (h2hip_solve_gmres(addeval_device, A->H, FIELD_DOUBLES == 2, n,
                        A->d_x, A->d_y, eps, maxiter, kmax, NULL, &it,
                        &res) != H2HIP_OK)
    hip_fail("gmres");
  if (h2hip_memcpy(x->v, A->d_y, sizeof(field) * n, 1) != H2HIP_OK)
    hip_fail("download");
  return it;
}
