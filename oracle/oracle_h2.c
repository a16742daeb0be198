/* CPU restatement of the H2-matrix products (TEST USE ONLY; see oracle.h).
 *
 * Same flattened description as the device library's h2hip_h2matrix_*
 * calls: per cluster (offset, size, father) into the index array, a rank,
 * leaf matrices V (size x rank, column-major, at voff[node]), transfer
 * matrices E (rank x rank_father, column-major, at node * eld), coupling
 * matrices S (rank_r x rank_c at block * sld), near-field blocks packed one
 * after the other (rows x cols, column-major).  V and E are real, S, N, x
 * and y are real or complex (interleaved re, im).
 *
 * The order of operations follows the reference:
 *   forward_clusterbasis_avector   Library/clusterbasis.c:523-573
 *   fastaddeval_h2matrix_avector   Library/h2matrix.c:1210-1266
 *   backward_clusterbasis_avector  Library/clusterbasis.c:748-797
 *   addeval_h2matrix_avector       Library/h2matrix.c:1268-1286
 *   addevaltrans_h2matrix_avector  Library/h2matrix.c:1288-1364
 * with the coupling and near-field blocks taken in the order given (the
 * reference walks its block tree; any order gives the same sums up to
 * rounding), leaf near-field results collected in a per-leaf buffer that the
 * backward transform adds V yc to before scattering (the reference keeps it
 * in the "size" slots after the coefficients of a leaf).
 */
#include <complex.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NONE 0xffffffffu

typedef double complex cplx;

typedef struct {
  uint32_t n;
  const uint32_t *tree, *rank, *idx;
  const double *V, *E;
  const uint64_t *voff;
  uint64_t eld;
  uint32_t *nsons, *sonbeg, *sons;
  uint64_t *coef, *leafoff;
  uint64_t ncoef, nleaf;
} side_t;

static int side_init(side_t *s, const orc_cbasis *b) {
  s->n = b->nnodes;
  s->tree = b->tree;
  s->rank = b->rank;
  s->idx = b->idx;
  s->V = b->V;
  s->E = b->E;
  s->voff = b->voff;
  s->eld = b->eld;
  s->nsons = calloc(s->n, sizeof(uint32_t));
  s->sonbeg = calloc(s->n + 1, sizeof(uint32_t));
  s->sons = calloc(s->n ? s->n : 1, sizeof(uint32_t));
  s->coef = calloc(s->n, sizeof(uint64_t));
  s->leafoff = calloc(s->n, sizeof(uint64_t));
  if (!s->nsons || !s->sonbeg || !s->sons || !s->coef || !s->leafoff)
    return -1;
  for (uint32_t i = 0; i < s->n; i++) {
    uint32_t f = s->tree[3 * i + 2];
    if (f != NONE) {
      if (f >= s->n) return -1;
      s->nsons[f]++;
    }
  }
  for (uint32_t i = 0; i < s->n; i++)
    s->sonbeg[i + 1] = s->sonbeg[i] + s->nsons[i];
  uint32_t *fill = calloc(s->n ? s->n : 1, sizeof(uint32_t));
  if (!fill) return -1;
  /* sons in increasing node number, as the clustering numbers them */
  for (uint32_t i = 0; i < s->n; i++) {
    uint32_t f = s->tree[3 * i + 2];
    if (f != NONE) s->sons[s->sonbeg[f] + fill[f]++] = i;
  }
  free(fill);
  s->ncoef = 0;
  s->nleaf = 0;
  for (uint32_t i = 0; i < s->n; i++) {
    s->coef[i] = s->ncoef;
    s->ncoef += s->rank[i];
    if (s->nsons[i] == 0) {
      s->leafoff[i] = s->nleaf;
      s->nleaf += s->tree[3 * i + 1];
    }
  }
  return 0;
}

static void side_free(side_t *s) {
  free(s->nsons);
  free(s->sonbeg);
  free(s->sons);
  free(s->coef);
  free(s->leafoff);
}

static cplx val(const double *p, uint64_t i, int c) {
  return c ? p[2 * i] + I * p[2 * i + 1] : p[i];
}

/* Library/clusterbasis.c:523-573: leaf xc = V^T x|_t, inner
 * xc = sum_sons E_son^T xt_son (sons first) */
static void forward(const side_t *s, uint32_t t, const cplx *x, cplx *xt) {
  const uint32_t k = s->rank[t];
  cplx *xc = xt + s->coef[t];
  for (uint32_t j = 0; j < k; j++) xc[j] = 0.0;
  if (s->nsons[t]) {
    for (uint32_t so = 0; so < s->nsons[t]; so++) {
      const uint32_t c = s->sons[s->sonbeg[t] + so];
      forward(s, c, x, xt);
      const uint32_t kc = s->rank[c];
      const double *E = s->E + (uint64_t)c * s->eld; /* kc x k */
      const cplx *x1 = xt + s->coef[c];
      for (uint32_t j = 0; j < k; j++) {
        cplx sum = 0.0;
        for (uint32_t i = 0; i < kc; i++) sum += E[i + (uint64_t)j * kc] * x1[i];
        xc[j] += sum;
      }
    }
  } else {
    const uint32_t off = s->tree[3 * t], size = s->tree[3 * t + 1];
    const double *V = s->V + s->voff[t]; /* size x k */
    for (uint32_t j = 0; j < k; j++) {
      cplx sum = 0.0;
      for (uint32_t i = 0; i < size; i++)
        sum += V[i + (uint64_t)j * size] * x[s->idx[off + i]];
      xc[j] += sum;
    }
  }
}

/* Library/clusterbasis.c:748-797: inner yt_son += E_son yc, leaf
 * yp += V yc, y[idx] += yp */
static void backward(const side_t *s, uint32_t t, cplx *yt, cplx *yp,
                     cplx *y) {
  const uint32_t k = s->rank[t];
  const cplx *yc = yt + s->coef[t];
  if (s->nsons[t]) {
    for (uint32_t so = 0; so < s->nsons[t]; so++) {
      const uint32_t c = s->sons[s->sonbeg[t] + so];
      const uint32_t kc = s->rank[c];
      const double *E = s->E + (uint64_t)c * s->eld; /* kc x k */
      cplx *y1 = yt + s->coef[c];
      for (uint32_t j = 0; j < k; j++)
        for (uint32_t i = 0; i < kc; i++)
          y1[i] += E[i + (uint64_t)j * kc] * yc[j];
      backward(s, c, yt, yp, y);
    }
  } else {
    const uint32_t off = s->tree[3 * t], size = s->tree[3 * t + 1];
    const double *V = s->V + s->voff[t];
    cplx *p = yp + s->leafoff[t];
    for (uint32_t j = 0; j < k; j++)
      for (uint32_t i = 0; i < size; i++) p[i] += V[i + (uint64_t)j * size] * yc[j];
    for (uint32_t i = 0; i < size; i++) y[s->idx[off + i]] += p[i];
  }
}

int orc_h2_addeval(int is_complex, int trans, const double *alpha,
                   const orc_cbasis *rb, const orc_cbasis *cb, uint32_t nfar,
                   const uint32_t *far, const double *S, uint64_t sld,
                   uint32_t nnear, const uint32_t *near, const double *N,
                   const double *x, double *y) {
  side_t r, c;
  memset(&r, 0, sizeof r);
  memset(&c, 0, sizeof c);
  int rc = -1;
  if (side_init(&r, rb) || side_init(&c, cb)) goto out;
  /* in: side the product reads from, out: side it writes to */
  const side_t *in = trans ? &r : &c, *outs = trans ? &c : &r;
  const int cm = is_complex ? 1 : 0;
  const cplx a = alpha[0] + (is_complex ? I * alpha[1] : 0.0);
  uint64_t nin = 0, nout = 0;
  for (uint32_t i = 0; i < in->n; i++)
    if (in->tree[3 * i + 2] == NONE) nin += in->tree[3 * i + 1];
  for (uint32_t i = 0; i < outs->n; i++)
    if (outs->tree[3 * i + 2] == NONE) nout += outs->tree[3 * i + 1];
  cplx *xv = calloc(nin + 1, sizeof(cplx));
  cplx *yv = calloc(nout + 1, sizeof(cplx));
  cplx *xt = calloc(in->ncoef + 1, sizeof(cplx));
  cplx *yt = calloc(outs->ncoef + 1, sizeof(cplx));
  cplx *yp = calloc(outs->nleaf + 1, sizeof(cplx));
  if (!xv || !yv || !xt || !yt || !yp) goto out2;
  for (uint64_t i = 0; i < nin; i++) xv[i] = val(x, i, cm);
  /* forward transform from every root of the input side */
  for (uint32_t i = 0; i < in->n; i++)
    if (in->tree[3 * i + 2] == NONE) forward(in, i, xv, xt);
  /* Library/h2matrix.c:1210-1266 (u: yt += alpha S xt, f: yp += alpha f xp),
   * Library/h2matrix.c:1288-1343 for the adjoint */
  for (uint32_t b = 0; b < nfar; b++) {
    const uint32_t tr = far[2 * b], tc = far[2 * b + 1];
    const uint32_t kr = r.rank[tr], kc = c.rank[tc];
    const uint64_t so = (uint64_t)b * sld;
    if (!trans) {
      cplx *o = yt + r.coef[tr];
      const cplx *i0 = xt + c.coef[tc];
      for (uint32_t j = 0; j < kc; j++)
        for (uint32_t i = 0; i < kr; i++)
          o[i] += a * val(S, so + i + (uint64_t)j * kr, cm) * i0[j];
    } else {
      cplx *o = yt + c.coef[tc];
      const cplx *i0 = xt + r.coef[tr];
      for (uint32_t j = 0; j < kc; j++) {
        cplx sum = 0.0;
        for (uint32_t i = 0; i < kr; i++)
          sum += conj(val(S, so + i + (uint64_t)j * kr, cm)) * i0[i];
        o[j] += a * sum;
      }
    }
  }
  uint64_t no = 0;
  for (uint32_t b = 0; b < nnear; b++) {
    const uint32_t tr = near[2 * b], tc = near[2 * b + 1];
    const uint32_t roff = r.tree[3 * tr], rows = r.tree[3 * tr + 1];
    const uint32_t coff = c.tree[3 * tc], cols = c.tree[3 * tc + 1];
    if (!trans) {
      if (r.nsons[tr] || c.nsons[tc]) goto out2; /* near blocks are leaves */
      cplx *o = yp + r.leafoff[tr];
      for (uint32_t j = 0; j < cols; j++) {
        const cplx xj = xv[c.idx[coff + j]];
        for (uint32_t i = 0; i < rows; i++)
          o[i] += a * val(N, no + i + (uint64_t)j * rows, cm) * xj;
      }
    } else {
      if (r.nsons[tr] || c.nsons[tc]) goto out2;
      cplx *o = yp + c.leafoff[tc];
      for (uint32_t j = 0; j < cols; j++) {
        cplx sum = 0.0;
        for (uint32_t i = 0; i < rows; i++)
          sum += conj(val(N, no + i + (uint64_t)j * rows, cm)) *
                 xv[r.idx[roff + i]];
        o[j] += a * sum;
      }
    }
    no += (uint64_t)rows * cols;
  }
  for (uint32_t i = 0; i < outs->n; i++)
    if (outs->tree[3 * i + 2] == NONE) backward(outs, i, yt, yp, yv);
  for (uint64_t i = 0; i < nout; i++) {
    if (cm) {
      y[2 * i] += creal(yv[i]);
      y[2 * i + 1] += cimag(yv[i]);
    } else {
      y[i] += creal(yv[i]);
    }
  }
  rc = 0;
out2:
  free(xv);
  free(yv);
  free(xt);
  free(yt);
  free(yp);
out:
  side_free(&r);
  side_free(&c);
  return rc;
}

/* y += alpha H x for a self-adjoint H given by its lower block triangle
 * (addevalsymm_h2matrix_avector, Library/h2matrix.c:1570-1604, with
 * addevalsymm_diag / addevalsymm_offdiag, Library/h2matrix.c:1366-1567):
 * the blocks (r, c) with r > c (node numbers; sons are numbered in order,
 * so this is the reference's "row son after column son") act as
 * S xt_c -> yt_r and S^* xta_r -> yta_c, diagonal leaf blocks (r == c)
 * with their lower triangle and its conjugate; blocks with r < c are
 * skipped.  xt / yta use the column basis, xta / yt the row basis. */
int orc_h2_addevalsymm(int is_complex, const double *alpha,
                       const orc_cbasis *rb, const orc_cbasis *cb,
                       uint32_t nfar, const uint32_t *far, const double *S,
                       uint64_t sld, uint32_t nnear, const uint32_t *near,
                       const double *N, const double *x, double *y) {
  side_t r, c;
  memset(&r, 0, sizeof r);
  memset(&c, 0, sizeof c);
  int rc = -1;
  if (side_init(&r, rb) || side_init(&c, cb) || r.n != c.n) goto out;
  const int cm = is_complex ? 1 : 0;
  const cplx a = alpha[0] + (is_complex ? I * alpha[1] : 0.0);
  uint64_t n = 0;
  for (uint32_t i = 0; i < c.n; i++)
    if (c.tree[3 * i + 2] == NONE) n += c.tree[3 * i + 1];
  cplx *xv = calloc(n + 1, sizeof(cplx));
  cplx *yv = calloc(n + 1, sizeof(cplx));
  cplx *xt = calloc(c.ncoef + 1, sizeof(cplx));
  cplx *xta = calloc(r.ncoef + 1, sizeof(cplx));
  cplx *yt = calloc(r.ncoef + 1, sizeof(cplx));
  cplx *yta = calloc(c.ncoef + 1, sizeof(cplx));
  cplx *yp = calloc(r.nleaf + 1, sizeof(cplx));
  cplx *ypa = calloc(c.nleaf + 1, sizeof(cplx));
  if (!xv || !yv || !xt || !xta || !yt || !yta || !yp || !ypa) goto out2;
  for (uint64_t i = 0; i < n; i++) xv[i] = val(x, i, cm);
  for (uint32_t i = 0; i < c.n; i++)
    if (c.tree[3 * i + 2] == NONE) {
      forward(&c, i, xv, xt);
      forward(&r, i, xv, xta);
    }
  for (uint32_t b = 0; b < nfar; b++) {
    const uint32_t tr = far[2 * b], tc = far[2 * b + 1];
    if (tr <= tc) continue;
    const uint32_t kr = r.rank[tr], kc = c.rank[tc];
    const uint64_t so = (uint64_t)b * sld;
    cplx *o = yt + r.coef[tr], *oa = yta + c.coef[tc];
    const cplx *i0 = xt + c.coef[tc], *ia = xta + r.coef[tr];
    for (uint32_t j = 0; j < kc; j++) {
      cplx sum = 0.0;
      for (uint32_t i = 0; i < kr; i++) {
        const cplx sv = val(S, so + i + (uint64_t)j * kr, cm);
        o[i] += a * sv * i0[j];
        sum += conj(sv) * ia[i];
      }
      oa[j] += a * sum;
    }
  }
  uint64_t no = 0;
  for (uint32_t b = 0; b < nnear; b++) {
    const uint32_t tr = near[2 * b], tc = near[2 * b + 1];
    const uint32_t roff = r.tree[3 * tr], rows = r.tree[3 * tr + 1];
    const uint32_t coff = c.tree[3 * tc], cols = c.tree[3 * tc + 1];
    if (r.nsons[tr] || c.nsons[tc]) goto out2;
    if (tr >= tc) {
      cplx *o = yp + r.leafoff[tr], *oa = ypa + c.leafoff[tc];
      for (uint32_t j = 0; j < cols; j++) {
        const cplx xj = xv[c.idx[coff + j]];
        /* Library/h2matrix.c:1486-1492 on the diagonal: entry (j, j) once,
         * i > j as a_ij and conj(a_ij) */
        const uint32_t i0 = tr == tc ? j : 0;
        cplx sum = 0.0;
        for (uint32_t i = i0; i < rows; i++) {
          const cplx av = val(N, no + i + (uint64_t)j * rows, cm);
          o[i] += a * av * xj;
          if (tr != tc || i > j) sum += conj(av) * xv[r.idx[roff + i]];
        }
        oa[j] += a * sum;
      }
    }
    no += (uint64_t)rows * cols;
  }
  for (uint32_t i = 0; i < r.n; i++)
    if (r.tree[3 * i + 2] == NONE) {
      backward(&r, i, yt, yp, yv);
      backward(&c, i, yta, ypa, yv);
    }
  for (uint64_t i = 0; i < n; i++) {
    if (cm) {
      y[2 * i] += creal(yv[i]);
      y[2 * i + 1] += cimag(yv[i]);
    } else {
      y[i] += creal(yv[i]);
    }
  }
  rc = 0;
out2:
  free(xv);
  free(yv);
  free(xt);
  free(xta);
  free(yt);
  free(yta);
  free(yp);
  free(ypa);
out:
  side_free(&r);
  side_free(&c);
  return rc;
}
