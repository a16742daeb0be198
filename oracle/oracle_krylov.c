/* CPU restatement of the reference's restarted GMRES (real field) and
 * conjugate gradients (real or complex field), TEST USE ONLY (see oracle.h).
 *
 * Follows, step by step:
 *   solve_gmres_avector             Library/krylovsolvers.c:183-227
 *   init_gmres / step_gmres /
 *   finish_gmres                    Library/krylov.c:612-770
 *   findapply_givens / apply_givens Library/krylov.c:535-610
 *   qrdecomp_amatrix (no BLAS)      Library/factorizations.c:2179-2230
 *   qreval_amatrix_avector (no BLAS) Library/factorizations.c:2822-2870
 *   uppersolve_amatrix_avector      Library/factorizations.c:312-338
 * The Krylov matrix qr (n x kmax, column-major) holds the Householder
 * vectors below its diagonal, R shifted by one column above it, and the
 * Givens parameters on its diagonal, exactly as the reference keeps them.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define ALMOST_ZERO 1e-300

static double sgn1(double x) { return x < 0.0 ? -1.0 : 1.0; }

/* Householder QR of the rows x cols block at a (leading dimension lda);
 * Library/factorizations.c:2179-2230 */
static void qrdecomp(double *a, size_t lda, size_t rows, size_t cols,
                     double *tau) {
  const size_t refl = rows < cols ? rows : cols;
  for (size_t k = 0; k < refl; k++) {
    double norm2 = 0.0;
    for (size_t i = k; i < rows; i++) norm2 += a[i + k * lda] * a[i + k * lda];
    const double norm = sqrt(norm2);
    if (norm2 <= ALMOST_ZERO) {
      tau[k] = 0.0;
      continue;
    }
    const double diag = a[k + k * lda];
    const double adiag = fabs(diag) + norm;
    const double alpha = -sgn1(diag) * norm;
    double beta = 1.0 / (norm2 + fabs(diag) * norm);
    beta *= adiag * adiag;
    const double gamma = sgn1(diag) / adiag;
    for (size_t i = k + 1; i < rows; i++) a[i + k * lda] *= gamma;
    tau[k] = beta;
    a[k + k * lda] = alpha;
    for (size_t j = k + 1; j < cols; j++) {
      double g = a[k + j * lda];
      for (size_t i = k + 1; i < rows; i++) g += a[i + k * lda] * a[i + j * lda];
      g *= beta;
      a[k + j * lda] -= g;
      for (size_t i = k + 1; i < rows; i++) a[i + j * lda] -= g * a[i + k * lda];
    }
  }
}

/* x := Q^T x (qtrans) or Q x with the reflections stored in the rows x cols
 * block at a; Library/factorizations.c:2822-2870 */
static void qreval(int qtrans, const double *a, size_t lda, size_t rows,
                   size_t cols, const double *tau, size_t ntau, double *x) {
  size_t refl = rows < cols ? rows : cols;
  if (ntau < refl) refl = ntau;
  for (size_t kk = 0; kk < refl; kk++) {
    const size_t k = qtrans ? kk : refl - 1 - kk;
    const double beta = tau[k];
    if (beta == 0.0) continue;
    double g = x[k];
    for (size_t i = k + 1; i < rows; i++) g += a[i + k * lda] * x[i];
    g *= beta;
    x[k] -= g;
    for (size_t i = k + 1; i < rows; i++) x[i] -= g * a[i + k * lda];
  }
}

/* Library/krylov.c:535-585 */
static double findapply_givens(double *a, double *b) {
  const double aa = *a, bb = *b;
  double c, s, t, rho;
  if (bb * bb < aa * aa) {
    t = bb / aa;
    c = 1.0 / sqrt(t * t + 1.0);
    s = t * c;
    rho = s;
  } else if (aa * aa < bb * bb) {
    t = aa / bb;
    s = 1.0 / sqrt(t * t + 1.0);
    c = t * s;
    rho = 1.0 / c;
  } else {
    c = 1.0;
    s = 0.0;
    rho = 1.0;
  }
  *a = c * aa + s * bb;
  *b = 0.0;
  return rho;
}

/* Library/krylov.c:587-610 */
static void apply_givens(double rho, double *a, double *b) {
  const double aa = *a, bb = *b;
  double c, s;
  if (rho * rho < 1.0) {
    s = rho;
    c = sqrt(1.0 - s * s);
  } else if (rho * rho > 1.0) {
    c = 1.0 / rho;
    s = sqrt(1.0 - c * c);
  } else {
    c = 1.0;
    s = 0.0;
  }
  *a = c * aa + s * bb;
  *b = c * bb - s * aa;
}

typedef struct {
  orc_addeval_t addeval;
  void *A;
  size_t n;
  unsigned kmax;
  double *qr, *tau, *rhat, *q;
  unsigned k;
} gm;

/* Library/krylov.c:612-656 */
static void init_gmres(gm *G, const double *b, double *x) {
  const size_t n = G->n;
  double *r = G->qr;
  memcpy(r, b, sizeof(double) * n);
  G->addeval(-1.0, G->A, x, r);
  qrdecomp(G->qr, n, n, 1, G->tau);
  memset(G->q, 0, sizeof(double) * n);
  G->q[0] = 1.0;
  qreval(0, G->qr, n, n, 1, G->tau, G->kmax, G->q);
  memset(G->rhat, 0, sizeof(double) * n);
  G->rhat[0] = G->qr[0];
  G->k = 0;
}

/* Library/krylov.c:658-725 */
static void step_gmres(gm *G) {
  const size_t n = G->n, ld = n;
  const unsigned k = G->k;
  if (k + 1 >= G->kmax) return;
  double *a = G->qr + (size_t)(k + 1) * ld;
  memset(a, 0, sizeof(double) * n);
  G->addeval(1.0, G->A, G->q, a);
  qreval(1, G->qr, ld, n, k + 1, G->tau, G->kmax, a);
  qrdecomp(G->qr + (k + 1) + (size_t)(k + 1) * ld, ld, n - (k + 1), 1,
           G->tau + k + 1);
  memset(G->q, 0, sizeof(double) * n);
  G->q[k + 1] = 1.0;
  qreval(0, G->qr, ld, n, k + 2, G->tau, G->kmax, G->q);
  for (unsigned i = 0; i < k; i++) {
    const double rho = G->qr[(i + 1) + (size_t)(i + 1) * ld];
    apply_givens(rho, G->qr + i + (size_t)(k + 1) * ld,
                 G->qr + (i + 1) + (size_t)(k + 1) * ld);
  }
  const double rho = findapply_givens(G->qr + k + (size_t)(k + 1) * ld,
                                      G->qr + (k + 1) + (size_t)(k + 1) * ld);
  G->qr[(k + 1) + (size_t)(k + 1) * ld] = rho;
  apply_givens(rho, G->rhat + k, G->rhat + (k + 1));
  G->k = k + 1;
}

/* Library/krylov.c:727-762 */
static void finish_gmres(gm *G, const double *b, double *x) {
  const size_t n = G->n, ld = n;
  const unsigned k = G->k;
  /* R y = rhat_k with R in rows 0..k-1 of columns 1..k */
  for (unsigned j = k; j-- > 0;) {
    G->rhat[j] /= G->qr[j + (size_t)(j + 1) * ld];
    for (unsigned i = 0; i < j; i++)
      G->rhat[i] -= G->qr[i + (size_t)(j + 1) * ld] * G->rhat[j];
  }
  for (size_t i = k; i < n; i++) G->rhat[i] = 0.0;
  qreval(0, G->qr, ld, n, k, G->tau, G->kmax, G->rhat);
  for (size_t i = 0; i < n; i++) x[i] += G->rhat[i];
  init_gmres(G, b, x);
}

unsigned orc_solve_gmres(orc_addeval_t addeval, void *A, size_t n,
                         const double *b, double *x, double eps,
                         unsigned maxiter, unsigned kmax, double *residual) {
  gm G;
  G.addeval = addeval;
  G.A = A;
  G.n = n;
  G.kmax = kmax;
  G.qr = calloc(n * (size_t)kmax, sizeof(double));
  G.tau = calloc(kmax, sizeof(double));
  G.rhat = calloc(n, sizeof(double));
  G.q = calloc(n, sizeof(double));
  if (!G.qr || !G.tau || !G.rhat || !G.q || kmax < 1) {
    free(G.qr);
    free(G.tau);
    free(G.rhat);
    free(G.q);
    return 0;
  }
  double norm = 0.0;
  for (size_t i = 0; i < n; i++) norm += b[i] * b[i];
  norm = sqrt(norm);
  /* Library/krylovsolvers.c:183-227 */
  init_gmres(&G, b, x);
  double error = fabs(G.rhat[G.k]);
  unsigned iter = 0;
  while (error > eps * norm && iter + 1 != maxiter) {
    if (G.k + 1 >= kmax) finish_gmres(&G, b, x);
    step_gmres(&G);
    error = fabs(G.rhat[G.k]);
    iter++;
  }
  finish_gmres(&G, b, x);
  if (residual) *residual = error;
  free(G.qr);
  free(G.tau);
  free(G.rhat);
  free(G.q);
  return iter;
}

/* ---- conjugate gradients ------------------------------------------------
 * solve_cg_avector (Library/krylovsolvers.c:18-52) with init_cg and step_cg
 * (Library/krylov.c:162-196), real or complex field (dotprod_avector
 * conjugates its first argument, Library/avector.c:366-379). */

static double complex cg_dot(int cplx, const double *u, const double *v,
                             size_t n) {
  double complex s = 0.0;
  if (cplx) {
    const double complex *a = (const double complex *)u;
    const double complex *b = (const double complex *)v;
    for (size_t i = 0; i < n; i++) s += conj(a[i]) * b[i];
  } else {
    for (size_t i = 0; i < n; i++) s += u[i] * v[i];
  }
  return s;
}

/* y += alpha x */
static void cg_axpy(int cplx, double complex alpha, const double *x,
                    double *y, size_t n) {
  if (cplx) {
    const double complex *a = (const double complex *)x;
    double complex *b = (double complex *)y;
    for (size_t i = 0; i < n; i++) b[i] += alpha * a[i];
  } else {
    for (size_t i = 0; i < n; i++) y[i] += creal(alpha) * x[i];
  }
}

static void cg_scale(int cplx, double complex alpha, double *x, size_t n) {
  if (cplx) {
    double complex *a = (double complex *)x;
    for (size_t i = 0; i < n; i++) a[i] *= alpha;
  } else {
    for (size_t i = 0; i < n; i++) x[i] *= creal(alpha);
  }
}

static void cg_apply(orc_addeval_c_t addeval, void *A, double complex alpha,
                     const double *x, double *y) {
  const double a[2] = {creal(alpha), cimag(alpha)};
  addeval(a, A, x, y);
}

unsigned orc_solve_cg(orc_addeval_c_t addeval, void *A, int cplx, size_t n,
                      const double *b, double *x, double eps,
                      unsigned maxiter, double *residual) {
  const size_t len = (cplx ? 2 : 1) * n;
  double *r = calloc(len, sizeof(double));
  double *p = calloc(len, sizeof(double));
  double *a = calloc(len, sizeof(double));
  if (!r || !p || !a) {
    free(r);
    free(p);
    free(a);
    return 0;
  }
  const double norm = sqrt(creal(cg_dot(cplx, b, b, n)));
  /* init_cg: r = b - A x, p = r */
  memcpy(r, b, len * sizeof(double));
  cg_apply(addeval, A, -1.0, x, r);
  memcpy(p, r, len * sizeof(double));
  double error = sqrt(creal(cg_dot(cplx, r, r, n)));
  unsigned iter = 0;
  while (error > eps * norm && iter + 1 != maxiter) {
    /* step_cg */
    memset(a, 0, len * sizeof(double));
    cg_apply(addeval, A, 1.0, p, a);
    const double complex gamma = cg_dot(cplx, p, a, n);
    const double complex lambda = cg_dot(cplx, p, r, n) / gamma;
    cg_axpy(cplx, lambda, p, x, n);
    cg_axpy(cplx, -lambda, a, r, n);
    const double complex mu = cg_dot(cplx, a, r, n) / gamma;
    cg_scale(cplx, -mu, p, n);
    cg_axpy(cplx, 1.0, r, p, n);
    error = sqrt(creal(cg_dot(cplx, r, r, n)));
    iter++;
  }
  if (residual) *residual = error;
  free(r);
  free(p);
  free(a);
  return iter;
}
