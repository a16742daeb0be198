/* CPU restatement of the reference's near-field quadrature and H2-matrix
 * matrix-vector product, for TEST USE ONLY.
 *
 * This is the checker.  Only tests, the smoke check and the CPU-baseline leg
 * of the benchmark may load it; the product path never does.  It follows the
 * reference (H2Lib) loop by loop; each function cites the reference file and
 * line range it restates (paths relative to the reference root).
 *
 * Pinning (what checks each part independently of this code):
 *   orc_gauss            textbook Gauss-Legendre identities (sum of weights,
 *                        exactness on polynomials) and the reference tables
 *   orc_sq_new           the reference's tables (libh2ref), sum of weights = 1/4,
 *                        exactness on polynomials over the triangle pair
 *   orc_select           brute force over all vertex permutations
 *   orc_nearfield        the reference's nearfield callbacks on the same mesh
 *                        (libh2ref), row-sum identities of the single and
 *                        double layer potential on the sphere
 *   orc_h2_*             the reference's addeval_h2matrix_avector (libh2ref),
 *                        the dense product when every block is near-field,
 *                        adjointness <Hx, y> = <x, H^* y>
 *   orc_solve_gmres      the reference's solve_gmres_avector on the same H2
 *                        system (steps and solution), the residual bound
 *                        |b - A x| <= eps |b| and a direct solve on random
 *                        dense systems
 *   orc_solve_cg         the reference's solve_cg_avector on the same H2
 *                        systems (real: steps and solution; complex: the
 *                        iterates after a fixed number of steps), the
 *                        residual bound and a direct solve on random SPD
 *                        and Hermitian positive definite systems, and
 *                        termination in at most n steps on tiny ones
 */
#ifndef ORACLE_H
#define ORACLE_H

#include <complex.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Gauss-Legendre and the triangle-pair rules ------------------------ */

/* Gauss-Legendre nodes (ascending) and weights on [-1, 1].
 * Restates Library/gaussquad.c:12-37. */
void orc_gauss(unsigned m, double *x, double *w);

/* One quadrature rule on the pair of reference triangles.  Point k of the
 * first triangle is A (1 - x[k]) + B (x[k] - x[k+n]) + C x[k+n]; likewise y
 * for the second triangle.  w[k + 9n] is the plain weight; w[k + j n] for
 * j < 9 are the basis-weighted columns (see orc_sq_weight_*). */
typedef struct {
  unsigned n;
  double *x, *y, *w;
  double base;
} orc_rule;

typedef struct {
  unsigned q, q2;
  orc_rule dist, vert, edge, id;
  /* Single-triangle rule: point (x[k], y[k]), weight w[k + 3n]. */
  orc_rule single;
} orc_sq;

/* Restates build_singquad2d, Library/singquad2d.c:25-403. */
orc_sq *orc_sq_new(unsigned q, unsigned q2);
void orc_sq_del(orc_sq *sq);

/* Basis weighting, Library/singquad2d.c:456-530.  basis_r / basis_c:
 * 0 constant, 1 linear.  Applies the ll, cl or lc weighting to all four
 * pair rules and the l weighting to the single rule. */
void orc_sq_weight(orc_sq *sq, int basis_r, int basis_c);

/* Shared-vertex classification and vertex permutation,
 * Library/singquad2d.c:549-633.  Returns the number of shared vertices and
 * the selected rule. */
unsigned orc_select(const orc_sq *sq, const unsigned *tv, const unsigned *sv,
                    unsigned *tp, unsigned *sp, const orc_rule **rule);

/* Copy of a rule for export: cls 0 dist, 1 vert, 2 edge, 3 id, 4 single.
 * x, y get 2n values (single: n), w gets the plain weight column (n). */
int orc_sq_export(const orc_sq *sq, int cls, double *x, double *y, double *w,
                  unsigned *n, double *base);

/* ---- Near-field operators ----------------------------------------------- */

typedef struct orc_bem orc_bem;

/* kernel: 0 Laplace, 1 Helmholtz.  op: 0 SLP, 1 DLP, 2 ADLP (adjoint
 * double layer).  basis: 0 constant, 1 linear.  alpha, k: complex numbers
 * as (re, im).  Restates the constructors of Library/laplacebem3d.c:
 * 843-1345 and Library/helmholtzbem3d.c:1008-1500 (kernel constant, alpha,
 * weighting, mass matrices) and Library/surface3d.c:51-108 (normals and
 * g). */
orc_bem *orc_bem_new(int kernel, int op, int basis_r, int basis_c,
                     unsigned nv, const double *x, unsigned nt,
                     const unsigned *t, unsigned q, unsigned q2,
                     const double *alpha, const double *k);
void orc_bem_del(orc_bem *b);
int orc_bem_is_complex(const orc_bem *b);
const double *orc_bem_normals(const orc_bem *b);
const double *orc_bem_gram(const orc_bem *b);

/* The nearfield / nearfield_far callback, Library/bem3d.h:380,405:
 * fills the rows x cols window of column-major N (leading dimension ld;
 * the cols x rows window with ntrans).  N holds doubles (Laplace) or
 * interleaved (re, im) pairs (Helmholtz).  ridx / cidx may be NULL.
 * Restates assemble_{cc,cl,lc,ll}_{near,far}_bem3d,
 * Library/bem3d.c:1150-1247, 1438-1528, 1798-2127, 2394-2723, 3038-3436. */
void orc_nearfield(const orc_bem *b, int far, const unsigned *ridx,
                   unsigned rows, const unsigned *cidx, unsigned cols,
                   int ntrans, void *N, size_t ld);

/* Number of kernel evaluations a nearfield call performs (cost model). */
double orc_nearfield_evals(const orc_bem *b, int far, const unsigned *ridx,
                           unsigned rows, const unsigned *cidx,
                           unsigned cols);

/* ---- H2-matrix products -------------------------------------------------- */

/* One cluster basis over a flattened cluster tree: tree[3 i + 0..2] =
 * (offset into idx, size, father or 0xffffffff), rank[i]; leaf matrices V
 * (size x rank, column-major) at voff[i]; transfer matrices E (rank_i x
 * rank_father, column-major) at i * eld. */
typedef struct {
  uint32_t nnodes;
  const uint32_t *tree, *rank, *idx;
  const double *V;
  const uint64_t *voff;
  const double *E;
  uint64_t eld;
} orc_cbasis;

/* y += alpha H x (trans = 0) or y += alpha H^* x (trans = 1).  far / near:
 * (row node, column node) pairs; S: coupling block b (rank_r x rank_c) at
 * b * sld; N: near-field blocks packed in order (rows x cols each).  Complex
 * values (S, N, x, y, alpha) are interleaved (re, im).  Restates
 * addeval_h2matrix_avector / addevaltrans_h2matrix_avector,
 * Library/h2matrix.c:1210-1364, with forward / backward_clusterbasis_avector,
 * Library/clusterbasis.c:523-573, 748-797.  Returns 0, or -1 on bad input. */
int orc_h2_addeval(int is_complex, int trans, const double *alpha,
                   const orc_cbasis *rb, const orc_cbasis *cb, uint32_t nfar,
                   const uint32_t *far, const double *S, uint64_t sld,
                   uint32_t nnear, const uint32_t *near, const double *N,
                   const double *x, double *y);

/* y += alpha H x for a self-adjoint H from its lower block triangle
 * (addevalsymm_h2matrix_avector, Library/h2matrix.c:1366-1604): blocks with
 * row node > column node act with S and S^*, diagonal leaf blocks with their
 * lower triangle and its conjugate, blocks with row node < column node are
 * ignored.  Same arguments as orc_h2_addeval.  Returns 0, or -1. */
int orc_h2_addevalsymm(int is_complex, const double *alpha,
                       const orc_cbasis *rb, const orc_cbasis *cb,
                       uint32_t nfar, const uint32_t *far, const double *S,
                       uint64_t sld, uint32_t nnear, const uint32_t *near,
                       const double *N, const double *x, double *y);

/* ---- GMRES ----------------------------------------------------------------- */

/* y += alpha A x (the reference's addeval_t, Library/krylov.h:41, real). */
typedef void (*orc_addeval_t)(double alpha, void *A, const double *x,
                              double *y);

/* Restarted GMRES(kmax) with Householder orthogonalisation for A x = b, x
 * holding the start vector; returns the steps, *residual the final
 * estimate |rhat_k|.  Restates solve_gmres_avector
 * (Library/krylovsolvers.c:183-227) with init/step/finish_gmres
 * (Library/krylov.c:535-770). */
unsigned orc_solve_gmres(orc_addeval_t addeval, void *A, size_t n,
                         const double *b, double *x, double eps,
                         unsigned maxiter, unsigned kmax, double *residual);

/* ---- conjugate gradients -------------------------------------------------- */

/* y += alpha A x with alpha = (re, im); x and y interleaved complex when the
 * field is complex. */
typedef void (*orc_addeval_c_t)(const double *alpha, void *A, const double *x,
                                double *y);

/* Conjugate gradients for A x = b (cplx: complex vectors), x holding the
 * start vector; returns the steps, *residual the final |r|.  Restates
 * solve_cg_avector (Library/krylovsolvers.c:18-52) with init_cg / step_cg
 * (Library/krylov.c:162-196). */
unsigned orc_solve_cg(orc_addeval_c_t addeval, void *A, int cplx, size_t n,
                      const double *b, double *x, double eps,
                      unsigned maxiter, double *residual);

#ifdef __cplusplus
}
#endif

#endif
