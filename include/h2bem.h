/* h2bem -- C host library around h2hip: surfaces, cluster trees, block
 * trees, operators with the reference's near-field callback signature,
 * H2-matrices built by interpolation, and Krylov solvers.
 *
 * It mirrors the reference's (H2Lib's) interfaces on this path with an h2b_
 * prefix; each declaration names the reference function it corresponds to
 * (paths relative to the H2Lib source root).  All compute-heavy work goes
 * through include/h2hip.h.
 */
#ifndef H2BEM_H
#define H2BEM_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#include "h2hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- surfaces -----------------------------------------------------------
 * Mirror of surface3d (Library/surface3d.h:45-71). */
typedef struct {
  uint32_t vertices, triangles;
  double *x;   /* 3 * vertices */
  uint32_t *t; /* 3 * triangles, counter-clockwise */
  double *n;   /* 3 * triangles, unit normals */
  double *g;   /* triangles, twice the areas */
  double hmin, hmax;
} h2b_surface;

/* Refinement of the octahedral sphere: 8 split^2 triangles, points projected
 * to the unit sphere.  new_sphere_macrosurface3d (Library/macrosurface3d.c:81)
 * + build_from_macrosurface3d_surface3d (Library/macrosurface3d.c:707). */
h2b_surface *h2b_sphere_surface(uint32_t split);
/* Refinement of the cube [-1,1]^3: 12 split^2 triangles
 * (new_cube_macrosurface3d, Library/macrosurface3d.c:390). */
h2b_surface *h2b_cube_surface(uint32_t split);
/* Cylinder of radius 1 along the x axis, |x| <= 4, closed by discs
 * (new_cylinder_macrosurface3d, Library/macrosurface3d.c:431-700). */
h2b_surface *h2b_cylinder_surface(uint32_t split);
/* Surface from arrays; computes n, g, hmin, hmax (prepare_surface3d,
 * Library/surface3d.c:51-108). */
h2b_surface *h2b_surface_from_arrays(uint32_t vertices, const double *x,
                                     uint32_t triangles, const uint32_t *t);
/* Applies x -> s R x + c to all vertices (R row-major 3x3) and updates n, g. */
void h2b_surface_transform(h2b_surface *s, const double *R, double scale,
                           const double *shift);
void h2b_surface_free(h2b_surface *s);

/* ---- cluster trees ---------------------------------------------------------
 * Mirror of cluster (Library/cluster.h:40-70), flattened: nodes in
 * depth-first order, each owning the index range [off, off + size) of the
 * tree's permutation idx.  basis: 0 constant (DOFs are triangles,
 * characteristic points are centroids), 1 linear (DOFs are vertices). */
typedef struct {
  uint32_t off, size;
  uint32_t sons;     /* 0 or 2 */
  uint32_t son[2];   /* node numbers */
  uint32_t desc;     /* nodes in the subtree */
  uint32_t level;
  double bmin[3], bmax[3];
} h2b_cnode;

typedef struct {
  uint32_t n;        /* number of DOFs */
  uint32_t *idx;     /* permutation of the DOFs */
  uint32_t nnodes;
  h2b_cnode *nodes;  /* nodes[0] is the root */
} h2b_ctree;

/* build_bem3d_cluster (Library/bem3d.c:932-954) with
 * build_adaptive_cluster (Library/cluster.c:80-168). */
h2b_ctree *h2b_build_cluster(const h2b_surface *s, int basis, uint32_t clf);
void h2b_ctree_free(h2b_ctree *t);

/* ---- block trees -------------------------------------------------------
 * build_strict_block (Library/block.c:258-322) with admissibility
 * admis = 0: admissible_max_cluster (Library/block.c:52-63),
 * admis = 1: admissible_2_cluster (Library/block.c:29-50).
 * Leaves are listed in depth-first order of the block tree (sons in the
 * reference's storage order).  Each leaf is (row node, column node). */
typedef struct {
  uint32_t nnear, nfar;
  uint32_t *near; /* 2 per leaf: row node, column node */
  uint32_t *far;
} h2b_blocks;

h2b_blocks *h2b_build_strict_block(const h2b_ctree *rt, const h2b_ctree *ct,
                                   double eta, int admis);
void h2b_blocks_free(h2b_blocks *b);
/* Near-field leaves as (row offset, rows, column offset, cols) into the
 * trees' idx arrays: the block list of h2hip_nfplan_create. */
void h2b_near_block_list(const h2b_blocks *b, const h2b_ctree *rt,
                         const h2b_ctree *ct, uint32_t *out);

/* ---- H2-matrix approximation by interpolation -----------------------------
 * setup_h2matrix_aprx_inter_bem3d (Library/bem3d.c:9658-9686) with
 * assemble_bem3d_h2matrix_row/col_clusterbasis (Library/bem3d.c:10377-
 * 10446) and the couplings of assemble_bem3d_h2matrix
 * (Library/bem3d.c:10313-10346).  Rank k = m^3 on every cluster. */
typedef struct {
  uint32_t m, k;
  double *rpts, *cpts;     /* 3 m per node: x, y, z interpolation points */
  uint64_t *rvoff, *cvoff; /* per node: leaf matrix offset, UINT64_MAX inner */
  double *rV, *cV;         /* leaf matrices, |t| x k column-major */
  uint64_t rvlen, cvlen;
  double *rE, *cE;         /* per node: k x k transfer to the father */
  uint32_t nfar;
  double *S;               /* per far block: k x k coupling, column-major */
} h2b_h2interp;

/* Cluster bases of both trees; dn_side selects the normal derivative of
 * the Lagrange polynomials (assemble_bem3d_dn_lagrange_{c,l}_amatrix,
 * Library/bem3d.c:8024-8336) for the column basis (1: double layer
 * potential) or the row basis (2: adjoint double layer, whose kernel
 * differentiates in x, Library/laplacebem3d.c:1199-1222); 0: neither.
 * q: order of the single-triangle Gauss rule. */
h2b_h2interp *h2b_build_h2interp(const h2b_surface *s, const h2b_ctree *rt,
                                 const h2b_ctree *ct, const h2b_blocks *bl,
                                 int row_basis, int col_basis, int dn_side,
                                 unsigned m, unsigned q);
/* The tensor Chebyshev points of every cluster of T (3 m per node: x,
 * then y, then z; the points h2b_build_h2interp uses, bit for bit), for
 * h2hip_h2matrix_interp_bases and h2hip_h2matrix_interp_couplings
 * (setup_interpolation_bem3d, Library/bem3d.c:4978-4996, 5179-5283). */
void h2b_interp_points(const h2b_ctree *T, unsigned m, double *out);
/* Laplace couplings S = g(xi_r, xi_c) / (4 pi) for all far blocks. */
void h2b_h2interp_coupling_laplace(h2b_h2interp *H, const h2b_blocks *bl);
/* Helmholtz couplings (complex, re / im interleaved, 2 k^2 doubles per
 * admissible block in the order of the block list) for k = kr + i ki into
 * S_out (assemble_bem3d_inter_uniform with fill_kernel_helmholtzbem3d,
 * Library/bem3d.c:6079-6108, Library/helmholtzbem3d.c:828-832). */
void h2b_h2interp_coupling_helmholtz(const h2b_h2interp *H,
                                     const h2b_blocks *bl, double kr,
                                     double ki, double *S_out);
void h2b_h2interp_free(h2b_h2interp *H);

/* ---- boundary data and functionals (piecewise linear basis) ---------------
 * kind 0: Dirichlet data 1/|x - x0|, kind 1: Neumann data
 * -<x - x0, n>/|x - x0|^3 (eval_{dirichlet,neumann}_fundamental_
 * laplacebem3d, Library/laplacebem3d.c:1390-1430, with x0 as parameter).
 * q: order of the single-triangle Gauss rule. */
/* out[i] = int f phi_i (integrate_bem3d_l_avector, Library/bem3d.c:8631) */
void h2b_integrate_l(const h2b_surface *s, unsigned q, int kind,
                     const double *x0, double *out);
/* y += alpha M x, M the mass matrix (Library/bem3d.c:8878-8908) */
void h2b_mass_l_addeval(const h2b_surface *s, double alpha, const double *x,
                        double *y);
/* L2 projection M^{-1} int f phi (projectL2_bem3d_l_avector,
 * Library/bem3d.c:9009-9088); returns the CG steps */
int h2b_projectL2_l(const h2b_surface *s, unsigned q, int kind,
                    const double *x0, double *out);
/* |sum_i x_i phi_i - f|_L2 (normL2diff_l_bem3d, Library/bem3d.c:8954) */
double h2b_normL2diff_l(const h2b_surface *s, unsigned q, const double *x,
                        int kind, const double *x0);
/* |f|_L2 (normL2_bem3d, Library/bem3d.c:8704) */
double h2b_normL2(const h2b_surface *s, unsigned q, int kind,
                  const double *x0);

/* Helmholtz point source at x0 with real wave number k, piecewise constant
 * basis, complex values as (re, im) pairs.  kind 0: Dirichlet data
 * e^{ikr}/r (rhs_dirichlet_point_helmholtzbem3d), 1: Neumann data
 * (rhs_neumann_point_helmholtzbem3d), Library/helmholtzbem3d.c:1521-1578.
 * Projection onto the constants: out[t] = mean of the data over triangle t
 * by the rule of order q (projectL2_bem3d_c_avector, Library/bem3d.c:
 * 8827-8876). */
void h2b_projectL2_c_helmholtz(const h2b_surface *s, unsigned q, int kind,
                               double k, const double *x0, double *out);
/* The integrals of the data against the constant basis: out[t] = g_t sum_q
 * w_q data(x_q) (integrate_bem3d_c_avector, Library/bem3d.c:8579-8629). */
void h2b_integrate_c_helmholtz(const h2b_surface *s, unsigned q, int kind,
                               double k, const double *x0, double *out);
/* The same data against the piecewise linear basis, out[i] = sum over the
 * triangles t at vertex i of g_t sum_q w_q phi_i(x_q) data(x_q)
 * (integrate_bem3d_l_avector, Library/bem3d.c:8631-8702); 2 vertices
 * doubles. */
void h2b_integrate_l_helmholtz(const h2b_surface *s, unsigned q, int kind,
                               double k, const double *x0, double *out);
/* out[0] = |x - data|_L2 (normL2diff_c_bem3d, Library/bem3d.c:8776-8825;
 * x may be NULL for 0), out[1] = |data|_L2 (normL2_bem3d,
 * Library/bem3d.c:8704-8751). */
void h2b_normL2_c_helmholtz(const h2b_surface *s, unsigned q,
                            const double *x, int kind, double k,
                            const double *x0, double *out);

#ifdef __cplusplus
}
#endif

#endif
