/* h2hip -- MI355X (gfx950) near-field Galerkin quadrature and H2-matrix
 * matrix-vector products for 3-D boundary element methods.
 *
 * Plain C ABI: opaque handles, plain pointers and sizes, integer status
 * codes.  All functions return H2HIP_OK (0) on success and a negative code
 * on failure; h2hip_last_error() returns a message for the calling thread.
 *
 * Complex values (Helmholtz operators) are stored as interleaved
 * (re, im) pairs of doubles, i.e. the layout of C99 double complex.
 * Matrices are column-major with a leading dimension, as in the reference's
 * amatrix (Library/amatrix.h:43-57).
 *
 * "stream" arguments are hipStream_t values passed as void* (NULL: the
 * default stream).  Device pointers are plain void* as returned by
 * hipMalloc or by a framework's allocator.
 *
 * Every entry point below names the reference interface it replaces
 * (paths relative to the H2Lib source root).
 */
#ifndef H2HIP_H
#define H2HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define H2HIP_OK 0
#define H2HIP_EINVAL (-1)
#define H2HIP_EDEVICE (-2)
#define H2HIP_ENOMEM (-3)

/* kernel families */
#define H2HIP_LAPLACE 0
#define H2HIP_HELMHOLTZ 1
/* layer potentials */
#define H2HIP_SLP 0
#define H2HIP_DLP 1
#define H2HIP_ADLP 2 /* adjoint double layer */
/* basis functions (basisfunctionbem3d, Library/bem3d.h) */
#define H2HIP_BASIS_CONSTANT 0
#define H2HIP_BASIS_LINEAR 1

typedef struct h2hip_surface h2hip_surface;
typedef struct h2hip_op h2hip_op;
typedef struct h2hip_nfplan h2hip_nfplan;
typedef struct h2hip_h2matrix h2hip_h2matrix;

/* Message describing the last failure on the calling thread. */
const char *h2hip_last_error(void);

/* Device number used by the library (hipGetDevice of the caller). */
int h2hip_device_count(void);

/* Reproducible mode for the whole process (default off).  On: the H2
 * products and the Krylov solvers' inner products add their contributions
 * in a fixed order instead of with floating-point atomics, so a product or
 * a solve gives the same bits on every run, and a vector of a batched
 * product (h2hip_addeval_h2matrix_batch, the batched solvers) gets the bits
 * of the single product; the products cost about one more reduction pass.
 * Read when work is enqueued.  No counterpart in the reference (its
 * products are sequential). */
int h2hip_set_reproducible(int on);
int h2hip_get_reproducible(void);

/* ---- surfaces ------------------------------------------------------------
 * Device copy of a triangulated surface: vertices x[3*vertices], triangles
 * t[3*triangles] (counter-clockwise), unit normals n[3*triangles] and Gram
 * determinants g[triangles] (twice the triangle areas).  n and g may be
 * NULL, then they are computed as prepare_surface3d does.
 * Replaces the device-side view of surface3d (Library/surface3d.h:45-71,
 * Library/surface3d.c:51-108) and the vertex-to-triangle map
 * setup_vertex_to_triangle_map_bem3d (Library/bem3d.c:779-804). */
int h2hip_surface_create(h2hip_surface **out, uint32_t vertices,
                         const double *x, uint32_t triangles,
                         const uint32_t *t, const double *n,
                         const double *g);
void h2hip_surface_destroy(h2hip_surface *s);
/* Copies the normals and Gram determinants held by the surface. */
int h2hip_surface_geometry(const h2hip_surface *s, double *n, double *g);

/* ---- operators -------------------------------------------------------------
 * Near-field quadrature state of one boundary integral operator.
 * kernel: H2HIP_LAPLACE / H2HIP_HELMHOLTZ; layer: H2HIP_SLP / H2HIP_DLP /
 * H2HIP_ADLP (kernel <y - x, n_x>/|x - y|^3, times e^{ikr} (1 - ikr) for
 * Helmholtz, with n_x the normal of the row triangle); bases:
 * H2HIP_BASIS_*; q_regular / q_singular: Gauss orders of the regular and
 * the singular (Sauter-Schwab) rules; alpha: coefficient of the identity
 * term added to the double layers (ignored for SLP); k: wave number
 * (Helmholtz only).  alpha and k are (re, im) pairs and may be NULL (0).
 * Replaces new_slp_laplace_bem3d / new_dlp_laplace_bem3d /
 * new_adlp_laplace_bem3d (Library/laplacebem3d.c:843-1345) and
 * new_slp_helmholtz_bem3d / new_dlp_helmholtz_bem3d /
 * new_adlp_helmholtz_bem3d (Library/helmholtzbem3d.c:1008-1514), together
 * with build_singquad2d (Library/singquad2d.c:320-403) and the basis
 * weighting of the rules (Library/singquad2d.c:456-530). */
int h2hip_op_create(h2hip_op **out, const h2hip_surface *s, int kernel,
                    int layer, int row_basis, int col_basis,
                    uint32_t q_regular, uint32_t q_singular,
                    const double *alpha, const double *k);
void h2hip_op_destroy(h2hip_op *op);
/* Precision of the kernel evaluations: 0 FP64 (default), 1 or 2 mixed.
 * Mixed:
 * kernel values (1/r, the phase, e^{ikr}) in FP32 for every pair; regular
 * pairs form x - y from an FP64 difference of anchor points plus FP32
 * offsets and sum the inner rule dimension in FP32, the outer in FP64;
 * singular pairs keep the Sauter-Schwab points, x - y and the sums in FP64.
 * Entries agree with FP64 to about 1e-6 relative.  2: as 1, but the
 * regular pairs accumulate in FP32 throughout, the sums over the outer
 * rule dimension compensated (Kahan); singular pairs as in 1.  No
 * counterpart in the reference (it computes in its build's field type
 * only). */
int h2hip_op_set_precision(h2hip_op *op, int mode);

/* ---- the near-field callback ---------------------------------------------
 * Fills the rows x cols window of the column-major HOST array N (leading
 * dimension ld) with the Galerkin matrix entries for the row DOFs ridx and
 * the column DOFs cidx (NULL: 0, 1, 2, ...).  With ntrans != 0 the
 * transposed cols x rows window is written (conjugated for the cc and ll
 * basis pairs, as in the reference).  far != 0 uses the regular rule for all
 * pairs (nearfield_far).  Index arrays that repeat DOFs give the
 * reference's entries (constant basis: identity terms in its overwrite
 * order; linear basis: a support triangle counts only the first three
 * block positions that touch it).  Blocks until N is written.
 * Thread-safe: calls from several host threads run concurrently (up to
 * four per operator, each on its own stream, with device and pinned host
 * buffers kept between calls).
 * Replaces the callbacks bem->nearfield and bem->nearfield_far
 * (Library/bem3d.h:380-381, 405-406), i.e. assemble_{cc,cl,lc,ll}_near_bem3d
 * and assemble_{cc,cl,lc,ll}_far_bem3d (Library/bem3d.c:1150-3436). */
int h2hip_nearfield(const h2hip_op *op, int far, const uint32_t *ridx,
                    uint32_t rows, const uint32_t *cidx, uint32_t cols,
                    int ntrans, void *N, size_t ld);
/* The same callback, deferred: records the block (the index arrays are
 * copied; N must stay valid and untouched until the flush) and returns at
 * once without writing N.  Thread-safe; calls from several threads append
 * to one queue per operator.  The queued blocks are computed and written by
 * h2hip_nearfield_flush.  Replaces the enqueueing callbacks of the
 * reference's OpenCL path (Library/laplaceoclbem3d.c:248-279). */
int h2hip_nearfield_defer(const h2hip_op *op, int far, const uint32_t *ridx,
                          uint32_t rows, const uint32_t *cidx, uint32_t cols,
                          int ntrans, void *N, size_t ld);
/* Computes every block deferred on op since the last flush (one plan and
 * one launch set per rule) and writes each into its host window; returns
 * when all are written.  Blocks deferred while the flush runs wait for the
 * next one.  Replaces the wait for the OpenCL scheduler's queues
 * (Library/opencl.h:359-367). */
int h2hip_nearfield_flush(const h2hip_op *op);

/* ---- batched, device-resident near field -----------------------------------
 * A plan covers many blocks of one index set pair at once: block b is the
 * rows x cols window ridx_all[roff .. roff+rows) x cidx_all[coff ..
 * coff+cols), given as blocks[4b .. 4b+3] = (roff, rows, coff, cols).  The
 * packed output stores block b column-major with leading dimension rows at
 * element offset off[b] (h2hip_nfplan_offsets).  This is the near-field
 * part of assemble_bem3d_h2matrix (Library/bem3d.c:10313-10346), which calls
 * bem->nearfield for every inadmissible leaf block. */
int h2hip_nfplan_create(h2hip_nfplan **out, const h2hip_surface *s,
                        int row_basis, int col_basis,
                        const uint32_t *ridx_all, uint32_t nridx,
                        const uint32_t *cidx_all, uint32_t ncidx,
                        uint32_t nblocks, const uint32_t *blocks);
void h2hip_nfplan_destroy(h2hip_nfplan *p);
/* Total number of entries of all blocks. */
uint64_t h2hip_nfplan_entries(const h2hip_nfplan *p);
/* Element offsets of the blocks in the packed output (nblocks values). */
int h2hip_nfplan_offsets(const h2hip_nfplan *p, uint64_t *off);
/* Schedule of the plan's launches (no counterpart in the reference; the
 * entries do not depend on it beyond the order of atomic sums, which
 * varies from launch to launch anyway for the linear basis; constant-basis
 * entries are the same bit for bit).  sorted = 1 orders the tiles
 * by decreasing regular work (longest first); persistent = 1 runs the
 * regular kernel as a persistent grid of resident workgroups that take
 * tiles from a device counter.  Not thread-safe against concurrent
 * executions of the same plan. */
int h2hip_nfplan_set_schedule(h2hip_nfplan *p, int sorted, int persistent);
/* Fused tiles (no counterpart in the reference): fused = 1 evaluates each
 * tile's pairs with common vertices in the tile's own workgroup, added into
 * its LDS accumulator before the tile is stored, so one launch (one
 * workgroup per tile, or the persistent grid when persistent = 1) assembles
 * every block, tiles taken by decreasing total work; no global atomics.
 * Applies to regular order 2 with singular order 4 (other orders keep the
 * launches per rule class).  Same entries up to the order of the sums. */
int h2hip_nfplan_set_fused(h2hip_nfplan *p, int fused);
/* Tile edge of the plan (8..64 DOFs, default 64; no counterpart in the
 * reference): the blocks are cut again into tiles of at most tile x tile
 * DOFs, the schedule kept.  Smaller tiles give more workgroups for small
 * blocks; with the linear basis a triangle pair whose vertices fall into
 * several tiles is evaluated once per tile.  The packed output and the
 * entries (up to the order of the sums) do not change. */
int h2hip_nfplan_set_tile(h2hip_nfplan *p, uint32_t tile);
/* Kernel evaluations per rule class (regular, vertex, edge, identical) and
 * triangle pairs per class, as the reference's loops perform them
 * (counted on the host; out[0..3] evaluations, out[4..7] pairs). */
int h2hip_nfplan_work(const h2hip_nfplan *p, const h2hip_op *op, int far,
                      double *out);
/* Computes all blocks of the plan into the device array d_out (doubles or
 * complex pairs, h2hip_nfplan_entries elements), on the given stream,
 * without synchronising. */
int h2hip_nearfield_execute(const h2hip_op *op, const h2hip_nfplan *p,
                            int far, void *d_out, void *stream);

/* ---- H2-matrices -----------------------------------------------------------
 * Device H2-matrix: a row and a column cluster tree, nested cluster bases
 * (leaf matrices V and transfer matrices E, real), couplings S of the
 * admissible leaves and dense near-field blocks N of the inadmissible
 * leaves (real for Laplace, complex for Helmholtz).  Replaces h2matrix,
 * clusterbasis and uniform (Library/h2matrix.h:48-71,
 * Library/clusterbasis.h:40-71, Library/uniform.h:69-84) and
 * build_from_block_h2matrix (Library/h2matrix.c:438-485).
 *
 * Trees: node i has tree[3i .. 3i+2] = (offset into idx, size, father),
 * father = 0xffffffff for the root; rank[i] is its rank k_i; idx is the
 * DOF permutation.  Leaves are the nodes without sons.
 * near / far: 2 values per block, (row node, column node). */
int h2hip_h2matrix_create(h2hip_h2matrix **out, int is_complex,
                          uint32_t rnodes, const uint32_t *rtree,
                          const uint32_t *rrank, uint32_t nrdof,
                          const uint32_t *ridx, uint32_t cnodes,
                          const uint32_t *ctree, const uint32_t *crank,
                          uint32_t ncdof, const uint32_t *cidx,
                          uint32_t nnear, const uint32_t *near, uint32_t nfar,
                          const uint32_t *far);
void h2hip_h2matrix_destroy(h2hip_h2matrix *H);
/* Cluster basis of one side (0 row, 1 column): leaf matrices V_t
 * (size x rank, column-major) packed in node order at voff[t] (entries for
 * inner nodes ignored), transfer matrices E_t (rank_t x rank_father,
 * column-major) packed at node * eld where eld >= max rank^2. */
int h2hip_h2matrix_set_basis(h2hip_h2matrix *H, int side, const double *V,
                             const uint64_t *voff, const double *E,
                             uint64_t eld);
/* Couplings, block b (rank_r x rank_c, column-major) at S + b * sld
 * (doubles or complex pairs). */
int h2hip_h2matrix_set_couplings(h2hip_h2matrix *H, const void *S,
                                 uint64_t sld);
/* Near-field blocks packed in block order, rows x cols column-major
 * (host array), or the device storage for h2hip_nearfield_execute with a
 * plan over the same blocks. */
int h2hip_h2matrix_set_nearfield(h2hip_h2matrix *H, const void *N);
void *h2hip_h2matrix_nearfield_device(h2hip_h2matrix *H);
/* Transposed copies for the adjoint product (no counterpart in the
 * reference; keep = 1): the couplings and near blocks are also kept as
 * their conjugate transposes (twice the memory of both), made by the first
 * h2hip_addevaltrans_h2matrix after they were set, on its stream; the
 * adjoint then runs the kernels of the forward product (couplings computed
 * in the product still take the adjoint path).  Setting the couplings or
 * the near field (h2hip_h2matrix_set_couplings, _interp_couplings,
 * _set_nearfield, _nearfield_device) marks the copies stale; a caller that
 * writes the near field again through an earlier pointer calls this
 * function again before the next adjoint product.  keep = 0 frees them. */
int h2hip_h2matrix_set_adjoint_storage(h2hip_h2matrix *H, int keep);
/* Products of two to eight vectors (h2hip_addeval_h2matrix_batch) of a real
 * matrix whose clusters all have rank 64 multiply the stored couplings on
 * FP64 matrix cores (on = 1; no counterpart in the reference): 64
 * v_mfma_f64_16x16x4_f64 per coupling block.  Other products, computed
 * couplings and reproducible mode keep the vector kernels. */
int h2hip_h2matrix_set_mfma(h2hip_h2matrix *H, int on);
/* Couplings computed in the product (coupling mode 1) of an order-4 or -5
 * interpolation matrix (Laplace or Helmholtz) whose row and column boxes
 * have the same points and whose admissible blocks come in mirrored pairs
 * (r, c) / (c, r), as for the operators of one mesh: the one-vector
 * forward product evaluates each coupling value once for both blocks of a
 * pair (on = 1, the default; no counterpart in the reference).
 * Reproducible mode, adjoint and batched products keep one evaluation per
 * block. */
int h2hip_h2matrix_set_coupling_pairs(h2hip_h2matrix *H, int on);
/* The number of mirrored pairs found by h2hip_h2matrix_interp_couplings
 * (0: the structure is not symmetric, or not order 4 or 5; -1: H null). */
int h2hip_h2matrix_coupling_pairs(const h2hip_h2matrix *H);
/* Couplings of an interpolation H2-matrix computed on the device (instead
 * of h2hip_h2matrix_set_couplings): kernel_const G(xi, eta) at the tensor
 * Chebyshev points of both boxes of every admissible block, G = 1/r
 * (H2HIP_LAPLACE, real matrix) or e^{ik r}/r (H2HIP_HELMHOLTZ, complex
 * matrix, k = (re, im)), as assemble_bem3d_inter_uniform
 * (Library/bem3d.c:6079-6108) with fill_kernel_* fills them.  rpts / cpts:
 * 3 m points per node of each tree (x, y, z), ranks m^3.  Returns when
 * the couplings are written. */
/* y += alpha H x for a self-adjoint H (addevalsymm_h2matrix_avector,
 * Library/h2matrix.c:1570-1604): only the blocks on and below the block
 * diagonal are used (row cluster after column cluster; the matrix may be
 * created with just those, which halves its storage), each below the
 * diagonal also as its adjoint, diagonal near blocks with their lower
 * triangle.  Row and column trees and cluster bases must be equal (as the
 * interpolation bases of a single layer operator are). */
int h2hip_addevalsymm_h2matrix(const double *alpha, const h2hip_h2matrix *H,
                               const void *x, void *y, void *stream);
/* Schedule of the basis transforms of the products (results agree to the
 * rounding of the atomic sums): 0 one upward launch in which the last son
 * to finish computes its father (counters and fences), 1 one upward launch
 * for the leaves and the clusters whose sons are all leaves, then one
 * launch per height above them; the downward sweep walks each leaf's path
 * from the root.  2 flat transforms for the products y += alpha H x
 * (h2hip_addeval_h2matrix, h2hip_addevalsymm_h2matrix and the batched
 * h2hip_addeval_h2matrix_batch without trans): the bases
 * of all ancestors restricted to each leaf (W_a|_l = V_l E_l E_f ...) are
 * built once on the device (row-major for the column side, column-major
 * for the row side; h2hip_h2matrix_bytes reports them as the bases), and
 * both sweeps become one launch without dependence chains; the adjoint
 * products use schedule 1.  Needs the bases set and leaves that partition
 * the index positions; setting new bases rebuilds the blocks.  Not
 * thread-safe against products in flight.  No counterpart in the
 * reference. */
int h2hip_h2matrix_set_schedule(h2hip_h2matrix *H, int level_sync);
/* Couplings of the products y += alpha H x (h2hip_addeval_h2matrix and
 * h2hip_addeval_h2matrix_batch, not transposed): recompute = 0 reads the
 * stored couplings, 1 computes each entry in the product from the
 * interpolation points kept by h2hip_h2matrix_interp_couplings (the same
 * arithmetic, so the same values), which trades the product's largest read
 * (k^2 doubles per admissible block) for k^2 kernel evaluations, shared by
 * all vectors of a batched product.  Transposed and self-adjoint products
 * keep reading the stored couplings.  Needs couplings set by
 * h2hip_h2matrix_interp_couplings; h2hip_h2matrix_set_couplings switches
 * back to 0.  h2hip_h2matrix_bytes then counts no coupling bytes.  No
 * counterpart in the reference. */
int h2hip_h2matrix_set_coupling_mode(h2hip_h2matrix *H, int recompute);
/* Cluster bases of one side (0 rows, 1 columns) built on the device by
 * interpolation, in place of h2hip_h2matrix_set_basis: leaf matrices
 * V_t[i, j] = int phi_i L_j (dn != 0: int phi_i dL_j/dn, the column basis of
 * the double layer) with the single-triangle rule of order q, and transfer
 * matrices E_s[i, j] = L^father_j(xi^s_i), over the tensor Chebyshev points
 * pts (3 m per cluster: x, then y, then z; h2b_interp_points) of the
 * side's tree on the surface s (basis 0: DOFs are triangles, 1: vertices).
 * Every rank must be m^3 (m <= 5).  assemble_bem3d_h2matrix_row/
 * col_clusterbasis with setup_h2matrix_aprx_inter_bem3d
 * (Library/bem3d.c:9658-9686, 10377-10446, 5937-6005, 7555-8336).
 * Complete on return. */
int h2hip_h2matrix_interp_bases(h2hip_h2matrix *H, int side,
                                const h2hip_surface *s, int basis, int dn,
                                uint32_t m, uint32_t q, const double *pts,
                                void *stream);
int h2hip_h2matrix_interp_couplings(h2hip_h2matrix *H, int kernel,
                                    const double *k, uint32_t m,
                                    const double *rpts, const double *cpts,
                                    void *stream);
/* ---- dense matrices ----------------------------------------------------------
 * y += alpha A x (trans = 0) or y += alpha A^* x (trans != 0, conjugate
 * transpose) for a column-major device matrix A of rows x cols with leading
 * dimension ld >= rows, real or complex (is_complex: interleaved pairs),
 * device vectors x and y, alpha = (re, im); y is not cleared.  Ordered on
 * the stream; results do not depend on the launch schedule (partial sums
 * over column chunks are added in chunk order).  The partial sums of A x
 * live in stream-ordered device memory for the duration of the call.
 * Replaces addeval_amatrix_avector / addevaltrans_amatrix_avector
 * (Library/amatrix.c:862-925). */
int h2hip_addeval_amatrix(const double *alpha, int is_complex, int trans,
                          uint32_t rows, uint32_t cols, const void *A,
                          uint64_t ld, const void *x, void *y, void *stream);

/* Bytes one product y += alpha H x reads in the current schedule and
 * coupling mode: near field, couplings (0 when computed), bases (the
 * blocks W_a|_l of both sides with the flat transforms, else leaf and
 * transfer matrices). */
int h2hip_h2matrix_bytes(const h2hip_h2matrix *H, uint64_t *out3);
/* Storage of the matrix in the reference's terms and on the device:
 * out4[0] near-field entries, out4[1] couplings of all admissible blocks
 * (k_r k_c entries each, whether kept or computed in the products),
 * out4[2] leaf and transfer matrices of both sides, one copy each (the
 * storage of the reference's clusterbasis, Library/clusterbasis.h:40-71),
 * all in bytes; out4[3] bytes of every device buffer the matrix holds
 * (both orientations of the bases, flat blocks, coefficient vectors,
 * trees).  No counterpart in the reference. */
int h2hip_h2matrix_storage(const h2hip_h2matrix *H, uint64_t *out4);
/* y += alpha H x (addeval_h2matrix_avector, Library/h2matrix.c:1268-1286)
 * and y += alpha H^* x (addevaltrans_h2matrix_avector, Library/h2matrix.c:
 * 1345-1364) on device vectors in the original DOF order; alpha = (re, im).
 * Calls on the same matrix must not overlap (shared coefficient
 * workspace). */
int h2hip_addeval_h2matrix(const double *alpha, const h2hip_h2matrix *H,
                           const void *d_x, void *d_y, void *stream);
int h2hip_addevaltrans_h2matrix(const double *alpha, const h2hip_h2matrix *H,
                                const void *d_x, void *d_y, void *stream);

/* The same products for nvec (1..8) vectors at once, which read the
 * matrix once: y_v += alpha_v H x_v (trans: H^*), alpha = 2 nvec doubles
 * (re, im), d_x / d_y host arrays of nvec device pointers.  For more than
 * one vector every leaf and every rank must be at most 64.  Shares the
 * calls' coefficient workspace like the single products.  No counterpart
 * in the reference (it multiplies one vector at a time); serves batched
 * solvers. */
int h2hip_addeval_h2matrix_batch(const h2hip_h2matrix *H, uint32_t nvec,
                                 const double *alpha, const void *const *d_x,
                                 void *const *d_y, int trans, void *stream);

/* ---- Krylov solvers ----------------------------------------------------------
 * y += alpha A x on device vectors, the counterpart of the reference's
 * addeval_t (Library/krylov.h:41) with a stream; h2hip_addeval_h2matrix has
 * this form. */
typedef int (*h2hip_addeval_t)(const double *alpha, void *A, const void *x,
                               void *y, void *stream);
/* Restarted GMRES(kmax) for A x = b on device vectors (x holds the start
 * vector): stops when the least-squares residual is at most eps |b| or
 * after maxiter - 1 steps, restarts when the Krylov dimension reaches
 * kmax - 1.  solve_gmres_avector (Library/krylovsolvers.c:183-227) with
 * init/step/finish_gmres (Library/krylov.c:612-770).  FP64 inner products.
 * Returns the number of steps and the final residual estimate. */
int h2hip_solve_gmres(h2hip_addeval_t addeval, void *A, int is_complex,
                      uint64_t n, const void *d_b, void *d_x, double eps,
                      uint32_t maxiter, uint32_t kmax, void *stream,
                      uint32_t *iterations, double *residual);
/* Conjugate gradients for A x = b, A self-adjoint positive definite, on
 * device vectors (x holds the start vector): stops when |r| <= eps |b| or
 * after maxiter - 1 steps (maxiter 0: no limit).  solve_cg_avector
 * (Library/krylovsolvers.c:18-52) with init_cg / step_cg
 * (Library/krylov.c:162-196), <u, v> = sum conj(u_i) v_i.  The stopping
 * test runs on the device, the host looks at it every 8 steps;
 * use_graph != 0 replays the 8 steps as one captured graph (addeval must
 * then be capturable, as h2hip_addeval_h2matrix is).  Returns the steps
 * and the final |r|; the solve is complete on return. */
int h2hip_solve_cg(h2hip_addeval_t addeval, void *A, int is_complex,
                   uint64_t n, const void *d_b, void *d_x, double eps,
                   uint32_t maxiter, int use_graph, void *stream,
                   uint32_t *iterations, double *residual);
/* h2hip_solve_cg without allocations.  No counterpart in the reference: for
 * callers that cannot allocate (serving loops that must not call hipMalloc,
 * harnesses that forbid it).  The work vectors and the device state live in
 * a fixed device array of the library (24 MiB per GPU, held until the
 * process ends), so the call allocates no device or pinned memory and
 * creates no events, streams or graphs; systems of more than 1048576 real
 * or 524288 complex unknowns are refused with H2HIP_EINVAL.  addeval must
 * itself allocate nothing (an H2-matrix product allocates its buffers on
 * its first call in a mode: run one product first).  Same kernels, steps
 * and stopping rule as h2hip_solve_cg without use_graph.  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU race. */
int h2hip_solve_cg_noalloc(h2hip_addeval_t addeval, void *A, int is_complex,
                           uint64_t n, const void *d_b, void *d_x,
                           double eps, uint32_t maxiter, void *stream,
                           uint32_t *iterations, double *residual);

/* Helmholtz point-source data on the device (real wave number k, source
 * x0, Dirichlet data e^{ikr}/r for kind 0, Neumann data for kind 1,
 * rhs_dirichlet_point_helmholtzbem3d / rhs_neumann_point_helmholtzbem3d,
 * Library/helmholtzbem3d.c:1521-1578) with the single-triangle rule of
 * order q (1..8), complex output d_out on the device, ordered on the
 * stream without synchronising:
 *   mode 0: L2 projection onto the constants, one value per triangle
 *           (projectL2_bem3d_c_avector, Library/bem3d.c:8827-8876);
 *   mode 1: integrals against the constants, one per triangle
 *           (integrate_bem3d_c_avector, Library/bem3d.c:8579-8629);
 *   mode 2: integrals against the linear basis, one per vertex
 *           (integrate_bem3d_l_avector, Library/bem3d.c:8631-8702).
 * accumulate != 0 adds to d_out (data of several sources). */
int h2hip_helmholtz_point_data(const h2hip_surface *s, uint32_t q, int kind,
                               int mode, double k, const double *x0,
                               int accumulate, void *d_out, void *stream);
/* out[0] = |x - data|_L2 and out[1] = |data|_L2 for the constant-basis
 * device vector d_x (normL2diff_c_bem3d, Library/bem3d.c:8776-8825;
 * normL2_bem3d, Library/bem3d.c:8704-8751).  Returns when out is
 * written. */
int h2hip_helmholtz_point_errors_c(const h2hip_surface *s, uint32_t q,
                                   int kind, double k, const double *x0,
                                   const void *d_x, double *out,
                                   void *stream);

/* ---- operator combinations and the exterior potential ---------------------
 * A = c1 A1 + c2 A2 as one operator for the Krylov solvers: the
 * Brakhage-Werner operator KM - i eta V of the reference's Helmholtz test
 * (addeval_A, Tests/test_helmholtzbem3d.c:24-65) is
 * h2hip_combined_create(&C, h2hip_addeval_h2matrix, KM, {1, 0},
 * h2hip_addeval_h2matrix, V, {0, -eta}).  f2 may be NULL (A = c1 A1).
 * c1, c2 complex as (re, im). */
typedef struct h2hip_combined h2hip_combined;
int h2hip_combined_create(h2hip_combined **out, h2hip_addeval_t f1,
                          void *A1, const double *c1, h2hip_addeval_t f2,
                          void *A2, const double *c2);
void h2hip_combined_destroy(h2hip_combined *C);
/* y += alpha (c1 A1 + c2 A2) x; signature of h2hip_addeval_t. */
int h2hip_addeval_combined(const double *alpha, void *A, const void *x,
                           void *y, void *stream);
/* The Brakhage-Werner potential of the constant-basis density d_w (complex,
 * device) at npts points (host array, 3 doubles each) into out (host,
 * complex): kernel_const sum_t g_t w_t sum_q w_q e^{ikr}/r^3 ((1 - ikr)
 * <x - y_q, n_t> - i eta r^2) with the single-triangle rule of order q
 * (eval_brakhage_werner_c, Tests/test_helmholtzbem3d.c:67-132).  k and eta
 * complex as (re, im).  Returns when out is written. */
int h2hip_eval_brakhage_werner_c(const h2hip_surface *s, uint32_t q,
                                 const double *k, const double *eta,
                                 const void *d_w, uint32_t npts,
                                 const double *pts, void *out, void *stream);
/* The same for a piecewise linear density d_w (one complex value per
 * vertex): kernel_const sum_t g_t sum_i w_{t_i} sum_q w_q phi_i(y_q) ...
 * (eval_brakhage_werner_l, Tests/test_helmholtzbem3d.c:133-221, which
 * takes k real: only k[0] is used). */
int h2hip_eval_brakhage_werner_l(const h2hip_surface *s, uint32_t q,
                                 const double *k, const double *eta,
                                 const void *d_w, uint32_t npts,
                                 const double *pts, void *out, void *stream);

/* ---- boundary data ----------------------------------------------------------
 * L2 projection of the Dirichlet (kind 0: 1/|x - x0|) or Neumann (kind 1:
 * -<x - x0, n>/|x - x0|^3) data of a point source at x0 onto the piecewise
 * linear basis of the surface, written to the device vector d_out (one
 * entry per vertex): the load vector with the single-triangle Gauss rule of
 * order q (1..8) and conjugate gradients on the mass matrix to
 * |r| <= 1e-15 |b| (at most 500 steps), all on the device, ordered on the
 * stream; the stopping test runs on the device and the host reads its flag
 * every 8 steps, so the call returns when d_out is complete.
 * projectL2_bem3d_l_avector (Library/bem3d.c:9009-9088) with
 * eval_dirichlet_fundamental_laplacebem3d /
 * eval_neumann_fundamental_laplacebem3d (Library/laplacebem3d.c:1390-1430).
 * Calls on one surface share a workspace and are serialised by a lock. */
int h2hip_projectL2_point_source(const h2hip_surface *s, uint32_t q, int kind,
                                 const double *x0, double *d_out,
                                 void *stream);
/* The load vector alone: d_out[i] = int phi_i f (accumulate != 0: added
 * to d_out) for the same data f and rule, one entry per vertex, ordered on
 * the stream; needs no workspace, so calls on one surface may overlap.
 * integrate_bem3d_l_avector (Library/bem3d.c:8631-8702) with
 * eval_dirichlet_fundamental_laplacebem3d /
 * eval_neumann_fundamental_laplacebem3d (Library/laplacebem3d.c:1390-1430):
 * the right-hand side of the exterior Neumann problem. */
int h2hip_integrate_point_source(const h2hip_surface *s, uint32_t q,
                                 int kind, const double *x0, int accumulate,
                                 double *d_out, void *stream);

/* ---- batched GMRES (serving) --------------------------------------------------
 * Up to 8 independent GMRES(kmax) solves with one H2-matrix advance
 * together: every step is one h2hip_addeval_h2matrix_batch product for all
 * of them followed by the per-solve Gram-Schmidt and rotation kernels, each
 * launched once for all solves.  Each solve has the semantics of
 * h2hip_solve_gmres; solves enter free slots (submit) and leave when they
 * finish (run reports them), so requests arriving at different times share
 * the products (continuous batching).  No counterpart in the reference; for
 * callers that serve many right-hand sides.  The matrix must satisfy the
 * batched product's limits (leaves and ranks up to 64). */
typedef struct h2hip_gmres_batch h2hip_gmres_batch;
int h2hip_gmres_batch_create(h2hip_gmres_batch **out, const h2hip_h2matrix *H,
                             int is_complex, uint64_t n, uint32_t slots,
                             uint32_t kmax, void *stream);
void h2hip_gmres_batch_destroy(h2hip_gmres_batch *G);
/* The batch's operator becomes c1 H + c2 H2 (complex coefficients as
 * (re, im); H2 of the same size and batched-product limits), e.g. the
 * Brakhage-Werner operator KM - i eta V of the reference's Helmholtz test
 * (addeval_A, Tests/test_helmholtzbem3d.c:24-65): every step makes two
 * batched products.  Only while no solve is in progress and with GMRES as
 * the method (the combination is not Hermitian; H2HIP_EINVAL otherwise). */
int h2hip_gmres_batch_combine(h2hip_gmres_batch *G, const double *c1,
                              const h2hip_h2matrix *H2, const double *c2);
/* The batch's solves use method 0, GMRES(kmax) (the default), or 1,
 * conjugate gradients with the semantics of h2hip_solve_cg
 * (solve_cg_avector, Library/krylovsolvers.c:18-52, the reference Laplace
 * test's solver; for Hermitian positive definite operators; kmax >= 3, as
 * CG keeps its vectors in the first three Krylov columns).  With CG a chunk
 * is always 8 steps and the device's stopping flag freezes each solve at
 * its stopping step.  Only while no solve is in progress; CG is refused
 * once h2hip_gmres_batch_combine has set a combination. */
int h2hip_gmres_batch_set_method(h2hip_gmres_batch *G, int method);
/* Number of free slots. */
int h2hip_gmres_batch_free_slots(const h2hip_gmres_batch *G);
/* Starts a solve of A x = b (device vectors; x holds the start vector and
 * receives the solution); returns the slot (>= 0) or an error code. */
int h2hip_gmres_batch_submit(h2hip_gmres_batch *G, const void *d_b,
                             void *d_x, double eps, uint32_t maxiter,
                             uint64_t tag);
/* Advances every active solve by one chunk of steps (at most 8, never past
 * a restart) and reports the solves that finished: their tags, steps and
 * residual estimates (up to cap of them; *ndone counts all).  Solutions of
 * finished solves are complete on return. */
int h2hip_gmres_batch_run(h2hip_gmres_batch *G, uint32_t *ndone,
                          uint64_t *tags, uint32_t *iterations,
                          double *residuals, uint32_t cap);
/* The same chunk in two halves, so that solves with different matrices
 * (each batch on its own stream) overlap on the device: enqueue launches
 * the chunk and returns at once (nothing happens if a chunk is in flight or
 * no solve is active); poll returns 1 while the chunk runs and otherwise
 * completes it like run (0, with *ndone and the finished solves).  submit
 * is refused while a chunk is in flight.  busy: bit 0 some solve active,
 * bit 1 a chunk in flight.  No counterpart in the reference. */
int h2hip_gmres_batch_enqueue(h2hip_gmres_batch *G);
int h2hip_gmres_batch_poll(h2hip_gmres_batch *G, uint32_t *ndone,
                           uint64_t *tags, uint32_t *iterations,
                           double *residuals, uint32_t cap);
int h2hip_gmres_batch_busy(const h2hip_gmres_batch *G);

/* ---- many independent systems, one graph launch per Arnoldi step ----------
 * Solves A_s x_s = b_s for count systems (typically one H2-matrix per mesh)
 * with the semantics of solve_gmres_avector (Library/krylovsolvers.c:
 * 183-227) for each of them, all with the same eps, maxiter and kmax.  The
 * systems step in lock step; one step of every active system is captured
 * once into a HIP graph (one branch per distinct matrix) and replayed for
 * each step (use_graph = 1), or launched directly (use_graph = 0; the
 * same kernels: results agree to the rounding of the product's atomic
 * sums).  The graph is captured again only
 * when a system leaves.  stream must not be the default stream (it is
 * captured).  x_s holds the start vector on entry.  No counterpart in the
 * reference (it solves one system at a time). */
typedef struct h2hip_gmres_multi h2hip_gmres_multi;
int h2hip_gmres_multi_create(h2hip_gmres_multi **out, uint32_t count,
                             const h2hip_h2matrix *const *H,
                             const uint64_t *n, int is_complex, uint32_t kmax,
                             void *stream);
void h2hip_gmres_multi_destroy(h2hip_gmres_multi *G);
int h2hip_gmres_multi_solve(h2hip_gmres_multi *G, const void *const *d_b,
                            void *const *d_x, double eps, uint32_t maxiter,
                            int use_graph, uint32_t *iterations,
                            double *residuals);
/* Number of graphs captured so far (one per change of the active set). */
uint64_t h2hip_gmres_multi_graph_builds(const h2hip_gmres_multi *G);

/* Waits for all work queued on the stream. */
int h2hip_stream_sync(void *stream);

/* ---- device memory, streams and events -----------------------------------
 * Thin wrappers over the HIP runtime for callers without their own device
 * allocator (tests, benchmarks, host programs).  No counterpart in the
 * reference, which keeps all data in host memory (allocmem,
 * Library/basic.h:512).  kind: 0 host to device, 1 device to host,
 * 2 device to device. */
int h2hip_malloc(void **ptr, size_t bytes);
int h2hip_free(void *ptr);
int h2hip_memcpy(void *dst, const void *src, size_t bytes, int kind);
int h2hip_memset(void *ptr, int value, size_t bytes);
/* Fill ordered on a stream (not for code under graph capture: a captured
 * fill node is not replayed reliably; use a kernel there). */
int h2hip_memset_async(void *ptr, int value, size_t bytes, void *stream);
int h2hip_stream_create(void **stream);
int h2hip_stream_destroy(void *stream);
int h2hip_event_create(void **event);
int h2hip_event_destroy(void *event);
int h2hip_event_record(void *event, void *stream);
/* Milliseconds between two recorded events (waits for the second). */
int h2hip_event_elapsed(void *start, void *stop, float *ms);

#ifdef __cplusplus
}
#endif

#endif
