/* H2Lib-side binding of the HIP near field and H2-matrix product.
 *
 * This file and h2lib_hip.c are added to an H2Lib build (they include
 * H2Lib's headers from Library/ and link against libh2hip.so).  They follow
 * the pattern of H2Lib's own accelerator constructors
 * (new_slp_laplace_ocl_bem3d, Library/laplaceoclbem3d.c:764-783): build the
 * operator with the CPU constructor, then overwrite bem->nearfield and
 * bem->nearfield_far (Library/bem3d.h:380-381, 405-406).  Every caller of
 * the callbacks (assemble_bem3d_amatrix, assemble_bem3d_h2matrix, the
 * Green / ACA approximations) then computes its near field on the GPU
 * without further changes.
 *
 * The H2-matrix product is offered as a packed device copy of an assembled
 * h2matrix (Library/h2matrix.h:48-71) with the semantics of
 * addeval_h2matrix_avector / addevaltrans_h2matrix_avector
 * (Library/h2matrix.c:1268-1364) on host avectors, callbacks of the types
 * addeval_t and mvm_t (Library/krylov.h:41, 60) for H2Lib's Krylov solvers,
 * and a device-resident GMRES(kmax) with solve_gmres_avector's semantics
 * (Library/krylovsolvers.c:183-227).
 *
 * Errors: as in H2Lib, a failed call prints the library's message and
 * aborts.
 */
#ifndef H2LIB_HIP_H
#define H2LIB_HIP_H

#include "bem3d.h"
#include "h2matrix.h"
#include "krylov.h"

/* Installs the GPU near field into bem (built by one of the reference
 * constructors for the given kernel: helmholtz 0/1, layer 0 single, 1
 * double, 2 adjoint double layer).  The surface and the quadrature orders
 * are taken from bem; alpha and k likewise. */
void setup_hip_bem3d(pbem3d bem, int helmholtz, int layer);

/* The same with deferred callbacks: bem->nearfield and bem->nearfield_far
 * only record their blocks (N is not written yet), and flush_hip_bem3d
 * computes all recorded blocks in one plan and writes them.  Callers flush
 * after the assembly and before they read the near-field blocks, e.g.
 *   assemble_bem3d_h2matrix(bem, h2); flush_hip_bem3d(bem);
 * the deferred completion of the reference's OpenCL callbacks
 * (Library/laplaceoclbem3d.c:248-279, Library/opencl.h:359-367). */
void setup_hip_deferred_bem3d(pbem3d bem, int helmholtz, int layer);
void flush_hip_bem3d(pcbem3d bem);

/* Releases the device data of bem (call before del_bem3d). */
void release_hip_bem3d(pcbem3d bem);

/* Precision of the GPU near field of bem (set up by setup_hip_bem3d):
 * 0 FP64 (default), 1 mixed (h2hip_op_set_precision).  No counterpart in
 * the reference. */
void set_precision_hip_bem3d(pcbem3d bem, int mode);

pbem3d new_slp_laplace_hip_bem3d(pcsurface3d gr, uint q_regular,
                                 uint q_singular,
                                 basisfunctionbem3d row_basis,
                                 basisfunctionbem3d col_basis);
pbem3d new_dlp_laplace_hip_bem3d(pcsurface3d gr, uint q_regular,
                                 uint q_singular,
                                 basisfunctionbem3d row_basis,
                                 basisfunctionbem3d col_basis, field alpha);
pbem3d new_adlp_laplace_hip_bem3d(pcsurface3d gr, uint q_regular,
                                  uint q_singular,
                                  basisfunctionbem3d row_basis,
                                  basisfunctionbem3d col_basis, field alpha);
#ifdef USE_COMPLEX
pbem3d new_slp_helmholtz_hip_bem3d(field k, pcsurface3d gr, uint q_regular,
                                   uint q_singular,
                                   basisfunctionbem3d row_basis,
                                   basisfunctionbem3d col_basis);
pbem3d new_dlp_helmholtz_hip_bem3d(field k, pcsurface3d gr, uint q_regular,
                                   uint q_singular,
                                   basisfunctionbem3d row_basis,
                                   basisfunctionbem3d col_basis,
                                   field alpha);
pbem3d new_adlp_helmholtz_hip_bem3d(field k, pcsurface3d gr,
                                    uint q_regular, uint q_singular,
                                    basisfunctionbem3d row_basis,
                                    basisfunctionbem3d col_basis,
                                    field alpha);
#endif

/* Device copy of an assembled H2-matrix (cluster bases, couplings, near
 * field).  The h2matrix may be deleted afterwards. */
typedef struct _hiph2matrix hiph2matrix;
typedef hiph2matrix *phiph2matrix;

phiph2matrix new_hip_h2matrix(pch2matrix h2);
void del_hip_h2matrix(phiph2matrix A);

/* y += alpha A x and y += alpha A^* x with host vectors. */
void addeval_hip_h2matrix_avector(field alpha, phiph2matrix A, pcavector x,
                                  pavector y);
void addevaltrans_hip_h2matrix_avector(field alpha, phiph2matrix A,
                                       pcavector x, pavector y);

/* addeval_t and mvm_t callbacks (matrix = phiph2matrix). */
void addeval_hip_h2matrix_callback(field alpha, void *A, pcavector x,
                                   pavector y);
void mvm_hip_h2matrix_callback(field alpha, bool trans, void *A, pcavector x,
                               pavector y);

/* GMRES(kmax) for A x = b with every vector on the device; x holds the
 * start vector.  Returns the number of steps. */
uint solve_gmres_hip_h2matrix_avector(phiph2matrix A, pcavector b, pavector x,
                                      real eps, uint maxiter, uint kmax);

#endif
