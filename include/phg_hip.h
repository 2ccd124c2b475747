/* phg_hip.h -- C ABI of the MI355X (gfx950) finite-element hot path.
 *
 * The library libphg_hip.so implements, as hand-written HIP kernels, the part
 * of PHG that the element loop and the Krylov solve run through:
 *   element integration  quad.c:5260 phgQuadGradBasAGradBas,
 *                        quad.c:5582 phgQuadBasABas,
 *                        quad.c:5481 phgQuadDofTimesBas,
 *                        geom.c:103  compute_jacobian,
 *   row insertion        matvec.c:5348 phgMatAddEntries_ / :3045
 *                        add_columns_to_row (insertion order, exact zeros
 *                        dropped), matvec.c:5661 phgVecAddEntries_,
 *   Dirichlet handling   matvec.c:4430 save_bdry_columns, solver.c:508
 *                        update_rhs,
 *   Krylov               matvec.c:552 phgMatVec, matvec.c:1193 phgVecDot,
 *                        matvec.c:51 phgVecNorm2, matvec.c:1237 phgVecAXPBY,
 *                        solver-phg.c:485 solver_pcg with
 *                        solver-phg.c:207 jacobi_pc.
 * PHG's own element loop (examples/poisson.c:101-170) calls the quad
 * functions once per basis pair; here the loop is lowered to one batched
 * call (phg_hip_assemble) over all elements, which produces the same CSR
 * storage order and the same values as PHG's row lists.
 *
 * Conventions: every function returns PHG_HIP_OK (0) or a negative error
 * code; phg_hip_last_error() gives the message.  Host pointers are plain
 * arrays; the library owns all device memory behind the opaque handles.
 * All work is ordered on one HIP stream per process (phg_hip_init).
 */
#ifndef PHG_HIP_H
#define PHG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHG_HIP_OK              0
#define PHG_HIP_E_HIP         (-1)   /* a HIP runtime call failed */
#define PHG_HIP_E_ARG         (-2)   /* invalid argument / size */
#define PHG_HIP_E_NOMEM       (-3)   /* device allocation failed */
#define PHG_HIP_E_ZERO_DIAG   (-4)   /* solver-phg.c:233 zero diagonal */
#define PHG_HIP_E_STATE       (-5)   /* library not initialised */

typedef struct phg_hip_space phg_hip_space;   /* mesh + FE space on device */
typedef struct phg_hip_mat   phg_hip_mat;     /* assembled CSR system     */
typedef struct phg_hip_vec   phg_hip_vec;     /* FP64 device vector       */

/* Problem selector for the element integrals (the integrand a driver such
 * as examples/poisson.c, simplest.c or elastic.c builds). */
typedef struct {
    int32_t dim;          /* 1: scalar (poisson.c), 3: elasticity (elastic.c) */
    int32_t use_kappa;    /* scalar: phgQuadGradBasAGradBas with a P0 kappa */
    double  mass_coef;    /* scalar: + mass_coef * phgQuadBasDotBas (simplest.c:70) */
    int32_t rhs_order;    /* quadrature order of the load, <0: 2*order */
    double  E, nu;        /* elasticity moduli (elastic.c:39-40) */
    int32_t qcache;       /* scalar Laplacian (no kappa, no mass): the integrals
                             of examples/poisson.c -qcache (quad-cache.c:1016
                             phgQCIntegrate_, one rule of order 2p) */
} phg_hip_problem;

/*--------------------------------------------------------------- runtime */

/* Replaces the GPU side of utils.c:555 phgInit: selects the device and
 * creates the library stream. */
int phg_hip_init(int32_t device);
/* utils.c:681 phgFinalize counterpart: releases the stream. */
int phg_hip_finalize(void);
const char *phg_hip_last_error(void);
int phg_hip_synchronize(void);

/*---------------------------------------------------------- mesh / space */

/* Uploads the mesh and the Lagrange Pk space (order 1..4):
 * verts[nvert][3] (utils.c:5470 phgImport_ result), elem_verts[nelem][4],
 * map[nelem][nbas] = dof.c:4821 phgDofMapE2D_ of every local basis,
 * bdry[ndof] = 1 for Dirichlet DOFs (dof-utils.c:1269 nodal path).
 * Everything lives on the device; the caller's arrays are not kept. */
int phg_hip_space_create(int32_t order, int64_t nvert, const double *verts,
                         int64_t nelem, const int32_t *elem_verts,
                         const int32_t *map, int64_t ndof, const uint8_t *bdry,
                         phg_hip_space **sp);
int phg_hip_space_destroy(phg_hip_space *sp);

/*--------------------------------------------------------------- assembly */

/* Element matrices only, copied to the host: K[nelem][M][M], b[nelem][M]
 * with M = nbas*dim, exactly as the drivers hold them before rows are
 * inserted (replaces the per-pair calls of quad.c:5260 / :5582 / :5481 in
 * examples/poisson.c:124-150 and elastic.c:160-196). kappa[nelem] may be
 * NULL; f[ndof*dim] is the load DOF (nodal interpolant). */
int phg_hip_element_matrices(phg_hip_space *sp, const phg_hip_problem *pb,
                             const double *kappa, const double *f,
                             double *K, double *b);

/* Full assembly: element integrals, PHG row insertion (matvec.c:5348),
 * Dirichlet rows e_i with right-hand side g (dof-utils.c:1369) and RHS
 * accumulation (matvec.c:5661).  g[ndof*dim] holds the Dirichlet values.
 * The result is the matrix before phgMatAssemble's boundary treatment.
 * The space keeps the scratch of its last assembly for the next one (the
 * element matrices, the row incidence and the row segments of the
 * insertion: about 12 M^2 + 8 M bytes per element, M = nbas * dim; 3.3 GB
 * for P2 on 1.57M elements) until phg_hip_space_destroy. */
int phg_hip_assemble(phg_hip_space *sp, const phg_hip_problem *pb,
                     const double *kappa, const double *f, const double *g,
                     phg_hip_mat **A);

/* phgSolverAssemble + phgSolverUpdateRHS on one process
 * (matvec.c:4602 phgMatAssemble -> :4430 save_bdry_columns(TRUE),
 * solver.c:734 -> :508 update_rhs). */
int phg_hip_mat_eliminate(phg_hip_mat *A);

/* A system built elsewhere (e.g. PHG's own packed matrix, matvec.c:279
 * phgMatPack: packed_ind / packed_cols / packed_data) uploaded as CSR:
 * rowptr[n+1] (0-based), cols[nnz], vals[nnz], rhs[n], bdry[n] (1 for
 * Dirichlet rows).  eliminated = 0: the matrix before phgMatAssemble's
 * boundary treatment (call phg_hip_mat_eliminate next); eliminated = 1: the
 * matrix after it (boundary columns already moved, Dirichlet rows e_i),
 * ready for the solvers (diagonal as matvec.c:500 phgMatSetupDiagonal).
 * No node coordinates are given, so the solvers keep the given numbering. */
int phg_hip_mat_from_csr(int64_t n, const int64_t *rowptr, const int32_t *cols,
                         const double *vals, const double *rhs,
                         const uint8_t *bdry, int32_t eliminated,
                         phg_hip_mat **A);

int phg_hip_mat_info(const phg_hip_mat *A, int64_t *nrows, int64_t *nnz);
/* matvec.c:6134 phgMatDumpCSR equivalent: copies the CSR (0-based, row
 * pointers widened to int64) and the right-hand side to the host. */
int phg_hip_mat_download(const phg_hip_mat *A, int64_t *rowptr, int32_t *cols,
                         double *vals, double *rhs);
int phg_hip_mat_destroy(phg_hip_mat *A);
/* matvec.c:6134 phgMatDumpCSR: writes fn as the size, the n+1 row pointers
 * (0-based), the column indices and the values ("%0.17lg"), one number per
 * line, rows in storage order; eliminates A first when it was not (as
 * PHG assembles first).  fn == NULL does nothing. */
int phg_hip_mat_dump_csr(phg_hip_mat *A, const char *fn);
/* matvec.c:6245 phgVecDump: the size, then the values ("%0.17lg"). */
int phg_hip_vec_dump(const phg_hip_vec *v, const char *fn);

/*---------------------------------------------------------------- vectors */

/* matvec.c:2690 phgMapCreateVec counterpart (one FP64 vector of n). */
int phg_hip_vec_create(int64_t n, phg_hip_vec **v);
int phg_hip_vec_destroy(phg_hip_vec *v);
int phg_hip_vec_upload(phg_hip_vec *v, const double *host);
int phg_hip_vec_download(const phg_hip_vec *v, double *host);
/* The assembled right-hand side of A as a vector (copy). */
int phg_hip_mat_rhs(const phg_hip_mat *A, phg_hip_vec *b);

/* matvec.c:552 phgMatVec(MAT_OP_N, alpha, A, x, beta, &y). */
int phg_hip_matvec(double alpha, const phg_hip_mat *A, const phg_hip_vec *x,
                   double beta, phg_hip_vec *y);
/* matvec.c:1193 phgVecDot, :51 phgVecNorm2, :1237 phgVecAXPBY, :157 phgVecCopy */
int phg_hip_vec_dot(const phg_hip_vec *x, const phg_hip_vec *y, double *res);
int phg_hip_vec_norm2(const phg_hip_vec *x, double *res);
int phg_hip_vec_axpby(double a, const phg_hip_vec *x, double b, phg_hip_vec *y);
int phg_hip_vec_copy(const phg_hip_vec *src, phg_hip_vec *dst);
/* solver-phg.c:207 jacobi_pc: x = b ./ diag(A) for an eliminated A (the
 * diagonal of matvec.c:500 phgMatSetupDiagonal); PHG_HIP_E_ZERO_DIAG if a
 * diagonal entry is zero (solver-phg.c:233 stops there). */
int phg_hip_pc_jacobi(const phg_hip_mat *A, const phg_hip_vec *b, phg_hip_vec *x);

/*---------------------------------------------------------------- solvers */

/* solver-phg.c:2430 PCG_Solve / :485 solver_pcg with the Jacobi PC of
 * solver-phg.c:207, FP64 throughout: x is the initial guess on input and
 * the solution on output, b the right-hand side.  Stops when
 * |r| <= max(rtol*|r0|, btol*|b|, atol) or after maxit iterations.
 * pAp = 0 or rz = 0 ends the iterations (solver-phg.c:547-570, a warning
 * there, not an error).  |b| = 0 gives x = 0 and 0 iterations and leaves
 * *residual as the caller set it (solver-phg.c:497-501; PCG_Solve starts
 * from 1.0).  A zero diagonal entry gives PHG_HIP_E_ZERO_DIAG.
 * Returns the iteration count (>= 0) or an error code. */
int phg_hip_pcg(const phg_hip_mat *A, const phg_hip_vec *b, phg_hip_vec *x,
                int32_t maxit, double rtol, double btol, double atol,
                double *residual);

/* solver-phg.c:2457 BiCGstab_Solve / :607 solver_BiCGstab with the Jacobi
 * PC (pcg_pc_type default), FP64, for matrices that need not be
 * symmetric: same arguments and stopping rule as phg_hip_pcg.  The
 * |s| <= 1e-60 exit returns the residual of the updated x, as the
 * reference does; b = 0 gives x = 0, 0 iterations and residual 1 (the
 * reference leaves its caller's initial value).  Returns the iteration
 * count or an error code. */
int phg_hip_bicgstab(const phg_hip_mat *A, const phg_hip_vec *b,
                     phg_hip_vec *x, int32_t maxit, double rtol, double btol,
                     double atol, double *residual);

/* solver-phg.c:2484 GMRES_Solve with PHG's default -gmres_type flgmres_r
 * (solver-phg.c:2063: right preconditioned, flexible, the search space
 * augmented by the corrections of the last augk cycles) and the Jacobi
 * PC.  restart and augk are -gmres_restart (PHG default 10) and
 * -lgmres_augk (default 2); restart <= 0 means maxit + 1 and augk <= 0
 * means 2 as in the reference, and restart + augk must stay below 4096 (augk at most 64).
 * Stopping rule as phg_hip_pcg on the Givens residual estimate.  As in the
 * reference the residual reported is |r0| unless the tolerance was
 * reached, and b = 0 gives x = 0 and residual 1.  Returns the number of
 * Arnoldi steps or an error code. */
int phg_hip_gmres(const phg_hip_mat *A, const phg_hip_vec *b, phg_hip_vec *x,
                  int32_t restart, int32_t augk, int32_t maxit, double rtol,
                  double btol, double atol, double *residual);

/* -gmres_type values, in the order of solver-phg.c:71-77 gmres_names */
enum {
    PHG_HIP_GMRES_L = 0,    /* solver-phg.c:783  left preconditioned */
    PHG_HIP_GMRES_R = 1,    /* solver-phg.c:1051 right preconditioned */
    PHG_HIP_FGMRES_R = 2,   /* solver-phg.c:1263 right, flexible */
    PHG_HIP_LGMRES_L = 3,   /* solver-phg.c:1481 left, augmented */
    PHG_HIP_LGMRES_R = 4,   /* solver-phg.c:1801 right, augmented */
    PHG_HIP_FLGMRES_R = 5   /* solver-phg.c:2063 right, flexible, augmented */
};

/* solver-phg.c:2484 GMRES_Solve with any -gmres_type (gmres_funcs,
 * solver-phg.c:74) and the Jacobi PC; phg_hip_gmres is gmres_type
 * PHG_HIP_FLGMRES_R.  The plain variants ignore augk.  Left variants work
 * on z = M^-1 (b - A x): when the Givens estimate |z| meets the tolerance
 * (taken from the unpreconditioned |r0|, |b|) they compute b - A x; if that
 * is still above it they restart and run the next cycle until the estimate
 * has dropped by the ratio of the two (solver-phg.c:904-908, :979-1003),
 * and they report the last true residual.  Right variants report the
 * Givens estimate.  Both report |r0| when no check was made.  A zero
 * diagonal is PHG_HIP_E_ZERO_DIAG where jacobi_pc would first run: at once
 * for the left variants, in the first cycle for gmres_r and lgmres_r, at
 * the first Arnoldi step for the flexible ones. */
int phg_hip_gmres_type(const phg_hip_mat *A, const phg_hip_vec *b,
                       phg_hip_vec *x, int32_t gmres_type, int32_t restart,
                       int32_t augk, int32_t maxit, double rtol, double btol,
                       double atol, double *residual);

/* Mixed precision (no reference counterpart): FP64 iterative refinement
 * r = b - A x in FP64, correction from Jacobi-PCG in FP32 arithmetic on an
 * FP32 copy of A with FP64 dot products, x += d in FP64.  Stops on the
 * FP64 residual with the same tolerance rule as phg_hip_pcg.  inner_rtol is
 * the relative reduction asked of each FP32 solve.  *inner_its receives the
 * total FP32 iterations.  Returns the number of refinement steps. */
int phg_hip_pcg_mixed(const phg_hip_mat *A, const phg_hip_vec *b,
                      phg_hip_vec *x, int32_t maxit, double rtol, double btol,
                      double atol, double inner_rtol, int32_t *inner_its,
                      double *residual);

/* Mixed precision with reliable updates (no reference counterpart): one
 * FP32 Jacobi-PCG recurrence (FP32 SpMV on the FP32 copy of A, FP32
 * vectors, FP64 dot products) whose FP32 increment is folded into x (FP64)
 * and whose residual is replaced by the FP64 residual b - A x whenever the
 * FP32 residual has fallen below delta times its largest value since the
 * previous replacement, or below the tolerance.  The search direction is
 * kept across replacements.  Stops on an FP64 residual with the tolerance
 * rule of phg_hip_pcg; maxit bounds the FP32 iterations.  *updates receives
 * the number of FP64 residuals.  Returns the FP32 iteration count. */
int phg_hip_pcg_reliable(const phg_hip_mat *A, const phg_hip_vec *b,
                         phg_hip_vec *x, int32_t maxit, double rtol,
                         double btol, double atol, double delta,
                         int32_t *updates, double *residual);

/*------------------------------------------------------- many matrices */

/* Batched solve of several independent systems A[i] x[i] = b[i] (no
 * reference counterpart), each with the mixed-precision PCG of
 * phg_hip_pcg_reliable (its own layout, tolerance and iteration count).
 * The systems run concurrently on a pool of streams; with use_graph = 1 a
 * fixed chunk of iterations of every system is captured once into a HIP
 * graph (fork / join over the streams) and the graph is replayed until
 * every system has converged, so each replay is one launch for the whole
 * batch.  nmat <= 16; a matrix may appear once.  nits[i], residual[i]
 * (optional) per system.  Returns the largest iteration count. */
int phg_hip_pcg_multi(int32_t nmat, phg_hip_mat *const *A,
                      const phg_hip_vec *const *b, phg_hip_vec *const *x,
                      int32_t maxit, double rtol, double btol, double atol,
                      double delta, int32_t use_graph, int32_t *nits,
                      double *residual);

/* Re-assembly and solve of several problems in captured HIP graphs (no
 * reference counterpart; the per-system steps are phg_hip_assemble +
 * phg_hip_mat_eliminate on a fixed pattern, then phg_hip_pcg_reliable).
 * System i is the eliminated matrix A[i] that phg_hip_assemble made on the
 * space sp[i] for the problem pb[i].  kappa[i] (nelem entries; ignored
 * unless pb[i]->use_kappa), f[i] (ndof * dim) and g[i] (n) are device
 * vectors with the coefficients, the load and the Dirichlet values as
 * phg_hip_assemble takes them; the batch keeps their addresses, and every
 * phg_hip_batch_run assembles A[i] again from their current contents (the
 * element kernel, PHG's row insertion and elimination into A[i]'s arrays,
 * its solver copy, FP32 copies and layouts, and the element operator's
 * coefficients when A[i] has one), then solves A[i] x[i] = rhs(A[i]) with
 * the reliable-update PCG from the current x[i] (the initial guess; zero it
 * for a cold start).  The pattern must not change: coefficients that make
 * element-matrix entries exactly zero where they were not (or the reverse)
 * make phg_hip_batch_run return PHG_HIP_E_ARG, and A[i] must then be
 * assembled again.  create runs one full assembly pass with the vectors'
 * current contents to check the pattern against A[i], and captures two
 * graphs: refill + solver start + 16 iterations of every system (a stream
 * per system), and 16 further iterations.  run launches the first and then
 * the second until every system has converged on the device, with two
 * launches in flight: the host reads each launch's solver states from pinned
 * memory after it ends, while the next runs, and never waits on an idle
 * GPU.  Returns the largest iteration count; nits, residual per system
 * (optional).  While a batch exists its matrices must not be eliminated
 * again, given a new element operator or a new FP32 format (run returns
 * PHG_HIP_E_STATE); layouts the refill does not update (SELL, block copies)
 * are dropped and rebuilt from the new values when next used.  nmat <= 16. */
typedef struct phg_hip_batch phg_hip_batch;
int phg_hip_batch_create(int32_t nmat, phg_hip_space *const *sp,
                         const phg_hip_problem *const *pb, phg_hip_mat *const *A,
                         const phg_hip_vec *const *kappa, const phg_hip_vec *const *f,
                         const phg_hip_vec *const *g, phg_hip_vec *const *x,
                         int32_t maxit, double rtol, double btol, double atol,
                         double delta, phg_hip_batch **batch);
int phg_hip_batch_run(phg_hip_batch *batch, int32_t *nits, double *residual);
int phg_hip_batch_destroy(phg_hip_batch *batch);

/*------------------------------------------------- matrix-free operator */

/* Matrix-free FP32 operator for the inner iterations of
 * phg_hip_pcg_reliable / phg_hip_pcg_multi (no reference counterpart).
 * For the eliminated matrix A of a scalar problem (pb->dim 1; P1-P3 on the
 * space sp, kappa as given to phg_hip_assemble, optional mass term
 * pb->mass_coef) it stores per element the six coefficients
 * kappa_e vol_e grad lambda_a . grad lambda_b (a <= b in 1..3), plus
 * mass_coef vol_e with a mass term, and applies
 * K_e = sum_m c_m T^m (the stiffness of quad.c:5260 phgQuadGradBasAGradBas
 * and the mass of quad.c:5582 phgQuadBasDotBas, factored through fixed
 * reference tensors T^m) to x on the matrix cores (P1 / P2: FP32 MFMA; P3:
 * bf16 MFMA with operands split into three bf16 parts so products keep FP32
 * precision).  Elements are taken in
 * blocks of consecutive elements (Hilbert-curve order of their centroids); each row's sum over its
 * elements runs in a fixed order (in LDS within a block, then over blocks
 * for rows shared between blocks), so results are deterministic.  Dirichlet
 * rows and columns are treated as
 * phg_hip_mat_eliminate leaves them.  The FP64 residuals of the solvers keep
 * using the assembled matrix.  After the call the solvers use the operator;
 * phg_hip_mat_use_element_operator(A, 0 / 1) switches it off / on. */
int phg_hip_mat_element_operator(phg_hip_mat *A, const phg_hip_space *sp,
                                 const phg_hip_problem *pb, const double *kappa);
int phg_hip_mat_use_element_operator(phg_hip_mat *A, int32_t on);
/* y = A x through the operator (FP32 arithmetic, FP64 vectors in PHG's
 * numbering); for checks against phg_hip_matvec. */
int phg_hip_element_apply(const phg_hip_mat *A, const phg_hip_vec *x,
                          phg_hip_vec *y);
/* Algorithmic bytes of one application: block kernel (per element packed
 * indices and coefficients; per touched row its index, slot start, x and
 * partial slot; interior rows of y, partials of shared rows) and fix-up
 * (shared rows from their partials, Dirichlet rows). */
int phg_hip_element_bytes(const phg_hip_mat *A, int64_t *elem_bytes,
                          int64_t *gather_bytes);
/* Average time (us) of `reps` launches of the block kernel and of the
 * fix-up kernel, each measured with HIP events on the library stream. */
int phg_hip_element_time(const phg_hip_mat *A, int32_t reps, double *elem_us,
                         double *gather_us);
/* Block layout of the operator: out[0] blocks, [1] touched-row entries,
 * [2] interior rows, [3] shared rows, [4] partial sums, [5] Dirichlet rows,
 * [6] 1 when the layout was built on the device (blocks of a fixed number of
 * elements), 0 when by the host's split. */
int phg_hip_element_layout(const phg_hip_mat *A, int64_t out[7]);

/*------------------------------------------------ many right-hand sides */

/* Block of right-hand sides (no reference counterpart: solver-phg.c:495
 * solver_pcg asserts nvec == 1, so PHG solves them one after another).
 * Solves A x_c = b_c for c < nrhs <= 16 with nrhs Jacobi-PCG recurrences
 * that share one SpMM per iteration over row-major multivectors
 * ([n][16], solver numbering); 3-component problems use a 3x3 block copy
 * of the matrix.  Each column has its own tolerance
 * max(rtol*|r0|, btol*|b|, atol), alpha, beta and iteration count, as
 * phg_hip_pcg would give it.  mixed = 0: FP64 throughout; mixed = 1: FP32
 * SpMM and vectors with reliable FP64 residual updates (threshold delta),
 * convergence tested on FP64 residuals.  x holds the initial guesses on
 * input.  nits[c], residual[c] (optional) receive each column's iteration
 * count and final residual.  Returns the largest iteration count. */
int phg_hip_pcg_block(const phg_hip_mat *A, int32_t nrhs,
                      const phg_hip_vec *const *b, phg_hip_vec *const *x,
                      int32_t maxit, double rtol, double btol, double atol,
                      int32_t mixed, double delta, int32_t *nits,
                      double *residual);
/* SpMM autotuner: times the wavefront shapes -- index 0..5 = (slots per
 * block row, columns per lane) (1,1) (2,1) (4,1) (4,4) (8,4) (16,4); a lane
 * with 4 columns loads them with one 16-byte access -- in FP64 and FP32 and
 * keeps the fastest for phg_hip_pcg_block.  phg_hip_spmm_tune_time reports the
 * measured microseconds; phg_hip_spmm_bytes the bytes one SpMM moves. */
int phg_hip_spmm_tune(phg_hip_mat *A, int32_t reps);
int phg_hip_spmm_tune_time(const phg_hip_mat *A, int32_t shape, int32_t fp32,
                           double *us);
int phg_hip_spmm_bytes(const phg_hip_mat *A, int32_t fp32, int64_t *bytes);
int phg_hip_spmm_set_shape(phg_hip_mat *A, int32_t shape64, int32_t shape32);

/*------------------------------------------------------------ DOF norms */

/* dof-utils.c:2872 phgDofNormL2_ of a scalar Pk DOF v (values in PHG
 * numbering) and the L2 norm of its gradient (dof-utils.c:875
 * phgDofGradient_ then phgDofNormL2_), by element quadrature of orders 2p
 * and 2p-2, summed in a fixed order.  Drivers report the error of a
 * solution as the norms of u_h - I_h u (examples/poisson.c:437-460). */
int phg_hip_dof_norms(phg_hip_space *sp, const phg_hip_vec *v, double *l2,
                      double *grad_l2);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving or time-stepping
 * loops, harnesses that forbid hipMalloc): phg_hip_dof_norms with its
 * per-element values and partial sums in a fixed scratch array of the
 * library instead of buffers allocated per call.  Same kernels, same
 * summation order, so the same results bit for bit.  Meshes of at most
 * 2^21 elements (larger ones give PHG_HIP_E_ARG); the array (32 MiB) is held
 * by the GPU until the process ends.  All calls on the GPU share the array:
 * calls that overlap on the GPU (from several host threads) race. */
int phg_hip_dof_norms_noalloc(phg_hip_space *sp, const phg_hip_vec *v,
                              double *l2, double *grad_l2);

/*---------------------------------------------------------------- adjoint */

/* Adjoint gradient (no reference counterpart; PHG has no adjoint).  For a
 * scalar problem assembled with a P0 coefficient kappa (use_kappa = 1) and
 * eliminated (A), the forward solution u and a target ud (both FP64, PHG
 * numbering, Dirichlet values included), with
 *     J = 1/2 \int (u_h - ud_h)^2 dx = 1/2 (u - ud)^T M (u - ud)
 * (M: the Pk mass matrix with the rule of quad.c:5582 phgQuadBasDotBas):
 *   - the adjoint right-hand side M (u - ud), zero on Dirichlet rows, is
 *     accumulated in element order,
 *   - the transposed system A^T lambda = M (u - ud) is solved with the FP64
 *     PCG of phg_hip_pcg (A is symmetric after elimination) from 0 to the
 *     relative tolerance rtol,
 *   - grad[e] = dJ/dlog(kappa_e) = -kappa_e \int_e grad u_h . grad lambda_h
 *     with the stiffness rule of quad.c:5260.
 * *J receives the objective, lambda (optional) the adjoint solution,
 * *nits the adjoint PCG iterations; grad is a host array [nelem]. */
int phg_hip_adjoint_logkappa(phg_hip_space *sp, const phg_hip_problem *pb,
                             const double *kappa, const phg_hip_mat *A,
                             const phg_hip_vec *u, const phg_hip_vec *ud,
                             int32_t maxit, double rtol, double *J,
                             double *grad, phg_hip_vec *lambda, int32_t *nits);

/*--------------------------------------------------- SpMV format autotuner */

/* No reference counterpart (PHG has one CSR kernel, matvec.c:887).  The
 * solvers run their SpMV on one of these layouts of the eliminated matrix,
 * all in the solver numbering (DOFs along a Hilbert curve of their nodes):
 *   B2K, B4K   work-group CSR stream: row blocks of at most 2048 / 4096
 *              entries are loaded coalesced and staged in LDS as products;
 *              one work-item sums a row in storage order;
 *   B2K_PAR, B4K_PAR
 *              the same with several lanes per row for the sums;
 *   WAVE       the same per wavefront (512-entry blocks, wavefront-private
 *              LDS, no work-group barriers);
 *   SELL       SELL-64: slices of 64 rows sorted by length inside windows of
 *              4096 rows, stored column-major;
 *   DICT, DICT_NP
 *              dictionary-compressed CSR: blocks of <= 4096 entries and
 *              <= 256 rows, each with a dictionary of <= 256 column chunks
 *              of width 256; an entry stores its column in 16 bits
 *              (slot, offset), 6 bytes per FP32 entry instead of 8.  DICT
 *              loads the next block while reducing the current one, DICT_NP
 *              does not (fewer registers).  Not available for matrices with
 *              a row that does not fit a block (phg_hip_mat_tune_time then
 *              reports -1);
 *   XS, XS_PF  staged lines: the same blocks with a dictionary of <= 128
 *              16-column lines of x; the SpMV reads those lines into LDS
 *              with coalesced loads and an entry's 16-bit code (slot,
 *              offset) is its LDS index (XS_PF loads the next block while
 *              reducing the current one).
 * phg_hip_mat_tune times each layout on the matrix (HIP events, `reps`
 * launches) in FP64 and FP32 and keeps the fastest for phg_hip_pcg and
 * phg_hip_pcg_mixed respectively. */
#define PHG_HIP_FMT_B2K        0
#define PHG_HIP_FMT_B4K        1
#define PHG_HIP_FMT_B2K_PAR    2
#define PHG_HIP_FMT_B4K_PAR    3
#define PHG_HIP_FMT_WAVE       4
#define PHG_HIP_FMT_SELL       5
#define PHG_HIP_FMT_DICT       6
#define PHG_HIP_FMT_DICT_NP    7
#define PHG_HIP_FMT_XS         8
#define PHG_HIP_FMT_XS_PF      9
#define PHG_HIP_FMT_COUNT      10
int phg_hip_mat_tune(phg_hip_mat *A, int32_t reps);
int phg_hip_mat_tune_time(const phg_hip_mat *A, int32_t fmt, int32_t fp32,
                          double *us);
int phg_hip_mat_set_format(phg_hip_mat *A, int32_t fmt64, int32_t fmt32);
/* Bytes one SpMV moves in a layout (matrix streams + one read of x + one
 * write of y; x gathers counted once): the algorithmic bytes for the
 * bandwidth figures.  The layout must have been built (by the tuner or by
 * a solve that uses it). */
int phg_hip_mat_format_bytes(const phg_hip_mat *A, int32_t fmt, int32_t fp32,
                             int64_t *bytes);
int phg_hip_mat_get_format(const phg_hip_mat *A, int32_t *fmt64,
                           int32_t *fmt32);

/*------------------------------------------------------------- profiling */

/* Per-kernel timing, the role of perf.c's counters (perf.c:490
 * phgPerfGetMflops) for the GPU path: with profiling on, every PCG kernel
 * launch is bracketed by HIP events on the library stream and solves run
 * one iteration per host round trip.  phg_hip_profile resets the totals. */
#define PHG_HIP_K_PCG_P64     0   /* p = z + beta p, FP64            */
#define PHG_HIP_K_SPMV64      1   /* Ap = A p and (p, Ap), FP64       */
#define PHG_HIP_K_UPDATE64    2   /* x, r updates and norms, FP64     */
#define PHG_HIP_K_PCG_P32     3
#define PHG_HIP_K_SPMV32      4   /* FP32 inner solve of pcg_mixed    */
#define PHG_HIP_K_UPDATE32    5
int phg_hip_profile(int32_t on);
/* Bandwidth reference: `reps` launches of a plain copy of n doubles
 * (16 n bytes each); *us = average launch time.  Used to scale HBM
 * counters and to state the attainable copy bandwidth. */
int phg_hip_stream_copy(int64_t n, int32_t reps, double *us);
int phg_hip_profile_read(int32_t kind, int64_t *launches, double *ms);

#ifdef __cplusplus
}
#endif
#endif
