/* phg_oem_hip.h -- PHG's solver plugin interface for the MI355X path.
 *
 * PHG selects a linear solver through an OEM_SOLVER vtable (solver.h:60-90)
 * registered in phgSolverList_ (solver.c:46-160), and hooks matrix-free
 * operators and preconditioners through MV_FUNC (matvec.h:60) and PC_PROC
 * (solver.h:95).  libphg_oem_hip.so provides, on top of the C ABI of
 * phg_hip.h:
 *   phgSolverHipPCG_       "hip_pcg": solver-phg.c:2430 PCG_Solve semantics
 *                          (Jacobi PC, FP64; optionally the mixed-precision
 *                          reliable-update PCG), on the GPU
 *   phgSolverHipBiCGstab_  "hip_bicgstab": solver-phg.c:2457 BiCGstab_Solve
 *   phgSolverHipGMRES_     "hip_gmres": solver-phg.c:2484 GMRES_Solve with
 *                          PHG's default flgmres_r
 *   phgHipMatVec           an MV_FUNC: y = A x on the GPU for a matrix-free
 *                          MAT wrapping a device matrix (phgMatCreateHip)
 *   phgHipJacobiPC         a PC_PROC: solver-phg.c:207 jacobi_pc on the GPU
 * and, for programs built without PHG, serial stand-ins of the few PHG
 * functions these call and of PHG's solver driver (phgSolverMat2Solver /
 * Assemble / VecSolve / Destroy, solver.c:840-1178; phg_amd/host/phg_mini.c,
 * part of the same library; inside PHG's tree only phg_oem_hip.c is added
 * and PHG's own functions are used).
 *
 * The types below re-declare, in PHG's field order, the structs of a serial
 * PHG build (USE_MPI 0, USE_OMP 0; FLOAT double, INT int as configure
 * defaults them), so a binary of the plugin built here reads and writes the
 * same fields as one built inside PHG's tree.  Inside PHG the plugin source
 * is compiled against phg.h instead of this type section.
 */
#ifndef PHG_OEM_HIP_H
#define PHG_OEM_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "phg_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/*------------------------------------------------ PHG base types (serial) */
typedef double FLOAT;            /* phg.h:81-146, FT_PHG == FT_DOUBLE */
typedef int INT;                 /* phg.h:97 TYPEOF_PHG_INT (int) */
typedef int BOOLEAN;             /* phg.h:263 */
typedef unsigned char BYTE;
typedef int MPI_Comm;            /* phg-mpi.h:75, no MPI */
#ifndef TRUE
#define TRUE 1
#define FALSE 0
#endif
#define MPI_COMM_WORLD 0         /* phg-mpi.h:77 */

/* matvec.h:26-30 */
typedef enum { MAT_OP_N, MAT_OP_T, MAT_OP_D } MAT_OP;
typedef enum { M_UNSYM = 0, M_SSYM = 1, M_SYM = 2, M_SPD = 3 } SYMMETRY;
/* matvec.h:78-92 */
typedef enum {
    PHG_UNPACKED = 0, PHG_PACKED = 1, PHG_DENSE = 2, PHG_MATRIX_FREE = 3,
    PHG_DESTROYED = 4
} MAT_TYPE;
typedef enum { PHG_UNDEFINED = 0, PHG_ADD = 1, PHG_REPLACE = 2 } MAT_MODE;

struct DOF_;

/* map.h:83-140 (serial: no cinfo) */
typedef struct MAP_ {
    FLOAT lif;
    struct DOF_ **dofs;
    struct MAP_ *map_nb;
    INT *D2Lmap, *L2Vmap, *O2Gmap, *ordering;
    INT *partition;              /* [nprocs + 1] */
    INT *offset;
    INT *bdry;                   /* sorted vector indices of boundary rows */
    INT *bdry_partition;
    INT bdry_nlocal, bdry_localsize, bdry_nglobal;
    INT D2Lmap_size, L2Vmap_size;
    INT nlocal, nglobal, localsize;
    MPI_Comm comm;
    int rank, nprocs;
    int halo_type;
    int refcount;
    BYTE ndof;
    BOOLEAN bogus_Lmap;
} MAP;

typedef struct MAT_ MAT;

/* matvec.h:33-55 (serial: no O2Gmap, ordering, localsize, alloc) */
typedef struct VEC_ {
    int magic;
    struct MAP_ *map;
    FLOAT *data;
    FLOAT *offp_data;
    MAT *mat;
    int nvec;
    BOOLEAN assembled;
} VEC;

/* matvec.h:60: y = op(A) x of a matrix-free matrix */
typedef int (*MV_FUNC)(MAT_OP op, struct MAT_ *A, VEC *x, VEC *y);

/* matvec.h:63-68 */
typedef struct {
    FLOAT *data;
    INT *cols;
    INT ncols;
    INT alloc;
} MAT_ROW;

/* matvec.h:72-77 */
typedef struct {
    FLOAT *lu;
    int *pvt;
    INT *map;
    int n;
} BDRY_EQNS;

/* matvec.h:94-170 (serial, no OpenMP locks) */
struct MAT_ {
    int magic;
    struct MAP_ *rmap, *cmap;
    MAT_ROW *rows, row_buffer;
    FLOAT *diag;
    FLOAT *diag1;
    MAT_TYPE type;
    INT *packed_cols;
    FLOAT *packed_data;
    size_t *packed_ind;
    size_t nnz_d, nnz_o;
    INT localsize;
    BOOLEAN handle_bdry_eqns;
    FLOAT mv_a, mv_b;
    MAT *mv_x;
    MV_FUNC mv_func;
    void **mv_data;
    INT bdry_localsize, *bdry;
    BDRY_EQNS *bdry_eqns;
    INT bdry_eqns_count, bdry_eqns_allocated;
    MAT_ROW *bdry_rows;
    void *blocks;
    int refcount;
    MAT_MODE mode;
    BOOLEAN assembled;
};

/* solver.h:26-54 */
typedef struct SOLVER_ {
    FLOAT residual;
    FLOAT rtol;
    FLOAT atol;
    FLOAT btol;
    double cond;
    MAT *mat, *mat_bak;
    VEC *rhs;
    FLOAT *D;
    struct OEM_SOLVER_ *oem_solver;
    void *oem_data;
    struct SOLVER_PC_ *pc;
    struct SOLVER_ *St;
    int symmetry;
    int matrix_free;
    int nits;
    int maxit;
    BOOLEAN monitor;
    BOOLEAN scaling;
    BOOLEAN warn_maxit;
    BOOLEAN pc_option_flag;
    BOOLEAN rhs_updated;
    BOOLEAN sent_to_oem;
    BOOLEAN oem_created;
    BOOLEAN assembled;
} SOLVER;

/* solver.h:60-90 */
typedef struct OEM_SOLVER_ {
    const char *name;
    int (*RegisterOptions)(void);
    int (*Initialize)(int *argc, char ***argv);
    int (*Finalize)(void);
    int (*Init)(SOLVER *solver);
    int (*Create)(SOLVER *solver);
    int (*Destroy)(SOLVER *solver);
    int (*AddMatrixEntries)(SOLVER *solver, INT nrows, INT *rows, INT ncols,
                            INT *cols, FLOAT *values);
    int (*AddRHSEntries)(SOLVER *solver, INT n, INT *ni, FLOAT *values);
    int (*Assemble)(SOLVER *solver);
    int (*SetPC)(SOLVER *solver);
    int (*Solve)(SOLVER *solver, VEC *x, BOOLEAN destroy);
    void **(*GetSolver)(SOLVER *solver);
    void **(*GetMatrix)(SOLVER *solver);
    void **(*GetPC)(SOLVER *solver);
    SYMMETRY req_symmetry;
    BOOLEAN iterative;
    BOOLEAN parallel;
    BOOLEAN allow_matrix_free;
    BOOLEAN allow_extern_pc;
} OEM_SOLVER;

/* solver.h:95-103 */
typedef void (*PC_PROC)(void *ctx, VEC *b, VEC **x);
typedef struct SOLVER_PC_ {
    char *name;
    PC_PROC pc_proc;
    void *ctx;
    SOLVER *solver_owned;
} SOLVER_PC;

/*------------------------------------------------------------ the plugin */

/* The vtables, registered next to phgSolverPCG_ in solver.c:46-160 (or
 * passed to phgSolverMat2Solver / phgSolverCreate directly).  Flow, as
 * solver.c drives any OEM solver:
 *   Init       (phgSolverMat2Solver)  private data;
 *   Create     (first phgSolverAssemble) row buffer for rmap->nlocal rows;
 *   AddMatrixEntries (send_matrix_to_oem_solver, solver.c:223): the rows of
 *              the assembled matrix (after phgMatAssemble's boundary
 *              treatment), one call per row in row order;
 *   Assemble   uploads them to the GPU as an eliminated CSR (the boundary
 *              rows rmap->bdry flagged), or -- for a matrix-free MAT whose
 *              mv_func is phgHipMatVec -- takes its device matrix as is;
 *   Solve      uploads solver->rhs->data (lifted by phgSolverUpdateRHS, which
 *              runs after Assemble, solver.c:1125) and x, solves on the GPU,
 *              downloads x; returns nits >= 0 (< 0 on error) and sets
 *              solver->residual / nits; destroy frees the matrices
 *              (phgMatDestroy(&solver->mat) and the device copy); |b| = 0
 *              gives x = 0, 0 iterations and residual 1.0, the value
 *              PCG_Solve starts from;
 *   Destroy    frees the private data and the device copy.
 * A zero diagonal ends the run through phgError, as jacobi_pc does
 * (solver-phg.c:233). */
extern OEM_SOLVER phgSolverHipPCG_;
extern OEM_SOLVER phgSolverHipBiCGstab_;
extern OEM_SOLVER phgSolverHipGMRES_;
#define SOLVER_HIP_PCG (&phgSolverHipPCG_)
#define SOLVER_HIP_BiCGstab (&phgSolverHipBiCGstab_)
#define SOLVER_HIP_GMRES (&phgSolverHipGMRES_)

/* Options of a solver made with one of the vtables above, the role of the
 * -hip_pcg_mixed / -gmres_restart / -lgmres_augk options PHG would register
 * (solver-phg.c:2343 Init reads them).  mixed (hip_pcg only): 1 runs the
 * FP32 reliable-update PCG (FP64 residuals, delta 0.1); restart / augk
 * (hip_gmres): PHG's defaults 10 and 2; tune: 1 times the SpMV layouts at
 * Assemble and keeps the fastest.  Call after phgSolverMat2Solver. */
int phgSolverHipSetOptions(SOLVER *solver, int mixed, int restart, int augk,
                           int tune);

/* The role of -gmres_type (solver-phg.c:458 GMRESRegisterOptions keyword
 * list gmres_names, solver-phg.c:71) for a hip_gmres solver: one of
 * "gmres_l", "gmres_r", "fgmres_r", "lgmres_l", "lgmres_r", "flgmres_r"
 * (the default).  Returns 0, or -1 for another solver or an unknown name. */
int phgSolverHipSetGMRESType(SOLVER *solver, const char *name);

/* MV_FUNC (matvec.h:60) for a matrix-free MAT made by phgMatCreateHip:
 * y = A x (MAT_OP_N; MAT_OP_T for a symmetric device matrix; MAT_OP_D: the
 * diagonal) with x and y host vectors of PHG; the product runs on the GPU in
 * PHG's row order (phg_hip_matvec).  Returns 0, or -1 (message through
 * phgError with code 0) for an operation the device matrix does not hold. */
int phgHipMatVec(MAT_OP op, MAT *A, VEC *x, VEC *y);

/* A matrix-free PHG MAT (phgMatCreateMatrixFree, matvec.c:1419) around an
 * assembled and eliminated device matrix, e.g. one phg_hip_assemble built
 * from PHG's element loop lowered to the GPU.  symmetric: the matrix is
 * symmetric (MAT_OP_T allowed).  The MAT borrows A; phgMatDestroyHip frees
 * the MAT and its context, not A. */
MAT *phgMatCreateHip(phg_hip_mat *A, BOOLEAN symmetric);
void phgMatDestroyHip(MAT **mat);

/* PC_PROC (solver.h:95): solver-phg.c:207 jacobi_pc, x = b ./ diag(A), on
 * the GPU, for ctx = a MAT made by phgMatCreateHip; installed with
 * phgSolverSetPC(solver, mat, phgHipJacobiPC).  A zero diagonal is fatal
 * (phgError, solver-phg.c:233). */
void phgHipJacobiPC(void *ctx, VEC *b, VEC **x);

/*------------------- serial stand-ins of PHG (phg_mini.c; PHG's own inside PHG) */
/* utils.c:1996 phgError (code 0: message only), utils.c:2062 phgWarning */
void phgError(int code, const char *fmt, ...);
void phgWarning(const char *fmt, ...);
void *phgAlloc(size_t size);
void *phgCalloc(size_t nmemb, size_t size);
void phgFree(void *ptr);
/* map.c phgMapCreateSimpleMap / phgMapDestroy (one process) */
MAP *phgMapCreateSimpleMap(MPI_Comm comm, INT n, INT N);
void phgMapDestroy(MAP **map);
/* matvec.c phgMapCreateVec / phgVecDestroy */
VEC *phgMapCreateVec(MAP *map, int nvec);
void phgVecDestroy(VEC **vec);
/* matvec.c:1419 phgMatCreateMatrixFree (the varargs are further mv_data
 * pointers, NULL-terminated) and phgMatDestroy (reference counted) */
MAT *phgMatCreateMatrixFree(MPI_Comm comm, INT m, INT M, MV_FUNC mv_func,
                            void *mv_data, ...);
void phgMatDestroy(MAT **mat);
/* solver.c:840 phgSolverMat2Solver (PHG's option defaults, solver.c:164-168:
 * maxit 10000, rtol 1e-10, btol 1e-10, atol 0), :928 phgSolverDestroy, :961
 * phgSolverAssemble, :1102 phgSolverVecSolve, :1232-1245 SetMaxIt / SetTol,
 * solver.c:1303 phgSolverSetPC.  phgSolverUpdateRHS: PHG's update_rhs lifts
 * the right-hand side with the boundary columns phgMatAssemble saved; these
 * stand-ins hold no PHG matrix rows, so the caller hands over the lifted
 * right-hand side and the stand-in only marks it so. */
SOLVER *phgSolverMat2Solver(OEM_SOLVER *oem_solver, MAT *mat);
int phgSolverDestroy(SOLVER **solver);
int phgSolverAssemble(SOLVER *solver);
void phgSolverUpdateRHS(SOLVER *solver);
int phgSolverVecSolve(SOLVER *solver, BOOLEAN destroy, VEC *x);
void phgSolverSetMaxIt(SOLVER *solver, int maxit);
void phgSolverSetTol(SOLVER *solver, FLOAT user_tol);
void phgSolverSetPC(SOLVER *solver, void *ctx, PC_PROC pc_func);

#ifdef __cplusplus
}
#endif
#endif
