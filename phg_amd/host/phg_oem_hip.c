/* phg_oem_hip.c -- the OEM_SOLVER plugin (solver.h:60-90) that runs PHG's
 * Krylov solvers on the MI355X through the C ABI of phg_hip.h, with the
 * MV_FUNC (matvec.h:60) and PC_PROC (solver.h:95) hooks of the same path.
 *
 * This file uses only what PHG's public headers declare (SOLVER, OEM_SOLVER,
 * MAT, VEC, MAP fields; phgError, phgCalloc, phgFree, phgMatDestroy,
 * phgMatCreateMatrixFree), so it is the file a maintainer adds to PHG's
 * src/ (as solver-hip.c) and registers in solver.c:46-160.  Built here it
 * links against the serial stand-ins of phg_mini.c.
 */
#include "phg_oem_hip.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { HIP_PCG = 0, HIP_BICGSTAB = 1, HIP_GMRES = 2 };

/* private data of one solver (solver->oem_data) */
typedef struct {
    int kind;                 /* HIP_PCG / HIP_BICGSTAB / HIP_GMRES */
    int mixed, restart, augk, tune;
    int gmres_type;           /* PHG_HIP_GMRES_L .. PHG_HIP_FLGMRES_R */
    phg_hip_mat *A;           /* device system */
    int owns_A;               /* 0: borrowed from a matrix-free MAT */
    phg_hip_vec *b, *x;
    /* rows received through AddMatrixEntries (send_matrix_to_oem_solver) */
    INT n, next_row;
    int64_t *rowptr;          /* [n + 1] */
    int32_t *cols;
    double *vals;
    size_t nnz, cap;
    int bad_rows;             /* a row arrived out of order or twice */
} HIP_DATA;

/* context of a matrix-free MAT around a device matrix (mv_data[0]) */
typedef struct {
    phg_hip_mat *A;
    BOOLEAN symmetric;
    phg_hip_vec *x, *y;       /* device staging of the host vectors */
} HIP_MV;

#define HIP_CHECK(call, what)                                                 \
    do {                                                                      \
        if ((call) != PHG_HIP_OK)                                             \
            phgError(1, "%s: %s\n", what, phg_hip_last_error());              \
    } while (0)

static int HipInitialize(int *argc, char ***argv)
{
    (void)argc;
    (void)argv;
    /* utils.c:555 phgInit role on the GPU side: device 0 of this process */
    return phg_hip_init(0) == PHG_HIP_OK ? 0 : -1;
}

static int HipFinalize(void) { return phg_hip_finalize(); }

static int init_kind(SOLVER *solver, int kind)
{
    HIP_DATA *d = phgCalloc(1, sizeof(*d));
    d->kind = kind;
    d->restart = 10;          /* -gmres_restart default (solver-phg.c) */
    d->augk = 2;              /* -lgmres_augk default */
    d->gmres_type = PHG_HIP_FLGMRES_R;  /* -gmres_type default */
    solver->oem_data = d;
    /* the library needs its stream before any matrix is uploaded */
    if (phg_hip_init(0) != PHG_HIP_OK)
        phgError(1, "phg_hip_init: %s\n", phg_hip_last_error());
    return 0;
}

static int HipInitPCG(SOLVER *solver) { return init_kind(solver, HIP_PCG); }
static int HipInitBiCGstab(SOLVER *solver) { return init_kind(solver, HIP_BICGSTAB); }
static int HipInitGMRES(SOLVER *solver) { return init_kind(solver, HIP_GMRES); }

static void free_rows(HIP_DATA *d)
{
    phgFree(d->rowptr);
    phgFree(d->cols);
    phgFree(d->vals);
    d->rowptr = NULL;
    d->cols = NULL;
    d->vals = NULL;
    d->nnz = d->cap = 0;
}

static void free_device(HIP_DATA *d)
{
    if (d->owns_A)
        phg_hip_mat_destroy(d->A);
    d->A = NULL;
    d->owns_A = 0;
    phg_hip_vec_destroy(d->b);
    phg_hip_vec_destroy(d->x);
    d->b = d->x = NULL;
}

/* called once, by the first phgSolverAssemble (solver.c:236-239) */
static int HipCreate(SOLVER *solver)
{
    HIP_DATA *d = solver->oem_data;
    d->n = solver->mat->rmap->nlocal;
    d->next_row = 0;
    free_rows(d);
    d->rowptr = phgCalloc((size_t)d->n + 1, sizeof(*d->rowptr));
    return 0;
}

static int HipDestroy(SOLVER *solver)
{
    HIP_DATA *d = solver->oem_data;
    if (d == NULL)
        return 0;
    free_rows(d);
    free_device(d);
    phgFree(d);
    solver->oem_data = NULL;
    return 0;
}

/* send_matrix_to_oem_solver (solver.c:255-273): each row of the assembled
 * matrix once, in order, with its columns in PHG's storage order */
static int HipAddMatrixEntries(SOLVER *solver, INT nrows, INT *rows, INT ncols,
                               INT *cols, FLOAT *values)
{
    HIP_DATA *d = solver->oem_data;
    INT i0 = solver->mat->rmap->partition[solver->mat->rmap->rank];
    for (INT k = 0; k < nrows; k++) {
        if (rows[k] - i0 != d->next_row || d->next_row >= d->n) {
            d->bad_rows = 1;
            return -1;
        }
        if (d->nnz + (size_t)ncols > d->cap) {
            size_t cap = d->cap ? 2 * d->cap : 1024;
            while (cap < d->nnz + (size_t)ncols)
                cap *= 2;
            int32_t *c = realloc(d->cols, cap * sizeof(*c));
            double *v = realloc(d->vals, cap * sizeof(*v));
            if (c == NULL || v == NULL)
                phgError(1, "hip solver: out of host memory (%zu entries)\n", cap);
            d->cols = c;
            d->vals = v;
            d->cap = cap;
        }
        for (INT j = 0; j < ncols; j++) {
            d->cols[d->nnz + (size_t)j] = cols[j];
            d->vals[d->nnz + (size_t)j] = values[(size_t)k * ncols + j];
        }
        d->nnz += (size_t)ncols;
        d->rowptr[++d->next_row] = (int64_t)d->nnz;
    }
    return 0;
}

static void setup_vectors(HIP_DATA *d, INT n)
{
    HIP_CHECK(phg_hip_vec_create(n, &d->b), "hip solver");
    HIP_CHECK(phg_hip_vec_create(n, &d->x), "hip solver");
}

/* phgSolverAssemble (solver.c:1075), after the rows were sent */
static int HipAssemble(SOLVER *solver)
{
    HIP_DATA *d = solver->oem_data;
    MAT *m = solver->mat;
    if (d->A != NULL)
        return 0;
    if (m->type == PHG_MATRIX_FREE) {
        if (m->mv_func != phgHipMatVec)
            phgError(1, "%s: a matrix-free matrix must wrap a device matrix "
                     "(phgMatCreateHip)\n", solver->oem_solver->name);
        HIP_MV *mv = m->mv_data[0];
        d->A = mv->A;
        d->owns_A = 0;
        setup_vectors(d, m->rmap->nlocal);
        return 0;
    }
    if (d->bad_rows || d->next_row != d->n)
        phgError(1, "%s: matrix rows must arrive once each, in order (%d of %d)\n",
                 solver->oem_solver->name, (int)d->next_row, (int)d->n);
    /* Dirichlet rows (map.h:115 bdry: sorted vector indices) */
    uint8_t *bdry = phgCalloc((size_t)d->n, 1);
    MAP *map = m->rmap;
    for (INT k = 0; map->bdry != NULL && k < map->bdry_nlocal; k++)
        if (map->bdry[k] >= 0 && map->bdry[k] < d->n)
            bdry[map->bdry[k]] = 1;
    double *zero = phgCalloc((size_t)d->n, sizeof(double));
    /* rows after phgMatAssemble: boundary columns moved, boundary rows
     * holding their diagonal, i.e. the eliminated system */
    int rc = phg_hip_mat_from_csr(d->n, d->rowptr, d->cols, d->vals, zero, bdry, 1,
                                  &d->A);
    phgFree(zero);
    phgFree(bdry);
    if (rc != PHG_HIP_OK)
        phgError(1, "%s: %s\n", solver->oem_solver->name, phg_hip_last_error());
    d->owns_A = 1;
    free_rows(d);
    if (d->tune)
        HIP_CHECK(phg_hip_mat_tune(d->A, 10), solver->oem_solver->name);
    setup_vectors(d, d->n);
    return 0;
}

/* solver-phg.c:2430 PCG_Solve / :2457 BiCGstab_Solve / :2484 GMRES_Solve */
static int HipSolve(SOLVER *solver, VEC *x, BOOLEAN destroy)
{
    HIP_DATA *d = solver->oem_data;
    FLOAT residual = 1.0;     /* the value PCG_Solve starts from */
    int nits;
    if (d->A == NULL)
        phgError(1, "%s: Solve before Assemble\n", solver->oem_solver->name);
    /* the right-hand side as phgSolverUpdateRHS left it (solver.c:1125) */
    HIP_CHECK(phg_hip_vec_upload(d->b, solver->rhs->data), solver->oem_solver->name);
    HIP_CHECK(phg_hip_vec_upload(d->x, x->data), solver->oem_solver->name);
    switch (d->kind) {
    case HIP_PCG:
        if (d->mixed) {
            int32_t upd = 0;
            nits = phg_hip_pcg_reliable(d->A, d->b, d->x, solver->maxit, solver->rtol,
                                        solver->btol, solver->atol, 0.1, &upd, &residual);
        }
        else
            nits = phg_hip_pcg(d->A, d->b, d->x, solver->maxit, solver->rtol,
                               solver->btol, solver->atol, &residual);
        break;
    case HIP_BICGSTAB:
        nits = phg_hip_bicgstab(d->A, d->b, d->x, solver->maxit, solver->rtol,
                                solver->btol, solver->atol, &residual);
        break;
    default:
        nits = phg_hip_gmres_type(d->A, d->b, d->x, d->gmres_type, d->restart, d->augk,
                                  solver->maxit, solver->rtol, solver->btol, solver->atol,
                                  &residual);
        break;
    }
    if (nits == PHG_HIP_E_ZERO_DIAG)
        /* jacobi_pc, solver-phg.c:233 */
        phgError(1, "%s: zero diagonal (%s)\n", solver->oem_solver->name,
                 phg_hip_last_error());
    if (nits < 0) {
        phgError(0, "%s: %s\n", solver->oem_solver->name, phg_hip_last_error());
        return nits;
    }
    HIP_CHECK(phg_hip_vec_download(d->x, x->data), solver->oem_solver->name);
    if (destroy) {
        free_device(d);
        phgMatDestroy(&solver->mat);
        solver->sent_to_oem = TRUE;
    }
    solver->residual = residual;
    return solver->nits = nits;
}

static void **HipGetMatrix(SOLVER *solver)
{
    HIP_DATA *d = solver->oem_data;
    return (void **)&d->A;
}

int phgSolverHipSetOptions(SOLVER *solver, int mixed, int restart, int augk, int tune)
{
    OEM_SOLVER *o = solver->oem_solver;
    if (o != &phgSolverHipPCG_ && o != &phgSolverHipBiCGstab_ && o != &phgSolverHipGMRES_)
        return -1;
    HIP_DATA *d = solver->oem_data;
    d->mixed = mixed;
    d->restart = restart > 0 ? restart : 10;
    d->augk = augk > 0 ? augk : 2;
    d->tune = tune;
    return 0;
}

int phgSolverHipSetGMRESType(SOLVER *solver, const char *name)
{
    /* the keywords of -gmres_type, solver-phg.c:71 gmres_names */
    static const char *names[] = {"gmres_l", "gmres_r", "fgmres_r",
                                  "lgmres_l", "lgmres_r", "flgmres_r"};
    if (solver->oem_solver != &phgSolverHipGMRES_ || name == NULL)
        return -1;
    for (int i = 0; i < (int)(sizeof(names) / sizeof(names[0])); i++)
        if (strcmp(name, names[i]) == 0) {
            ((HIP_DATA *)solver->oem_data)->gmres_type = PHG_HIP_GMRES_L + i;
            return 0;
        }
    return -1;
}

/*-------------------------------------------------------- MV_FUNC, PC_PROC */

static HIP_MV *hip_mv(MAT *A)
{
    if (A == NULL || A->type != PHG_MATRIX_FREE || A->mv_func != phgHipMatVec ||
        A->mv_data == NULL)
        return NULL;
    return A->mv_data[0];
}

int phgHipMatVec(MAT_OP op, MAT *A, VEC *x, VEC *y)
{
    HIP_MV *mv = hip_mv(A);
    if (mv == NULL) {
        phgError(0, "phgHipMatVec: not a matrix made by phgMatCreateHip\n");
        return -1;
    }
    if (op == MAT_OP_D || (op == MAT_OP_T && !mv->symmetric)) {
        phgError(0, "phgHipMatVec: operation %d not held by the device matrix\n", (int)op);
        return -1;
    }
    HIP_CHECK(phg_hip_vec_upload(mv->x, x->data), "phgHipMatVec");
    HIP_CHECK(phg_hip_matvec(1.0, mv->A, mv->x, 0.0, mv->y), "phgHipMatVec");
    HIP_CHECK(phg_hip_vec_download(mv->y, y->data), "phgHipMatVec");
    return 0;
}

MAT *phgMatCreateHip(phg_hip_mat *A, BOOLEAN symmetric)
{
    int64_t n = 0, nnz = 0;
    HIP_CHECK(phg_hip_mat_info(A, &n, &nnz), "phgMatCreateHip");
    HIP_MV *mv = phgCalloc(1, sizeof(*mv));
    mv->A = A;
    mv->symmetric = symmetric;
    HIP_CHECK(phg_hip_vec_create(n, &mv->x), "phgMatCreateHip");
    HIP_CHECK(phg_hip_vec_create(n, &mv->y), "phgMatCreateHip");
    MAT *m = phgMatCreateMatrixFree(MPI_COMM_WORLD, (INT)n, (INT)n, phgHipMatVec, mv, NULL);
    /* the device matrix is eliminated: nothing left for phgMatAssemble */
    m->assembled = TRUE;
    return m;
}

void phgMatDestroyHip(MAT **mat)
{
    HIP_MV *mv = hip_mv(*mat);
    if (mv != NULL && (*mat)->refcount == 0) {
        phg_hip_vec_destroy(mv->x);
        phg_hip_vec_destroy(mv->y);
        phgFree(mv);
    }
    phgMatDestroy(mat);
}

void phgHipJacobiPC(void *ctx, VEC *b, VEC **x)
{
    HIP_MV *mv = hip_mv((MAT *)ctx);
    if (mv == NULL)
        phgError(1, "phgHipJacobiPC: ctx must be a matrix made by phgMatCreateHip\n");
    HIP_CHECK(phg_hip_vec_upload(mv->x, b->data), "phgHipJacobiPC");
    int rc = phg_hip_pc_jacobi(mv->A, mv->x, mv->y);
    if (rc == PHG_HIP_E_ZERO_DIAG)
        phgError(1, "phgHipJacobiPC: %s\n", phg_hip_last_error());
    HIP_CHECK(rc, "phgHipJacobiPC");
    HIP_CHECK(phg_hip_vec_download(mv->y, (*x)->data), "phgHipJacobiPC");
}

/*---------------------------------------------------------------- vtables */

OEM_SOLVER phgSolverHipPCG_ = {
    "hip_pcg", NULL, HipInitialize, HipFinalize, HipInitPCG, HipCreate, HipDestroy,
    HipAddMatrixEntries, NULL, HipAssemble, NULL, HipSolve, NULL, HipGetMatrix, NULL,
    M_SPD, TRUE, FALSE, TRUE, FALSE
};

OEM_SOLVER phgSolverHipBiCGstab_ = {
    "hip_bicgstab", NULL, HipInitialize, HipFinalize, HipInitBiCGstab, HipCreate,
    HipDestroy, HipAddMatrixEntries, NULL, HipAssemble, NULL, HipSolve, NULL,
    HipGetMatrix, NULL, M_UNSYM, TRUE, FALSE, TRUE, FALSE
};

OEM_SOLVER phgSolverHipGMRES_ = {
    "hip_gmres", NULL, HipInitialize, HipFinalize, HipInitGMRES, HipCreate, HipDestroy,
    HipAddMatrixEntries, NULL, HipAssemble, NULL, HipSolve, NULL, HipGetMatrix, NULL,
    M_UNSYM, TRUE, FALSE, TRUE, FALSE
};
