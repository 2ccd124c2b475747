/* phg_mini.c -- serial stand-ins of the PHG functions the OEM plugin
 * (phg_oem_hip.c) calls and of PHG's solver driver, for programs that use
 * the plugin without PHG.  Inside PHG's tree this file is not compiled:
 * PHG's own utils.c, map.c, matvec.c and solver.c provide these functions.
 *
 * Matrices reach the driver already assembled (PHG's row insertion and
 * phgMatAssemble's boundary treatment are PHG's, or the device assembly's):
 * either in PHG's packed form (packed_ind / packed_cols / packed_data,
 * matvec.h:102-115) or row form (rows[]), or matrix-free.  The right-hand
 * side handed to phgSolverVecSolve is the lifted one (see
 * phgSolverUpdateRHS below).
 */
#include "phg_oem_hip.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/*------------------------------------------------ utils.c:1996-2070 roles */

void phgError(int code, const char *fmt, ...)
{
    va_list ap;
    fprintf(stderr, "\n* *** ERROR: ");
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fflush(stderr);
    if (code == 0)
        return;
    exit(code);     /* utils.c:1931 phgAbort with one process */
}

void phgWarning(const char *fmt, ...)
{
    va_list ap;
    fprintf(stderr, "*** WARNING: ");
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fflush(stderr);
}

void *phgAlloc(size_t size)
{
    void *p = malloc(size ? size : 1);
    if (p == NULL)
        phgError(1, "out of memory (%zu bytes)\n", size);
    return p;
}

void *phgCalloc(size_t nmemb, size_t size)
{
    void *p = calloc(nmemb ? nmemb : 1, size ? size : 1);
    if (p == NULL)
        phgError(1, "out of memory (%zu x %zu bytes)\n", nmemb, size);
    return p;
}

void phgFree(void *ptr) { free(ptr); }

/*----------------------------------------------------- maps and vectors */

MAP *phgMapCreateSimpleMap(MPI_Comm comm, INT n, INT N)
{
    if (N != n && N != -1)
        phgError(1, "phgMapCreateSimpleMap: one process, n (%d) != N (%d)\n", n, N);
    MAP *map = phgCalloc(1, sizeof(*map));
    map->nlocal = map->nglobal = map->localsize = n;
    map->partition = phgCalloc(2, sizeof(INT));
    map->partition[1] = n;
    map->comm = comm;
    map->rank = 0;
    map->nprocs = 1;
    return map;
}

void phgMapDestroy(MAP **map_ptr)
{
    MAP *map = *map_ptr;
    if (map == NULL)
        return;
    *map_ptr = NULL;
    if (map->refcount > 0) {        /* map.c:1575-1580 */
        map->refcount--;
        return;
    }
    phgFree(map->partition);
    phgFree(map->bdry);
    phgFree(map);
}

VEC *phgMapCreateVec(MAP *map, int nvec)
{
    VEC *v = phgCalloc(1, sizeof(*v));
    map->refcount++;
    v->map = map;
    v->nvec = nvec;
    v->data = nvec > 0 ? phgCalloc((size_t)map->nlocal * nvec, sizeof(FLOAT)) : NULL;
    v->assembled = TRUE;
    return v;
}

void phgVecDestroy(VEC **vec_ptr)
{
    VEC *v = *vec_ptr;
    if (v == NULL)
        return;
    phgFree(v->data);
    phgFree(v->offp_data);
    phgMapDestroy(&v->map);
    phgFree(v);
    *vec_ptr = NULL;
}

/*------------------------------------------------------------- matrices */

MAT *phgMatCreateMatrixFree(MPI_Comm comm, INT m, INT M, MV_FUNC mv_func,
                            void *mv_data, ...)
{
    MAT *mat = phgCalloc(1, sizeof(*mat));
    mat->rmap = mat->cmap = phgMapCreateSimpleMap(comm, m, M);
    mat->rmap->refcount++;
    mat->type = PHG_MATRIX_FREE;
    mat->mv_func = mv_func;
    mat->mv_a = 0.0;
    mat->mv_b = 1.0;
    /* mv_data and the further pointers, NULL-terminated (matvec.c:1429) */
    int nd = 0;
    va_list ap;
    if (mv_data != NULL) {
        nd = 1;
        va_start(ap, mv_data);
        while (va_arg(ap, void *) != NULL)
            nd++;
        va_end(ap);
    }
    mat->mv_data = phgCalloc((size_t)nd + 1, sizeof(void *));
    if (nd > 0) {
        mat->mv_data[0] = mv_data;
        va_start(ap, mv_data);
        for (int i = 1; i < nd; i++)
            mat->mv_data[i] = va_arg(ap, void *);
        va_end(ap);
    }
    mat->localsize = m;
    mat->mode = PHG_ADD;
    return mat;
}

/* matvec.c:2292: a reference count > 0 only drops a reference; matrices
 * made here free their maps and mv_data, the arrays of a host matrix stay
 * with whoever filled them */
void phgMatDestroy(MAT **mat_ptr)
{
    if (mat_ptr == NULL || *mat_ptr == NULL)
        return;
    MAT *mat = *mat_ptr;
    *mat_ptr = NULL;
    if (mat->refcount > 0) {
        mat->refcount--;
        return;
    }
    if (mat->type == PHG_MATRIX_FREE) {
        phgMapDestroy(&mat->rmap);
        phgMapDestroy(&mat->cmap);
        phgFree(mat->mv_data);
        phgFree(mat);
    }
}

/*---------------------------------------------------------- solver.c roles */

SOLVER *phgSolverMat2Solver(OEM_SOLVER *oem_solver, MAT *mat)
{
    if (oem_solver == NULL)
        phgError(1, "phgSolverMat2Solver: no default solver list here, pass a vtable\n");
    if (oem_solver->parallel == FALSE && mat->rmap->nprocs > 1)
        phgError(1, "solver \"%s\" can't be used for parallel systems.\n",
                 oem_solver->name);
    SOLVER *solver = phgCalloc(1, sizeof(*solver));
    solver->oem_solver = oem_solver;
    solver->symmetry = M_UNSYM;
    mat->refcount++;                 /* solver.c:879 */
    solver->mat = mat;
    solver->rhs = phgMapCreateVec(mat->rmap, 1);
    solver->rhs->mat = mat;
    solver->rtol = 1e-10;            /* solver.c:164-168 option defaults */
    solver->atol = 0.0;
    solver->btol = 1e-10;
    solver->maxit = 10000;
    solver->warn_maxit = TRUE;
    if (oem_solver->Init != NULL)
        oem_solver->Init(solver);
    return solver;
}

int phgSolverDestroy(SOLVER **solver_ptr)
{
    SOLVER *solver = *solver_ptr;
    if (solver == NULL)
        return 0;
    phgFree(solver->D);
    if (solver->St != NULL)
        phgSolverDestroy(&solver->St);
    solver->oem_solver->Destroy(solver);
    phgMatDestroy(&solver->mat_bak);
    phgMatDestroy(&solver->mat);
    phgVecDestroy(&solver->rhs);
    if (solver->pc != NULL) {
        phgSolverDestroy(&solver->pc->solver_owned);
        phgFree(solver->pc);
    }
    phgFree(solver);
    *solver_ptr = NULL;
    return 0;
}

/* solver.c:223 send_matrix_to_oem_solver */
static void send_matrix(SOLVER *solver)
{
    MAT *mat = solver->mat;
    if (solver->sent_to_oem)
        return;
    if (!solver->oem_created) {
        solver->oem_created = TRUE;
        solver->oem_solver->Create(solver);
    }
    if (solver->oem_solver->AddMatrixEntries == NULL)
        return;
    if (solver->oem_solver->allow_matrix_free &&
        ((solver->matrix_free == 0 && mat->type == PHG_MATRIX_FREE) ||
         solver->matrix_free == 2))
        return;
    if (mat->type == PHG_MATRIX_FREE)
        phgError(1, "solver \"%s\" needs the matrix entries\n", solver->oem_solver->name);
    INT j = mat->rmap->partition[mat->rmap->rank];
    for (INT i = 0; i < mat->rmap->nlocal; i++, j++) {
        INT ncols;
        INT *cols;
        FLOAT *data;
        if (mat->type == PHG_PACKED) {
            ncols = (INT)(mat->packed_ind[i + 1] - mat->packed_ind[i]);
            cols = mat->packed_cols + mat->packed_ind[i];
            data = mat->packed_data + mat->packed_ind[i];
        }
        else {
            ncols = mat->rows[i].ncols;
            cols = mat->rows[i].cols;
            data = mat->rows[i].data;
        }
        if (ncols <= 0)
            phgError(1, "send_matrix_to_oem_solver: matrix row %d is empty!\n", i);
        solver->oem_solver->AddMatrixEntries(solver, 1, &j, ncols, cols, data);
    }
    solver->sent_to_oem = TRUE;
}

int phgSolverAssemble(SOLVER *solver)
{
    if (solver->assembled)
        return 0;
    solver->assembled = TRUE;
    if (!solver->mat->assembled)
        phgError(1, "phgSolverAssemble: the matrix must come assembled (phgMatAssemble "
                 "is PHG's)\n");
    send_matrix(solver);
    return solver->oem_solver->Assemble == NULL ? 0 : solver->oem_solver->Assemble(solver);
}

/* PHG's update_rhs (solver.c:508) lifts the right-hand side with the
 * boundary columns phgMatAssemble saved in the matrix; the matrices handed
 * to this driver carry no such rows, so the caller provides the lifted
 * right-hand side and it is only marked as updated here */
void phgSolverUpdateRHS(SOLVER *solver) { solver->rhs_updated = TRUE; }

int phgSolverVecSolve(SOLVER *solver, BOOLEAN destroy, VEC *x)
{
    if (solver->mat->cmap->nlocal != x->map->nlocal)
        phgError(1, "(%s:%d) inconsistent RHS!\n", __FILE__, __LINE__);
    phgSolverAssemble(solver);
    solver->rhs->assembled = TRUE;
    if (!solver->rhs_updated)
        phgSolverUpdateRHS(solver);
    if (solver->oem_solver->AddRHSEntries != NULL) {
        INT j = solver->mat->rmap->partition[solver->mat->rmap->rank];
        FLOAT *rhs = solver->rhs->data;
        for (INT i = 0; i < solver->mat->rmap->nlocal; i++, j++)
            solver->oem_solver->AddRHSEntries(solver, 1, &j, rhs++);
    }
    int ret = solver->oem_solver->Solve(solver, x, destroy);
    if (ret < 0)
        phgError(1, "error occured in OEM solver \"%s\".\n", solver->oem_solver->name);
    if (solver->warn_maxit && ret >= solver->maxit && solver->rtol > 0.0 &&
        solver->oem_solver->iterative)
        phgWarning("maxit attained in OEM solver, the result may be inaccurate.\n");
    return ret;
}

void phgSolverSetMaxIt(SOLVER *solver, int maxit) { solver->maxit = maxit; }

void phgSolverSetTol(SOLVER *solver, FLOAT user_tol) { solver->rtol = user_tol; }

/* solver-phg.c:381 phgSolverSetPC_ */
void phgSolverSetPC(SOLVER *solver, void *ctx, PC_PROC pc_func)
{
    if (solver->pc == NULL)
        solver->pc = phgCalloc(1, sizeof(SOLVER_PC));
    solver->pc->ctx = ctx;
    solver->pc->pc_proc = pc_func;
    if (pc_func != NULL && solver->oem_solver->SetPC != NULL)
        solver->oem_solver->SetPC(solver);
}
