/* Oracle: CPU restatement of PHG's finite-element hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product never links or calls it.
 *
 * Every function follows the reference file:line given beside it, with the
 * same floating-point operation order (compile with -ffp-contract=off), so
 * element matrices, CSR storage order and exact-zero drops are reproduced bit
 * for bit.  PCG reductions are sequential as in matvec.c:1193.
 */
#ifndef PHG_ORACLE_H
#define PHG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* quad.c:2914 phgQuadGetQuad3D with the fill-down table of quad.c:2807;
 * orders 0..6.  Returns npoints; *pts = [np][4] barycentric, *wts = [np]. */
int orc_quad3d(int order, const double **pts, const double **wts);

/* lagrange.c:80-731 P1/P2/P3 basis values (full-block path) and lambda
 * gradients [nbas][4]. */
int orc_nbas(int p);
void orc_bas(int p, const double *lam, double *val);
void orc_grad_lambda(int p, const double *lam, double *grad);

/* geom.c:103 compute_jacobian: J[4][4], row a = (dlambda_a/dx, /dy, /dz, c_a) */
void orc_jacobian(const double x[4][3], double *J);
/* geom.c:46 _phg_compute_elem_data volume */
double orc_volume(const double x[4][3]);

/* Problem description shared by the element routines below. */
typedef struct {
    int p;                 /* Lagrange order 1..3 */
    int dim;               /* 1: scalar Poisson, 3: linear elasticity */
    int use_kappa;         /* scalar: stiffness with P0 coefficient
                              (phgQuadGradBasAGradBas) instead of A=NULL */
    double mass_coef;      /* scalar: + mass_coef * phgQuadBasDotBas (0: off) */
    int rhs_order;         /* quadrature order of the load; <0: 2*p
                              (phgQuadDofTimesBas QUAD_DEFAULT) */
    double E, nu;          /* elasticity (elastic.c:39) */
    int qcache;            /* scalar Laplacian through the quadrature cache
                              (examples/poisson.c:111-166 -qcache) */
} orc_problem;

/* Element matrix K[N*dim][N*dim] and load b[N*dim] for one element, as the
 * drivers compute them before Dirichlet rows are applied:
 *   scalar: examples/poisson.c:124-140 (+ simplest.c:66-73 mass),
 *           load quad.c:5481 phgQuadDofTimesBas on the nodal interpolant,
 *   elasticity: examples/elastic.c:90-196 (load negated, elastic.c:212).
 * f_loc holds the load DOF values of the element's bases [N][dim]. */
void orc_element(const orc_problem *pb, const double x[4][3], double kappa,
                 const double *f_loc, double *K, double *b);

/* Assembled system with PHG storage semantics. */
typedef struct {
    int64_t n;             /* rows */
    int64_t nnz;
    int64_t *rowptr;       /* [n+1] */
    int32_t *cols;         /* [nnz] insertion order (matvec.c:2815) */
    double *vals;
    double *rhs;           /* [n] */
} orc_csr;

/* Whole assembly: element loop in index order, PHG row insertion with
 * exact-zero drop (matvec.c:5348, :3045), Dirichlet rows e_i / g
 * (dof-utils.c:1369), RHS accumulation (matvec.c:5661).
 * map[nelem][N]: scalar DOF of each local basis; bdry[ndof_scalar];
 * g[ndof_scalar*dim], f[ndof_scalar*dim]; kappa[nelem] or NULL. */
orc_csr *orc_assemble(const orc_problem *pb, int64_t nelem, const double *verts,
                      const int32_t *elem_verts, const int32_t *map,
                      int64_t ndof, const uint8_t *bdry, const double *g,
                      const double *f, const double *kappa);

/* Dirichlet elimination, serial path of phgMatAssemble (matvec.c:4430
 * save_bdry_columns(TRUE)) followed by solver.c:508 update_rhs.
 * Modifies A in place.  bdry is per row. */
void orc_eliminate(orc_csr *A, const uint8_t *bdry_rows);

/* solver-phg.c:485 solver_pcg with jacobi_pc (solver-phg.c:207) and
 * tol = max(rtol*|r0|, btol*|b|, atol) (solver-phg.c:104).  x is the
 * initial guess on input.  Returns nits (<0: zero diagonal). */
int orc_pcg(const orc_csr *A, const double *b, double *x, int maxit,
            double rtol, double btol, double atol, double *residual);

/* solver-phg.c:607 solver_BiCGstab with jacobi_pc, same stopping rule;
 * returns nits (<0: zero diagonal); b = 0 gives x = 0, residual 1 (the
 * reference leaves the caller's initial value). */
int orc_bicgstab(const orc_csr *A, const double *b, double *x, int maxit,
                 double rtol, double btol, double atol, double *residual);

/* solver-phg.c:59-77 gmres_funcs with jacobi_pc: type 0 gmres_l, 1 gmres_r,
 * 2 fgmres_r, 3 lgmres_l, 4 lgmres_r, 5 flgmres_r (PHG's default).
 * restart <= 0 means maxit + 1; augk (augmented types only) <= 0 means 2.
 * Right variants report the Givens estimate once it reached the tolerance,
 * left ones the last true residual, both |r0| when no check was made.
 * Returns nits (-1: zero diagonal, -2: bad type). */
int orc_gmres(int type, const orc_csr *A, const double *b, double *x, int restart,
              int augk, int maxit, double rtol, double btol, double atol,
              double *residual);

/* y = A x (matvec.c:887-930, row sums in storage order). */
void orc_matvec(const orc_csr *A, const double *x, double *y);

void orc_csr_free(orc_csr *A);

/* Adjoint of J = 1/2 (u - ud)^T M (u - ud) for scalar Pk problems (no
 * reference counterpart).  orc_mass_apply: Me = M e accumulated in element
 * order, returns 1/2 e^T M e.  orc_adjoint_grad: grad[e] =
 * dJ/dlog(kappa_e) = -kappa_e lambda_e^T K1_e u_e (K1_e: stiffness with
 * kappa = 1). */
double orc_mass_apply(int p, int64_t nelem, const double *verts,
                      const int32_t *elem_verts, const int32_t *map,
                      int64_t ndof, const double *e, double *Me);
void orc_adjoint_grad(int p, int64_t nelem, const double *verts,
                      const int32_t *elem_verts, const int32_t *map,
                      const double *kappa, const double *u,
                      const double *lam, double *grad);

#ifdef __cplusplus
}
#endif
#endif
