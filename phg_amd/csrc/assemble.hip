// Element integration and PHG-order CSR assembly for gfx950.
//
// Compiled with -ffp-contract=off: every floating-point expression below is
// evaluated in the order of the reference source it cites, so the element
// matrices are bit-identical to PHG's and the exact-zero drop of
// matvec.c:5460 removes the same entries.
//
// Kernels
//   k_element<P,DIM>   one work-item per tetrahedron: Jacobian (geom.c:103)
//                      and volume (geom.c:46) in registers, reference-basis
//                      tables (lagrange.c) and quadrature rules (quad.c) staged
//                      in LDS once per work-group, element matrix and load
//                      written to HBM.
//   k_inc_*            element-row incidence (row -> (element, local index)
//                      sorted by element index), built with a counting sort.
//   k_row_walk_hash    a group of 16/32/64 lanes per row replays PHG's
//                      row-list insertion (matvec.c:3045 add_columns_to_row:
//                      linear search, add or append, zeros skipped) over the
//                      row's incident elements in element order, plus the RHS
//                      accumulation of matvec.c:5661, with the row's columns
//                      in an LDS hash table.  Dirichlet rows, and rows with
//                      more distinct columns than the table holds, are
//                      replayed serially.
//   k_elim_*           serial path of phgMatAssemble (matvec.c:4430
//                      save_bdry_columns(TRUE)) and solver.c:508 update_rhs.
#include "phg_hip_internal.h"

#include <stdio.h>
#include <string.h>

#include <atomic>
#include <vector>

#pragma clang fp contract(off)

namespace {

constexpr int NQ = 24;   // largest rule used by P1-P3 (order 6)
constexpr int NB = 20;   // P3
constexpr int NQ8 = 46;  // order-8 rule (P4 mass and load)

//--------------------------------------------------------------- quadrature
// quad.c:1799-2000 static 3D rules with the orbit order of
// include/phg/quad-permu.h:83-120; order 0 uses the order-1 rule
// (quad.c:2807 init_table).
template <int CAP>
struct RuleT {
    static constexpr int cap = CAP;
    int np;
    double pts[CAP][4];
    double w[CAP];
};
using Rule = RuleT<NQ>;
using Rule8 = RuleT<NQ8>;

template <typename Rule>
__host__ __device__ inline void r4(Rule &q, double w)
{
    q.pts[q.np][0] = q.pts[q.np][1] = q.pts[q.np][2] = q.pts[q.np][3] = .25;
    q.w[q.np++] = w;
}

template <typename Rule>
__host__ __device__ inline void r31(Rule &q, double a, double w)
{
    double b = 1.0 - 3.0 * a;
    double P[4][4] = {{b, a, a, a}, {a, b, a, a}, {a, a, b, a}, {a, a, a, b}};
    for (int i = 0; i < 4; i++) {
        for (int c = 0; c < 4; c++)
            q.pts[q.np][c] = P[i][c];
        q.w[q.np++] = w;
    }
}

template <typename Rule>
__host__ __device__ inline void r22(Rule &q, double a, double w)
{
    double h = 0.5 - a;
    double P[6][4] = {{a, a, h, h}, {a, h, a, h}, {a, h, h, a},
                      {h, a, h, a}, {h, a, a, h}, {h, h, a, a}};
    for (int i = 0; i < 6; i++) {
        for (int c = 0; c < 4; c++)
            q.pts[q.np][c] = P[i][c];
        q.w[q.np++] = w;
    }
}

template <typename Rule>
__host__ __device__ inline void r211(Rule &q, double a, double b, double w)
{
    double c = 1.0 - a - a - b;
    double P[12][4] = {{a, a, b, c}, {a, a, c, b}, {a, b, a, c}, {a, b, c, a},
                       {a, c, a, b}, {a, c, b, a}, {b, a, a, c}, {b, a, c, a},
                       {b, c, a, a}, {c, a, a, b}, {c, a, b, a}, {c, b, a, a}};
    for (int i = 0; i < 12; i++) {
        for (int k = 0; k < 4; k++)
            q.pts[q.np][k] = P[i][k];
        q.w[q.np++] = w;
    }
}

// orders above 6 need a rule of capacity NQ8
template <typename Rule>
__host__ __device__ inline void make_rule(int order, Rule &q)
{
    q.np = 0;
    if (order <= 1) {
        q.pts[0][0] = q.pts[0][1] = q.pts[0][2] = q.pts[0][3] = .25;
        q.w[0] = 1.;
        q.np = 1;
    }
    else if (order == 2) {
        r31(q, .13819660112501051517954131656343619, .25);
    }
    else if (order == 3) {
        r31(q, .32805469671142664733580581998119743,
            .13852796651186214232361769837564129);
        r31(q, .10695227393293068277170204157061650,
            .11147203348813785767638230162435871);
    }
    else if (order == 4) {
        r31(q, .09273525031089122628655892066032137,
            .07349304311636194934358694586367885);
        r31(q, .31088591926330060975814749494040332,
            .11268792571801585036501492847638892);
        r22(q, .04550370412564965000000000000000000,
            .04254602077708146686093208377328816);
    }
    else if (order == 5) {
        r31(q, .31088591926330060979734573376345783,
            .11268792571801585079918565233328633);
        r31(q, .09273525031089122640232391373703061,
            .07349304311636194954371020548632750);
        r22(q, .04550370412564964949188052627933943,
            .04254602077708146643806942812025744);
    }
    else if (order == 6 || Rule::cap < NQ8) {
        r31(q, .21460287125915202928883921938628499,
            .03992275025816749209969062755747998);
        r31(q, .04067395853461135311557944895641006,
            .01007721105532064294801323744593686);
        r31(q, .32233789014227551034399447076249213,
            .05535718154365472209515327785372602);
        r211(q, .06366100187501752529923552760572698,
             .60300566479164914136743113906093969, 27. / 560.);
    }
    else if (order == 7) {      // quad.c:1990
        r4(q, .09548528946413084886057843611722638);
        r31(q, .31570114977820279942342999959331149,
            .04232958120996702907628617079854674);
        r22(q, .05048982259839636876305382298656247,
            .03189692783285757993427482408294246);
        r211(q, .18883383102600104773643110385458576,
             .57517163758700002348324157702230752, .03720713072833462136961556119148112);
        r211(q, .02126547254148324598883610149981994,
             .81083024109854856111810537984823239, .00811077082990334156610343349109654);
    }
    else {                      // quad.c:2017, order 8
        r31(q, .03967542307038990126507132953938949,
            .00639714777990232132145142033517302);
        r31(q, .31448780069809631378416056269714830,
            .04019044802096617248816115847981783);
        r31(q, .10198669306270330000000000000000000,
            .02430797550477032117486910877192260);
        r31(q, .18420369694919151227594641734890918,
            .05485889241369744046692412399039144);
        r22(q, .06343628775453989240514123870189827,
            .03571961223409918246495096899661762);
        r211(q, .02169016206772800480266248262493018,
             .71993192203946593588943495335273478, .00718319069785253940945110521980376);
        r211(q, .20448008063679571424133557487274534,
             .58057719012880922417539817139062041, .01637218194531911754093813975611913);
    }
}

//------------------------------------------------------------------ bases
__device__ __constant__ int c_ev[6][2] = {{0, 1}, {0, 2}, {0, 3},
                                          {1, 2}, {1, 3}, {2, 3}};
__device__ __constant__ int c_fv[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3},
                                          {0, 1, 2}};
// the same tables for code that also runs on the host
__host__ __device__ inline void edge_verts(int k, int &a, int &b)
{
    constexpr int t[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    a = t[k][0];
    b = t[k][1];
}
__host__ __device__ inline void face_verts(int k, int &a, int &b, int &c)
{
    constexpr int t[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    a = t[k][0];
    b = t[k][1];
    c = t[k][2];
}

// lagrange.c:80 / :160 / :392 / :757 (full-block evaluation)
__host__ __device__ inline void bas_values(int p, const double *lam, double *v)
{
    double L0 = lam[0], L1 = lam[1], L2 = lam[2], L3 = lam[3];
    if (p == 4) {
        double L0T4M1 = 4. * L0 - 1., L1T4M1 = 4. * L1 - 1., L2T4M1 = 4. * L2 - 1.,
               L3T4M1 = 4. * L3 - 1.;
        v[0] = L0T4M1 * ((2. / 3.) * L0 - (1. / 3.)) * (L0T4M1 - 2.) * L0;
        v[1] = L1T4M1 * ((2. / 3.) * L1 - (1. / 3.)) * (L1T4M1 - 2.) * L1;
        v[2] = L2T4M1 * ((2. / 3.) * L2 - (1. / 3.)) * (L2T4M1 - 2.) * L2;
        v[3] = L3T4M1 * ((2. / 3.) * L3 - (1. / 3.)) * (L3T4M1 - 2.) * L3;
        const double t[4] = {L0T4M1 * ((32. / 3.) * L0 - (16. / 3.)),
                             L1T4M1 * ((32. / 3.) * L1 - (16. / 3.)),
                             L2T4M1 * ((32. / 3.) * L2 - (16. / 3.)),
                             L3T4M1 * ((32. / 3.) * L3 - (16. / 3.))};
        const double m[4] = {L0T4M1, L1T4M1, L2T4M1, L3T4M1};
        // edges: near the first vertex, middle, near the second
        for (int k = 0; k < 6; k++) {
            int a, b;
            edge_verts(k, a, b);
            const double d = lam[a] * lam[b];
            v[4 + 3 * k] = t[a] * d;
            v[5 + 3 * k] = 4. * m[a] * m[b] * d;
            v[6 + 3 * k] = t[b] * d;
        }
        const double m32[4] = {L0T4M1 * 32., L1T4M1 * 32., L2T4M1 * 32., L3T4M1 * 32.};
        const double d23 = L2 * L3, d13 = L1 * L3, d12 = L1 * L2;
        const double df[4] = {L1 * d23, L0 * d23, L0 * d13, L0 * d12};
        // faces: one node near each face vertex in face-vertex order
        for (int k = 0; k < 4; k++) {
            int a, b, c;
            face_verts(k, a, b, c);
            v[22 + 3 * k] = m32[a] * df[k];
            v[23 + 3 * k] = m32[b] * df[k];
            v[24 + 3 * k] = m32[c] * df[k];
        }
        v[34] = 256. * L0 * df[0];
    }
    else if (p == 1) {
        v[0] = L0; v[1] = L1; v[2] = L2; v[3] = L3;
    }
    else if (p == 2) {
        v[0] = L0 * (L0 + L0 - 1.);
        v[1] = L1 * (L1 + L1 - 1.);
        v[2] = L2 * (L2 + L2 - 1.);
        v[3] = L3 * (L3 + L3 - 1.);
        v[4] = 4. * L0 * L1;
        v[5] = 4. * L0 * L2;
        v[6] = 4. * L0 * L3;
        v[7] = 4. * L1 * L2;
        v[8] = 4. * L1 * L3;
        v[9] = 4. * L2 * L3;
    }
    else {
        double L03 = 3. * L0, L13 = 3. * L1, L23 = 3. * L2, L33 = 3. * L3, d;
        v[0] = .5 * (L03 - 1.) * (L03 - 2.) * L0;
        v[1] = .5 * (L13 - 1.) * (L13 - 2.) * L1;
        v[2] = .5 * (L23 - 1.) * (L23 - 2.) * L2;
        v[3] = .5 * (L33 - 1.) * (L33 - 2.) * L3;
        d = L0 * L1;
        v[4] = 4.5 * (L03 - 1.) * d;
        v[5] = 4.5 * (L13 - 1.) * d;
        d = L0 * L2;
        v[6] = 4.5 * (L03 - 1.) * d;
        v[7] = 4.5 * (L23 - 1.) * d;
        d = L0 * L3;
        v[8] = 4.5 * (L03 - 1.) * d;
        v[9] = 4.5 * (L33 - 1.) * d;
        d = L1 * L2;
        v[10] = 4.5 * (L13 - 1.) * d;
        v[11] = 4.5 * (L23 - 1.) * d;
        d = L1 * L3;
        v[12] = 4.5 * (L13 - 1.) * d;
        v[13] = 4.5 * (L33 - 1.) * d;
        d = L2 * L3;
        v[14] = 4.5 * (L23 - 1.) * d;
        v[15] = 4.5 * (L33 - 1.) * d;
        v[16] = 27. * L1 * L2 * L3;
        v[17] = 27. * L0 * L2 * L3;
        v[18] = 27. * L0 * L1 * L3;
        v[19] = 27. * L0 * L1 * L2;
    }
}

// lagrange.c:111 / :248 / :538 / :1085: gradients with respect to lambda,
// [nb][4]
__host__ __device__ inline void bas_grad_lambda(int p, const double *lam, double *g)
{
    int nb = p == 1 ? 4 : (p == 2 ? 10 : (p == 3 ? 20 : 35));
    for (int i = 0; i < 4 * nb; i++)
        g[i] = 0.;
    if (p == 4) {
        for (int k = 0; k < 4; k++) {
            const double L0 = lam[k];
            g[4 * k + k] = ((8. / 3.) * L0 - 1.) * (L0 * (16. * L0 - 12.) + 1.);
        }
        for (int k = 0; k < 6; k++) {
            int a, b;
            edge_verts(k, a, b);
            for (int t = 0; t < 3; t++) {
                // TEMP0 (a, b), TEMP1 (a, b), TEMP0 (b, a)
                const int v0 = t == 2 ? b : a, v1 = t == 2 ? a : b;
                const double L0 = lam[v0], L1 = lam[v1];
                double *gg = g + 4 * (4 + 3 * k + t);
                if (t == 1) {
                    gg[v0] = (32. * L0 - 4.) * (4. * L1 - 1.) * L1;
                    gg[v1] = L0 * (4. * L0 - 1.) * (32. * L1 - 4.);
                }
                else {
                    gg[v0] = (L0 * ((384. / 3.) * L0 - 64.) + (16. / 3.)) * L1;
                    gg[v1] = L0 * ((64. / 3.) * L0 - (16. / 3.)) * (2. * L0 - 1.);
                }
            }
        }
        for (int k = 0; k < 4; k++) {
            int a, b, c;
            face_verts(k, a, b, c);
            const int cyc[3][3] = {{a, b, c}, {b, c, a}, {c, a, b}};
            for (int t = 0; t < 3; t++) {
                const int v1 = cyc[t][0], v2 = cyc[t][1], v3 = cyc[t][2];
                const double L1 = lam[v1], L2 = lam[v2], L3 = lam[v3];
                double *gg = g + 4 * (22 + 3 * k + t);
                gg[v1] = (256. * L1 - 32.) * L2 * L3;
                gg[v2] = (128. * L1 - 32.) * L1 * L3;
                gg[v3] = (128. * L1 - 32.) * L1 * L2;
            }
        }
        g[4 * 34 + 0] = 256. * lam[1] * lam[2] * lam[3];
        g[4 * 34 + 1] = 256. * lam[0] * lam[2] * lam[3];
        g[4 * 34 + 2] = 256. * lam[0] * lam[1] * lam[3];
        g[4 * 34 + 3] = 256. * lam[0] * lam[1] * lam[2];
        return;
    }
    if (p == 1) {
        for (int k = 0; k < 4; k++)
            g[4 * k + k] = 1.0;
        return;
    }
    if (p == 2) {
        for (int k = 0; k < 4; k++)
            g[4 * k + k] = 4. * lam[k] - 1.;
        for (int k = 0; k < 6; k++) {
            int a, b;
            edge_verts(k, a, b);
            g[4 * (4 + k) + a] = 4. * lam[b];
            g[4 * (4 + k) + b] = 4. * lam[a];
        }
        return;
    }
    for (int k = 0; k < 4; k++)
        g[4 * k + k] = (13.5 * lam[k] - 9.) * lam[k] + 1.;
    for (int k = 0; k < 6; k++) {
        int a, b;
        edge_verts(k, a, b);
        double *g0 = g + 4 * (4 + 2 * k), *g1 = g0 + 4;
        g0[a] = (27. * lam[a] - 4.5) * lam[b];
        g0[b] = 4.5 * lam[a] * (3. * lam[a] - 1.);
        g1[a] = 4.5 * lam[b] * (3. * lam[b] - 1.);
        g1[b] = (27. * lam[b] - 4.5) * lam[a];
    }
    for (int k = 0; k < 4; k++) {
        int a, b, c;
        face_verts(k, a, b, c);
        double *gg = g + 4 * (16 + k);
        gg[a] = 27. * lam[b] * lam[c];
        gg[b] = 27. * lam[a] * lam[c];
        gg[c] = 27. * lam[a] * lam[b];
    }
}

//--------------------------------------------------------------- geometry
// geom.c:103 compute_jacobian, Gauss-Jordan with partial pivoting
__device__ inline void jacobian(const double x[4][3], double *J)
{
    const int n = 4;
    double d, t, a[4][8];
#pragma unroll
    for (int j = 0; j < n; j++)
#pragma unroll
        for (int i = 0; i < n; i++)
            a[j][n + i] = (i == j) ? 1.0 : 0.0;
#pragma unroll
    for (int j = 0; j < n; j++) {
        a[j][0] = x[j][0];
        a[j][1] = x[j][1];
        a[j][2] = x[j][2];
        a[j][3] = 1.0;
    }
#pragma unroll
    for (int j = 0; j < n; j++) {
        d = fabs(a[j][j]);
        int k = j;
#pragma unroll
        for (int i = j + 1; i < n; i++)
            if (d < (t = fabs(a[i][j]))) {
                d = t;
                k = i;
            }
        // row swap written as selects so the matrix stays in registers
#pragma unroll
        for (int i = j + 1; i < n; i++) {
            if (i == k) {
#pragma unroll
                for (int c = 0; c < 2 * n; c++) {
                    if (c < j)
                        continue;
                    t = a[j][c];
                    a[j][c] = a[i][c];
                    a[i][c] = t;
                }
            }
        }
        d = 1.0 / ((t = a[j][j]) == 0. ? 1. : t);
#pragma unroll
        for (int i = j + 1; i < 2 * n; i++)
            a[j][i] *= d;
#pragma unroll
        for (int i = j + 1; i < n; i++) {
            d = a[i][j];
#pragma unroll
            for (int c = j + 1; c < 2 * n; c++)
                a[i][c] -= d * a[j][c];
        }
    }
#pragma unroll
    for (int j = n - 2; j >= 0; j--)
#pragma unroll
        for (int i = j + 1; i < n; i++) {
            d = a[j][i];
#pragma unroll
            for (int c = n; c < 2 * n; c++)
                a[j][c] -= d * a[i][c];
        }
#pragma unroll
    for (int j = 0; j < n; j++)
#pragma unroll
        for (int i = 0; i < n; i++)
            J[4 * j + i] = a[i][j + n];
}

// geom.c:46 _phg_compute_elem_data (volume)
__device__ inline double volume(const double x[4][3])
{
    const double oned6 = 1.0 / (double)6.0;
    double x0 = x[0][0], y0 = x[0][1], z0 = x[0][2];
    double x1 = x[1][0] - x0, y1 = x[1][1] - y0, z1 = x[1][2] - z0;
    double x2 = x[2][0] - x0, y2 = x[2][1] - y0, z2 = x[2][2] - z0;
    double x3 = x[3][0] - x0, y3 = x[3][1] - y0, z3 = x[3][2] - z0;
    return fabs(x1 * y2 * z3 + x2 * y3 * z1 + y1 * z2 * x3 -
                (z1 * y2 * x3 + y1 * x2 * z3 + z2 * y3 * x1)) * oned6;
}

// quad.c:3314 phgQuadGetBasisGradient for basis k with lambda gradient G
template <int P>
__device__ inline void xgrad(int k, const double *G, const double *J,
                             double *o)
{
    if (P == 1) {
        o[0] = J[4 * k + 0];
        o[1] = J[4 * k + 1];
        o[2] = J[4 * k + 2];
    }
    else if (k < 4) {
        double d = G[k];
        o[0] = d * J[4 * k + 0];
        o[1] = d * J[4 * k + 1];
        o[2] = d * J[4 * k + 2];
    }
    else {
#pragma unroll
        for (int c = 0; c < 3; c++)
            o[c] = G[0] * J[0 * 4 + c] + G[1] * J[1 * 4 + c] +
                   G[2] * J[2 * 4 + c] + G[3] * J[3 * 4 + c];
    }
}

// diagonal of each row (matvec.c:500 phgMatSetupDiagonal: linear search;
// 0 when the row has no diagonal entry)
__global__ void k_diag(int64_t n, const int32_t *rowptr, const int32_t *cols,
                       const double *vals, double *diag)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    double d = 0.0;
    for (int32_t j = rowptr[r]; j < rowptr[r + 1]; j++)
        if (cols[j] == r) {
            d = vals[j];
            break;
        }
    diag[r] = d;
}

//--------------------------------------------------------- element kernel
struct ElemArgs {
    int64_t nelem;
    const double *verts;
    const int32_t *elem_verts;
    const int32_t *map;      // [nelem][N]
    const double *kappa;     // [nelem] or null
    const double *f;         // [ndof*dim]
    int use_kappa;
    double mass_coef;
    int rhs_order;
    double E, nu;
    double *K;               // [nelem][M][M]
    double *B;               // [nelem][M]
};

// table sizes per order: P1-P3 share one layout (rules up to order 6, 20
// bases); P4 has 35 bases, stiffness gradients on the order-6 rule and mass
// and load values on rules up to order 8
template <int P> struct Shape {
    static constexpr int N = P == 1 ? 4 : (P == 2 ? 10 : (P == 3 ? 20 : 35));
    static constexpr int NBT = P <= 3 ? NB : 35;
    static constexpr int NQG = NQ;                     // stiffness points
    static constexpr int NQT = P <= 3 ? NQ : NQ8;      // mass and load points
};

template <int P>
struct TablesT {
    int nq_s, nq_m, nq_r;
    double ws[Shape<P>::NQG], wm[Shape<P>::NQT], wr[Shape<P>::NQT];
    double gs[Shape<P>::NQG][Shape<P>::NBT][4];   // lambda gradients at the stiffness points
    double bm[Shape<P>::NQT][Shape<P>::NBT];      // basis values at the mass points
    double br[Shape<P>::NQT][Shape<P>::NBT];      // basis values at the load points
};

template <int P>
__device__ void build_tables(TablesT<P> &T, int dim, int rhs_order, int mass, bool qc)
{
    constexpr int N = Shape<P>::N;
    // all work-items build the same rules; thread 0 publishes them
    if (threadIdx.x == 0) {
        RuleT<Shape<P>::NQT> q;
        // stiffness (quad.c:5271; elastic.c:141); with the quadrature cache
        // the element's one rule of order 2p (poisson.c:126; P1-P3 only, the
        // host refuses it for P4)
        make_rule(qc ? 2 * P : 2 * (P - 1), q);
        T.nq_s = q.np < Shape<P>::NQG ? q.np : Shape<P>::NQG;
        for (int i = 0; i < T.nq_s; i++) {
            T.ws[i] = q.w[i];
            double g[Shape<P>::NBT * 4];
            bas_grad_lambda(P, q.pts[i], g);
            for (int k = 0; k < N; k++)
                for (int c = 0; c < 4; c++)
                    T.gs[i][k][c] = g[4 * k + c];
        }
        T.nq_m = 0;
        if (mass) {
            make_rule(2 * P, q);     // quad.c:5582 order p+p
            T.nq_m = q.np;
            for (int i = 0; i < q.np; i++) {
                T.wm[i] = q.w[i];
                bas_values(P, q.pts[i], T.bm[i]);
            }
        }
        make_rule(rhs_order < 0 ? 2 * P : rhs_order, q);   // quad.c:5495
        T.nq_r = q.np;
        for (int i = 0; i < q.np; i++) {
            T.wr[i] = q.w[i];
            bas_values(P, q.pts[i], T.br[i]);
        }
    }
    (void)dim;
    __syncthreads();
}

// QC: examples/poisson.c -qcache, the integrals through quad-cache.c:1016
// phgQCIntegrate_ with the rule of order 2p for stiffness and load
template <int P, int DIM, bool QC = false>
__global__ void __launch_bounds__(128) k_element(ElemArgs a)
{
    constexpr int N = Shape<P>::N;
    constexpr int M = N * DIM;
    __shared__ TablesT<P> T;
    build_tables<P>(T, DIM, a.rhs_order, a.mass_coef != 0.0, QC);

    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.nelem)
        return;

    double x[4][3];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        int64_t iv = a.elem_verts[4 * e + v];
        x[v][0] = a.verts[3 * iv + 0];
        x[v][1] = a.verts[3 * iv + 1];
        x[v][2] = a.verts[3 * iv + 2];
    }
    double J[16];
    jacobian(x, J);
    double vol = volume(x);
    double *K = a.K + e * (int64_t)(M * M);
    double *B = a.B + e * (int64_t)M;
    const int32_t *m = a.map + e * (int64_t)N;

    // points of the stiffness rule: order 2(P-1), or 2P with the cache
    constexpr int NQS = QC ? (P == 1 ? 4 : (P == 2 ? 14 : 24))
                           : (P == 1 ? 1 : (P == 2 ? 4 : (P == 3 ? 14 : 24)));
    if (DIM == 1) {
        const double kap = a.use_kappa ? a.kappa[e] : 0.0;
        const int nq = NQS;
        for (int i = 0; i < N; i++) {
            double gi[NQS][3];
            for (int q = 0; q < nq; q++)
                xgrad<P>(i, T.gs[q][i], J, gi[q]);
            for (int j = 0; j <= i; j++) {
                // phgQuadGradBasAGradBas(e, u, j, A, u, i): g1 = grad j
                double d = 0., d0;
                for (int q = 0; q < nq; q++) {
                    double g1[3];
                    xgrad<P>(j, T.gs[q][j], J, g1);
                    d0 = 0.;
                    d0 += g1[0] * gi[q][0] + g1[1] * gi[q][1] +
                          g1[2] * gi[q][2];
                    if (a.use_kappa)
                        d += d0 * T.ws[q] * kap;
                    else
                        d += d0 * T.ws[q];
                }
                double v = d * vol;
                K[i * N + j] = v;
                K[j * N + i] = v;
            }
        }
        if (a.mass_coef != 0.0) {
            // simplest.c:70 + a * phgQuadBasDotBas (quad.c:5582)
            for (int i = 0; i < N; i++)
                for (int j = 0; j <= i; j++) {
                    double d = 0., d0;
                    for (int q = 0; q < T.nq_m; q++) {
                        d0 = 0.;
                        d0 += T.bm[q][j] * T.bm[q][i];
                        d += d0 * T.wm[q];
                    }
                    double v = K[i * N + j] + a.mass_coef * (d * vol);
                    K[i * N + j] = v;
                    K[j * N + i] = v;
                }
        }
        // quad.c:5481 phgQuadDofTimesBas with f_h at the points from
        // dof-utils.c:59 phgDofEval_ (basis order)
        double fl[N];
#pragma unroll
        for (int k = 0; k < N; k++)
            fl[k] = a.f[m[k]];
        // points outer, local indices inner: each B[i] still sums its
        // terms in point order from 0, and f_h at a point stays in a
        // register instead of a per-point array
        double d[N];
#pragma unroll
        for (int i = 0; i < N; i++)
            d[i] = 0.0;
        for (int q = 0; q < T.nq_r; q++) {
            double fq = 0.0;
#pragma unroll
            for (int k = 0; k < N; k++)
                fq += T.br[q][k] * fl[k];
#pragma unroll
            for (int i = 0; i < N; i++) {
                if (QC)
                    d[i] += T.br[q][i] * fq * T.wr[q];   // phgQCIntegrate_: (phi_i f) w
                else
                    d[i] += T.br[q][i] * T.wr[q] * fq;
            }
        }
#pragma unroll
        for (int i = 0; i < N; i++)
            B[i] = d[i] * vol;
        return;
    }

    // linear elasticity, examples/elastic.c:90 stiffness_matrix
    {
        const double E = a.E, nu = a.nu;
        const double aa = nu / (1.0 - nu);
        const double bb = (1. - 2. * nu) / (2. * (1. - nu));
        const double c = E * (1. - nu) / ((1. + nu) * (1. - 2. * nu)) * vol;
        const int nq = NQS;
        for (int i = 0; i < N; i++) {
            double gi[NQS][3];
            for (int q = 0; q < nq; q++)
                xgrad<P>(i, T.gs[q][i], J, gi[q]);
            for (int j = 0; j <= i; j++) {
                double a00, a01, a02, a10, a11, a12, a20, a21, a22;
                a00 = a01 = a02 = a10 = a11 = a12 = a20 = a21 = a22 = 0.;
                for (int q = 0; q < nq; q++) {
                    double pn[3];
                    const double *pm = gi[q];
                    xgrad<P>(j, T.gs[q][j], J, pn);
                    double w = T.ws[q];
                    a00 += (pm[0] * pn[0] + bb * (pm[1] * pn[1] + pm[2] * pn[2])) * w;
                    a01 += (aa * pm[0] * pn[1] + bb * pm[1] * pn[0]) * w;
                    a02 += (aa * pm[0] * pn[2] + bb * pm[2] * pn[0]) * w;
                    a10 += (aa * pm[1] * pn[0] + bb * pm[0] * pn[1]) * w;
                    a11 += (pm[1] * pn[1] + bb * (pm[0] * pn[0] + pm[2] * pn[2])) * w;
                    a12 += (aa * pm[1] * pn[2] + bb * pm[2] * pn[1]) * w;
                    a20 += (aa * pm[2] * pn[0] + bb * pm[0] * pn[2]) * w;
                    a21 += (aa * pm[2] * pn[1] + bb * pm[1] * pn[2]) * w;
                    a22 += (pm[2] * pn[2] + bb * (pm[0] * pn[0] + pm[1] * pn[1])) * w;
                }
                double A3[3][3] = {{a00 * c, a01 * c, a02 * c},
                                   {a10 * c, a11 * c, a12 * c},
                                   {a20 * c, a21 * c, a22 * c}};
                for (int k = 0; k < 3; k++)
                    for (int l = 0; l < 3; l++) {
                        K[(i * 3 + k) * M + j * 3 + l] = A3[k][l];
                        if (j != i)
                            K[(j * 3 + k) * M + i * 3 + l] = A3[l][k];
                    }
            }
        }
        // load: vector path of phgQuadDofTimesBas, negated (elastic.c:212)
        double fq[Shape<P>::NQT][3];
        for (int q = 0; q < T.nq_r; q++) {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            for (int k = 0; k < N; k++) {
                int64_t d = (int64_t)m[k] * 3;
                double bv = T.br[q][k];
                s0 += bv * a.f[d + 0];
                s1 += bv * a.f[d + 1];
                s2 += bv * a.f[d + 2];
            }
            fq[q][0] = s0;
            fq[q][1] = s1;
            fq[q][2] = s2;
        }
        for (int i = 0; i < N; i++) {
            double r0 = 0, r1 = 0, r2 = 0, d;
            for (int q = 0; q < T.nq_r; q++) {
                d = T.br[q][i] * T.wr[q];
                r0 += d * fq[q][0];
                r1 += d * fq[q][1];
                r2 += d * fq[q][2];
            }
            r0 *= vol;
            r1 *= vol;
            r2 *= vol;
            B[3 * i + 0] = -r0;
            B[3 * i + 1] = -r1;
            B[3 * i + 2] = -r2;
        }
    }
}

//---------------------------------------------------------------- adjoint
// No reference counterpart (PHG has no adjoint).  For J = 1/2 (u - ud)^T M
// (u - ud), with M the Pk mass matrix integrated by the rule of
// phgQuadBasDotBas (quad.c:5582, order 2p):
//   k_adj_mass<P>   per element: Me_e = M_e (u - ud)_e into [nelem][N] and
//                   the element's share 1/2 e_e . Me_e;
//   k_adj_rows      per row: the adjoint right-hand side, the element
//                   contributions summed in element order (zero on Dirichlet
//                   rows, whose adjoint value is fixed at 0);
//   k_adj_grad<P>   per element: dJ/dlog(kappa_e) = -kappa_e vol
//                   sum_q w_q grad u(x_q) . grad lambda(x_q) with the
//                   stiffness rule of phgQuadGradBasAGradBas (quad.c:5260),
//                   i.e. -kappa_e lambda_e^T K1_e u_e without forming K1_e.
struct AdjArgs {
    int64_t nelem;
    const double *verts;
    const int32_t *elem_verts;
    const int32_t *map;
    const double *kappa;
    const double *u, *ud, *lam;
    double *Me;        // [nelem][N]
    double *Jel;       // [nelem]
    double *grad;      // [nelem]
};

template <int P>
__global__ void __launch_bounds__(128) k_adj_mass(AdjArgs a)
{
    constexpr int N = Shape<P>::N;
    __shared__ TablesT<P> T;
    build_tables<P>(T, 1, -1, 1, false);
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.nelem)
        return;
    double x[4][3];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        int64_t iv = a.elem_verts[4 * e + v];
        x[v][0] = a.verts[3 * iv + 0];
        x[v][1] = a.verts[3 * iv + 1];
        x[v][2] = a.verts[3 * iv + 2];
    }
    const double vol = volume(x);
    const int32_t *m = a.map + e * (int64_t)N;
    double el[N];
#pragma unroll
    for (int k = 0; k < N; k++)
        el[k] = a.u[m[k]] - a.ud[m[k]];
    // values of e_h at the mass points, then Me_i = vol sum_q w_q phi_i e_h
    double eq[Shape<P>::NQT];
    for (int q = 0; q < T.nq_m; q++) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < N; k++)
            t += T.bm[q][k] * el[k];
        eq[q] = t;
    }
    double J = 0.0;
    for (int i = 0; i < N; i++) {
        double d = 0.0;
        for (int q = 0; q < T.nq_m; q++)
            d += T.bm[q][i] * T.wm[q] * eq[q];
        d *= vol;
        a.Me[e * N + i] = d;
        J += el[i] * d;
    }
    a.Jel[e] = 0.5 * J;
}

__global__ void k_adj_rows(int64_t n, int N, const int64_t *off,
                           const int32_t *inc, const uint8_t *bdry,
                           const double *Me, double *rhs)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    double s = 0.0;
    if (!bdry[r])
        for (int64_t t = off[r]; t < off[r + 1]; t++)
            s += Me[inc[t]];          // code e * N + a indexes [nelem][N]
    rhs[r] = s;
}

template <int P>
__global__ void __launch_bounds__(128) k_adj_grad(AdjArgs a)
{
    constexpr int N = Shape<P>::N;
    constexpr int NQS = P == 1 ? 1 : (P == 2 ? 4 : (P == 3 ? 14 : 24));
    __shared__ TablesT<P> T;
    build_tables<P>(T, 1, -1, 0, false);
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.nelem)
        return;
    double x[4][3];
#pragma unroll
    for (int v = 0; v < 4; v++) {
        int64_t iv = a.elem_verts[4 * e + v];
        x[v][0] = a.verts[3 * iv + 0];
        x[v][1] = a.verts[3 * iv + 1];
        x[v][2] = a.verts[3 * iv + 2];
    }
    double J[16];
    jacobian(x, J);
    const double vol = volume(x);
    const int32_t *m = a.map + e * (int64_t)N;
    double gu[NQS][3], gl[NQS][3];
    for (int q = 0; q < NQS; q++)
        for (int c = 0; c < 3; c++)
            gu[q][c] = gl[q][c] = 0.0;
    for (int k = 0; k < N; k++) {
        const double uk = a.u[m[k]], lk = a.lam[m[k]];
        for (int q = 0; q < NQS; q++) {
            double g[3];
            xgrad<P>(k, T.gs[q][k], J, g);
            for (int c = 0; c < 3; c++) {
                gu[q][c] += uk * g[c];
                gl[q][c] += lk * g[c];
            }
        }
    }
    double s = 0.0;
    for (int q = 0; q < NQS; q++)
        s += (gu[q][0] * gl[q][0] + gu[q][1] * gl[q][1] + gu[q][2] * gl[q][2]) *
             T.ws[q];
    a.grad[e] = -a.kappa[e] * (s * vol);
}

// dof-utils.c:2872 phgDofNormL2_ and the L2 norm of dof-utils.c:875
// phgDofGradient_ for a scalar Pk DOF e: per element
//   l2[e] = vol sum_q w_q e_h(x_q)^2        (mass rule, order 2p)
//   h1[e] = vol sum_q w_q |grad e_h(x_q)|^2 (stiffness rule, order 2p-2)
template <int P>
__global__ void __launch_bounds__(128) k_dof_norms(int64_t nelem,
                                                   const double *verts,
                                                   const int32_t *elem_verts,
                                                   const int32_t *map,
                                                   const double *v,
                                                   double *l2, double *h1)
{
    constexpr int N = Shape<P>::N;
    constexpr int NQS = P == 1 ? 1 : (P == 2 ? 4 : (P == 3 ? 14 : 24));
    __shared__ TablesT<P> T;
    build_tables<P>(T, 1, -1, 1, false);
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nelem)
        return;
    double x[4][3];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int64_t iv = elem_verts[4 * e + k];
        x[k][0] = verts[3 * iv + 0];
        x[k][1] = verts[3 * iv + 1];
        x[k][2] = verts[3 * iv + 2];
    }
    double J[16];
    jacobian(x, J);
    const double vol = volume(x);
    const int32_t *m = map + e * (int64_t)N;
    double el[N];
#pragma unroll
    for (int k = 0; k < N; k++)
        el[k] = v[m[k]];
    double a = 0.0;
    for (int q = 0; q < T.nq_m; q++) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < N; k++)
            t += T.bm[q][k] * el[k];
        a += t * t * T.wm[q];
    }
    l2[e] = a * vol;
    double b = 0.0;
    for (int q = 0; q < NQS; q++) {
        double g[3] = {0.0, 0.0, 0.0};
        for (int k = 0; k < N; k++) {
            double gk[3];
            xgrad<P>(k, T.gs[q][k], J, gk);
            g[0] += el[k] * gk[0];
            g[1] += el[k] * gk[1];
            g[2] += el[k] * gk[2];
        }
        b += (g[0] * g[0] + g[1] * g[1] + g[2] * g[2]) * T.ws[q];
    }
    h1[e] = b * vol;
}

// fixed-order sum of n values: per-block partials, then one block
__global__ void __launch_bounds__(256) k_sum_blocks(int64_t n, const double *v,
                                                    double *part)
{
    __shared__ double sh[256];
    int64_t per = (n + gridDim.x - 1) / gridDim.x;
    int64_t i0 = blockIdx.x * per, i1 = i0 + per < n ? i0 + per : n;
    double t = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += 256)
        t += v[i];
    sh[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[blockIdx.x] = sh[0];
}

//------------------------------------------------------ nodal coordinates
// barycentric node of local basis b (dof.c:4925), combined in the element's
// local vertex order as geom.c:1259 phgGeomLambda2XYZ does
__global__ void k_dof_coords(int64_t nelem, int order, int N,
                             const double *verts, const int32_t *ev,
                             const int32_t *map, double *xyz)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nelem * N)
        return;
    int64_t e = t / N;
    int b = (int)(t % N);
    double lam[4] = {0., 0., 0., 0.};
    if (b < 4)
        lam[b] = 1.0;
    else if (order == 4) {
        // lagrange.c:1671 P4_points: edge nodes from the first vertex, face
        // node t near face vertex t, the centroid
        if (b < 22) {
            const int k = (b - 4) / 3, t = (b - 4) % 3;
            lam[c_ev[k][0]] = 0.75 - 0.25 * t;
            lam[c_ev[k][1]] = 0.25 + 0.25 * t;
        }
        else if (b < 34) {
            const int f = (b - 22) / 3, t = (b - 22) % 3;
            for (int i = 0; i < 3; i++)
                lam[c_fv[f][i]] = i == t ? 0.5 : 0.25;
        }
        else
            lam[0] = lam[1] = lam[2] = lam[3] = 0.25;
    }
    else if (order == 2) {
        lam[c_ev[b - 4][0]] = 0.5;
        lam[c_ev[b - 4][1]] = 0.5;
    }
    else if (b < 16) {
        int k = (b - 4) / 2, l = (b - 4) % 2;
        lam[c_ev[k][0]] = l == 0 ? 2. / 3. : 1. / 3.;
        lam[c_ev[k][1]] = l == 0 ? 1. / 3. : 2. / 3.;
    }
    else {
        int f = b - 16;
        lam[c_fv[f][0]] = lam[c_fv[f][1]] = lam[c_fv[f][2]] = 1. / 3.;
    }
    double x = 0., y = 0., z = 0.;
    for (int i = 0; i < 4; i++) {
        const double *c = verts + 3 * (int64_t)ev[4 * e + i];
        x += c[0] * lam[i];
        y += c[1] * lam[i];
        z += c[2] * lam[i];
    }
    int64_t d = map[e * N + b];
    xyz[3 * d] = x;
    xyz[3 * d + 1] = y;
    xyz[3 * d + 2] = z;
}

//-------------------------------------------------------- incidence build
// row of local index a of element e
__device__ inline int64_t row_of(const int32_t *map, int64_t e, int a, int N,
                                 int dim)
{
    return (int64_t)map[e * N + a / dim] * dim + a % dim;
}

__global__ void k_inc_count(int64_t nelem, int N, int dim,
                            const int32_t *map, int32_t *cnt)
{
    int M = N * dim;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nelem * M)
        return;
    int64_t e = t / M;
    int a = (int)(t % M);
    atomicAdd(&cnt[row_of(map, e, a, N, dim)], 1);
}

__global__ void k_inc_fill(int64_t nelem, int N, int dim, const int32_t *map,
                           const int64_t *off, int32_t *cursor, int32_t *inc)
{
    int M = N * dim;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nelem * M)
        return;
    int64_t e = t / M;
    int a = (int)(t % M);
    int64_t r = row_of(map, e, a, N, dim);
    int pos = atomicAdd(&cursor[r], 1);
    inc[off[r] + pos] = (int32_t)t;
}

// insertion sort of each row's incidence list (lists are short)
__global__ void k_inc_sort(int64_t n, const int64_t *off, int32_t *inc)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    int64_t s = off[r], t = off[r + 1];
    for (int64_t i = s + 1; i < t; i++) {
        int32_t v = inc[i];
        int64_t j = i - 1;
        while (j >= s && inc[j] > v) {
            inc[j + 1] = inc[j];
            j--;
        }
        inc[j + 1] = v;
    }
}

//---------------------------------------------------------------- row walk
struct WalkArgs {
    int64_t n;
    int N, dim;
    const int32_t *map;
    const int64_t *inc_off;
    const int32_t *inc;
    const double *K;        // [nelem][M][M]
    const double *B;        // [nelem][M]
    const uint8_t *bdry;    // per row
    const double *g;        // per row
    const int64_t *scr_off; // scratch segment of each row
    int32_t *scols;
    double *svals;
    int32_t *count;         // [n] entries of each row
    double *rhs;            // [n]
    int32_t *ovf;           // rows past the table: ovf[0] their number, then
                            // the rows (null: replay them serially)
};

// one row replayed serially (matvec.c:3045 add_columns_to_row)
__device__ inline void walk_row_serial(const WalkArgs &w, int64_t r)
{
    const int M = w.N * w.dim;
    int32_t *cols = w.scols + w.scr_off[r];
    double *vals = w.svals + w.scr_off[r];
    int32_t nc = 0;
    double rhs = 0.0;
    const bool isb = w.bdry[r] != 0;
    for (int64_t t = w.inc_off[r]; t < w.inc_off[r + 1]; t++) {
        int32_t code = w.inc[t];
        int64_t e = code / M;
        int a = code % M;
        double bval;
        if (isb) {
            // Dirichlet row e_i (dof-utils.c:1369): only the diagonal 1.0
            int32_t c = (int32_t)r;
            int32_t i;
            for (i = 0; i < nc; i++)
                if (cols[i] == c)
                    break;
            if (i < nc)
                vals[i] += 1.0;
            else {
                cols[nc] = c;
                vals[nc++] = 1.0;
            }
            bval = w.g[r];
        }
        else {
            const double *Krow = w.K + (e * M + a) * (int64_t)M;
            for (int b = 0; b < M; b++) {
                double v = Krow[b];
                if (v == 0.0)
                    continue;   // matvec.c:5460
                int32_t c = (int32_t)row_of(w.map, e, b, w.N, w.dim);
                int32_t i;
                for (i = 0; i < nc; i++)
                    if (cols[i] == c)
                        break;
                if (i < nc)
                    vals[i] += v;
                else {
                    cols[nc] = c;
                    vals[nc++] = v;
                }
            }
            bval = w.B[e * M + a];
        }
        if (bval != 0.0)        // matvec.c:5708
            rhs += bval;
    }
    w.count[r] = nc;
    w.rhs[r] = rhs;
}

// The same insertion with a hash table: a group of G lanes per row (64/G rows
// per wavefront) steps through the row's incident elements in order; lane b
// takes column b of the element row.  The columns of one element row are
// distinct, so within a step each lane owns a different entry: a column
// found in the row's table adds its value to that entry, a new one opens the
// next entry, numbered by lane order.  Entries thus open in PHG's append
// order and every entry sums its contributions in element order starting
// from the first, as the row list does.  Dirichlet rows take the serial
// path.  Rows with more than E distinct columns go to the list w.ovf, which
// the LIST instance (one row per wavefront, a larger table) walks the same
// way afterwards; a row that overflows that table too, or any row when there
// is no list, is replayed serially.
template <int G, int E, bool LIST = false>
__global__ void __launch_bounds__(64) k_row_walk_hash(WalkArgs w)
{
    constexpr int TS = 2 * E, NG = 64 / G;
    constexpr int CC = 64;        // incident-element codes of a row kept in LDS
    __shared__ int32_t key[NG][TS];
    __shared__ int32_t ent[NG][TS];
    __shared__ int32_t icode[NG][CC];
    __shared__ int32_t ecol[NG][E];
    __shared__ double eval[NG][E];
    const int lane = threadIdx.x, grp = lane / G, b = lane % G;
    const uint64_t gm = (G == 64 ? ~0ull : (1ull << (G % 64)) - 1ull) << (grp * G % 64);
    const uint64_t below = (1ull << lane) - 1ull;
    const int M = w.N * w.dim;
    const int64_t nrow = LIST ? (int64_t)w.ovf[0] : w.n;
    for (int64_t r0 = (int64_t)blockIdx.x * NG; r0 < nrow; r0 += (int64_t)gridDim.x * NG) {
        const bool live = r0 + grp < nrow;
        const int64_t r = LIST ? (live ? (int64_t)w.ovf[1 + r0 + grp] : 0) : r0 + grp;
        const int64_t t0 = live ? w.inc_off[r] : 0;
        const int ninc = live ? (int)(w.inc_off[r + 1] - t0) : 0;
        const bool serial = live && w.bdry[r] != 0;
        const int steps = serial ? 0 : ninc;
        int smax = steps;
        for (int o = 32; o >= 1; o >>= 1)
            smax = max(smax, __shfl_xor(smax, o, 64));
        for (int k = b; k < TS; k += G)
            key[grp][k] = -1;
        // the row's incident-element codes, so that a step's loads depend on
        // an LDS read only
        for (int k = b; k < min(steps, CC); k += G)
            icode[grp][k] = w.inc[t0 + k];
        __syncthreads();
        int32_t nc = 0;
        bool over = false;
        double rhs = 0.0;
        // loads of a step: element row value, its column (loaded whatever the
        // value, so it does not wait for it) and the element's load entry;
        // step t + 1's are issued before step t's insertion
        int32_t cr = -1;
        double vr = 0.0, br = 0.0;
        auto fetch = [&](int t) {
            if (t < steps) {
                const int32_t code = t < CC ? icode[grp][t] : w.inc[t0 + t];
                if (b < M) {
                    const int64_t e = code / M;
                    vr = w.K[(int64_t)code * M + b];  // [e][a][b], code = e M + a
                    cr = (int32_t)row_of(w.map, e, b, w.N, w.dim);
                }
                if (b == 0)
                    br = w.B[code];
            }
        };
        fetch(0);
        for (int t = 0; t < smax; t++) {
            int32_t c = -1;
            double v = 0.0;
            if (t < steps) {
                if (b < M) {
                    v = vr;
                    if (v != 0.0)                     // matvec.c:5460
                        c = cr;
                }
                if (b == 0) {
                    const double bval = br;
                    if (bval != 0.0)                  // matvec.c:5708
                        rhs += bval;
                }
            }
            fetch(t + 1);
            int slot = -1;
            bool fresh = false;
            if (c >= 0 && !over) {
                slot = (int)(((uint32_t)c * 2654435761u) >> 8) & (TS - 1);
                for (int probe = 0; probe < TS; probe++) {
                    const int32_t old = atomicCAS(&key[grp][slot], -1, c);
                    if (old == -1) {
                        fresh = true;
                        break;
                    }
                    if (old == c)
                        break;
                    slot = (slot + 1) & (TS - 1);
                }
            }
            const uint64_t fm = __ballot(fresh) & gm;
            if (fresh) {
                const int32_t p = nc + __popcll(fm & below);
                if (p < E) {
                    ent[grp][slot] = p;
                    ecol[grp][p] = c;
                    eval[grp][p] = v;
                }
            }
            nc += __popcll(fm);
            over = over || nc > E;
            __syncthreads();
            if (c >= 0 && !fresh && !over)
                eval[grp][ent[grp][slot]] += v;
            __syncthreads();
        }
        if (live && !serial && !over) {
            int32_t *ocols = w.scols + w.scr_off[r];
            double *ovals = w.svals + w.scr_off[r];
            for (int k = b; k < nc; k += G) {
                ocols[k] = ecol[grp][k];
                ovals[k] = eval[grp][k];
            }
            if (b == 0) {
                w.count[r] = nc;
                w.rhs[r] = rhs;
            }
        }
        else if (live && b == 0) {
            if (!serial && !LIST && w.ovf != nullptr)
                w.ovf[1 + atomicAdd(&w.ovf[0], 1)] = (int32_t)r;
            else
                walk_row_serial(w, r);
        }
        __syncthreads();
    }
}

__global__ void k_row_flags(int64_t n, int dim, const uint8_t *node, uint8_t *row)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        row[i] = node[i / dim];
}

// The row copies below give each row a group of RG consecutive lanes, so a
// row's entries move in coalesced runs; RPB rows per 256-thread block.
constexpr int RG = 16, RPB = 256 / RG;

__device__ inline uint64_t group_mask(int lane)
{
    return ((1ull << RG) - 1ull) << (lane & ~(RG - 1));
}

__global__ void __launch_bounds__(256) k_compact(int64_t n, const int64_t *scr_off,
                                                 const int32_t *scols, const double *svals,
                                                 const int64_t *rowptr, int32_t *cols,
                                                 double *vals)
{
    const int64_t r = (int64_t)blockIdx.x * RPB + threadIdx.x / RG;
    const int sl = threadIdx.x % RG;
    if (r >= n)
        return;
    const int64_t s = scr_off[r], d = rowptr[r], len = rowptr[r + 1] - d;
    for (int64_t i = sl; i < len; i += RG) {
        cols[d + i] = scols[s + i];
        vals[d + i] = svals[s + i];
    }
}

//-------------------------------------------------------------- elimination
// Entry j of row r stays when its column is r or not a Dirichlet row.
__global__ void __launch_bounds__(256) k_elim_count(int64_t n, const int32_t *rowptr,
                                                    const int32_t *cols, const uint8_t *bdry,
                                                    int32_t *count)
{
    const int64_t r = (int64_t)blockIdx.x * RPB + threadIdx.x / RG;
    const int sl = threadIdx.x % RG, lane = threadIdx.x % 64;
    if (r >= n)
        return;
    const uint64_t gm = group_mask(lane);
    int32_t k = 0;
    for (int32_t j0 = rowptr[r]; j0 < rowptr[r + 1]; j0 += RG) {
        const int32_t j = j0 + sl;
        bool keep = false;
        if (j < rowptr[r + 1]) {
            const int32_t c = cols[j];
            keep = c == r || !bdry[c];
        }
        k += __popcll(__ballot(keep) & gm);
    }
    if (sl == 0)
        count[r] = k;
}

// Kept entries are written in their order; the rhs update runs over the
// row's Dirichlet columns in column-list order, as the serial loop of
// solver.c:707-725 does, so its rounding is the reference's.
__global__ void __launch_bounds__(256) k_elim_fill(int64_t n, const int32_t *rowptr,
                                                   const int32_t *cols, const double *vals,
                                                   const double *rhs, const uint8_t *bdry,
                                                   const int64_t *nrowptr, int32_t *ncols,
                                                   double *nvals, double *nrhs, double *diag)
{
    const int64_t r = (int64_t)blockIdx.x * RPB + threadIdx.x / RG;
    const int sl = threadIdx.x % RG, lane = threadIdx.x % 64;
    if (r >= n)
        return;
    const uint64_t gm = group_mask(lane), below = (1ull << lane) - 1ull;
    const int32_t end = rowptr[r + 1];
    int64_t w = nrowptr[r];
    double b = rhs[r], dg = 0.0;
    bool have_diag = false;
    for (int32_t j0 = rowptr[r]; j0 < end; j0 += RG) {
        const int32_t j = j0 + sl;
        int32_t c = -1;
        double v = 0.0;
        bool keep = false, elim = false, isdiag = false;
        if (j < end) {
            c = cols[j];
            v = vals[j];
            keep = c == r || !bdry[c];
            elim = bdry[c] && c != r;
            isdiag = c == r;
        }
        const uint64_t km = __ballot(keep) & gm;
        if (keep) {
            const int64_t o = w + __popcll(km & below);
            ncols[o] = c;
            nvals[o] = v;
        }
        w += __popcll(km);
        const uint64_t dm = __ballot(isdiag) & gm;
        if (dm != 0 && !have_diag) {
            dg = __shfl(v, __ffsll((unsigned long long)dm) - 1, 64);
            have_diag = true;
        }
        // solver.c:722, d_k is the diagonal of boundary row k
        double t = 0.0;
        if (elim) {
            double dk = 0.0;
            for (int32_t jj = rowptr[c]; jj < rowptr[c + 1]; jj++)
                if (cols[jj] == c) {
                    dk = vals[jj];
                    break;
                }
            t = 1.0 * rhs[c] * v / dk;
        }
        uint64_t em = __ballot(elim) & gm;
        while (em != 0) {
            const int src = __ffsll((unsigned long long)em) - 1;
            b -= __shfl(t, src, 64);
            em &= em - 1;
        }
    }
    if (sl == 0) {
        nrhs[r] = b;
        diag[r] = dg;
    }
}

//------------------------------------------------------------------ scan
// exclusive prefix sum of int32 counts into int64 offsets (n+1 entries)
__global__ void k_scan_blocks(int64_t n, const int32_t *in, int64_t *out,
                              int64_t *bsum)
{
    __shared__ int64_t s[1024];
    int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    int64_t v = i < n ? in[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int64_t t = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n)
        out[i] = s[threadIdx.x] - v;
    if (threadIdx.x == 1023)
        bsum[blockIdx.x] = s[1023];
}

// exclusive scan of the nb block sums in place, total in *total: one
// work-group of 1024, each work-item a consecutive chunk (integer sums, so
// the result is the sequential scan's)
__global__ void __launch_bounds__(1024) k_scan_sums(int64_t nb, int64_t *bsum,
                                                    int64_t *total)
{
    __shared__ int64_t sh[1024];
    const int t = threadIdx.x;
    const int64_t per = (nb + 1023) / 1024;
    const int64_t i0 = min(nb, (int64_t)t * per), i1 = min(nb, i0 + per);
    int64_t acc = 0;
    for (int64_t i = i0; i < i1; i++)
        acc += bsum[i];
    sh[t] = acc;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int64_t v = t >= o ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += v;
        __syncthreads();
    }
    int64_t run = sh[t] - acc;
    for (int64_t i = i0; i < i1; i++) {
        const int64_t v = bsum[i];
        bsum[i] = run;
        run += v;
    }
    if (t == 1023)
        *total = sh[1023];
}

__global__ void k_scan_add(int64_t n, int64_t *out, const int64_t *bsum,
                           const int64_t *total)
{
    int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    if (i < n)
        out[i] += bsum[blockIdx.x];
    if (i == n)
        out[n] = *total;
}

__global__ void k_to_int32(int64_t n, const int64_t *in, int32_t *out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n)
        out[i] = (int32_t)in[i];
}

__global__ void k_scale_i32(int64_t n, const int32_t *in, int32_t mul,
                            int32_t *out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = in[i] * mul;
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// out[0..n] = exclusive scan of in[0..n-1]; returns the total in *total_h
int scan_i32(int64_t n, const int32_t *d_in, int64_t *d_out, int64_t *total_h,
             hipStream_t s)
{
    int64_t nb = (n + 1 + 1023) / 1024;
    int64_t *bsum = nullptr, *tot = nullptr;
    int rc = phg::dmalloc((void **)&bsum, (size_t)(nb + 1) * sizeof(int64_t));
    if (rc)
        return rc;
    tot = bsum + nb;
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(1024), 0, s, n,
                       d_in, d_out, bsum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, nb, bsum, tot);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(1024), 0, s, n,
                       d_out, bsum, tot);
    PHG_HIP_CHECK(hipGetLastError());
    if (total_h) {
        PHG_HIP_CHECK(hipMemcpyAsync(total_h, tot, sizeof(int64_t),
                                     hipMemcpyDeviceToHost, s));
    }
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    phg::dfree(bsum);
    return PHG_HIP_OK;
}

// launches k_element for the space's order and the problem's dim
int launch_element(const phg_hip_space *sp, const phg_hip_problem *pb,
                   const ElemArgs &a, hipStream_t s)
{
    dim3 grid(nblk(sp->nelem, 128)), block(128);
    if (sp->nelem == 0)
        return PHG_HIP_OK;
    if (pb->dim == 1 && pb->qcache) {
        if (sp->order == 1)
            hipLaunchKernelGGL((k_element<1, 1, true>), grid, block, 0, s, a);
        else if (sp->order == 2)
            hipLaunchKernelGGL((k_element<2, 1, true>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((k_element<3, 1, true>), grid, block, 0, s, a);
    }
    else if (pb->dim == 1) {
        if (sp->order == 1)
            hipLaunchKernelGGL((k_element<1, 1>), grid, block, 0, s, a);
        else if (sp->order == 2)
            hipLaunchKernelGGL((k_element<2, 1>), grid, block, 0, s, a);
        else if (sp->order == 3)
            hipLaunchKernelGGL((k_element<3, 1>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((k_element<4, 1>), grid, block, 0, s, a);
    }
    else {
        if (sp->order == 1)
            hipLaunchKernelGGL((k_element<1, 3>), grid, block, 0, s, a);
        else if (sp->order == 2)
            hipLaunchKernelGGL((k_element<2, 3>), grid, block, 0, s, a);
        else
            hipLaunchKernelGGL((k_element<3, 3>), grid, block, 0, s, a);
    }
    PHG_HIP_CHECK(hipGetLastError());
    return PHG_HIP_OK;
}

// PHG's row insertion over all rows (w.ovf[0] must be 0): 64/G rows per
// wavefront, G the smallest of 16, 32, 64 holding an element row; the rows
// past its table (high-valence vertices of refined P3 meshes: up to 343
// columns) then one per wavefront with a table of 2048
int launch_row_walk(const WalkArgs &wk, hipStream_t s)
{
    const int M = wk.N * wk.dim;
    const unsigned grid = (unsigned)std::min<int64_t>((wk.n + 3) / 4, 1 << 16);
    if (wk.n == 0)
        return PHG_HIP_OK;
    if (M <= 16)
        hipLaunchKernelGGL((k_row_walk_hash<16, 128>), dim3(grid), dim3(64), 0, s, wk);
    else if (M <= 32)
        hipLaunchKernelGGL((k_row_walk_hash<32, 128>), dim3(grid), dim3(64), 0, s, wk);
    else   // M <= 60 for P1-P3, scalar or vector (check_problem)
        hipLaunchKernelGGL((k_row_walk_hash<64, 512>), dim3(grid), dim3(64), 0, s, wk);
    hipLaunchKernelGGL((k_row_walk_hash<64, 2048, true>), dim3(2048), dim3(64), 0, s, wk);
    PHG_HIP_CHECK(hipGetLastError());
    return PHG_HIP_OK;
}

// scratch segment offsets of the row walk: scr_off = exclusive scan of
// cnt[r] * M (cnt: incidence sizes), total in *nscr
int scratch_offsets(int64_t n, const int32_t *cnt, int M, int64_t **scr_off, int64_t *nscr)
{
    hipStream_t s = phg::ctx()->stream;
    int32_t *count = nullptr;
    int rc;
    *scr_off = nullptr;
    if ((rc = phg::dmalloc((void **)&count, (size_t)(n + 1) * sizeof(int32_t))) ||
        (rc = phg::dmalloc((void **)scr_off, (size_t)(n + 1) * sizeof(int64_t)))) {
        phg::dfree(count);
        return rc;
    }
    hipLaunchKernelGGL(k_scale_i32, dim3(nblk(n, 256)), dim3(256), 0, s, n, cnt, (int32_t)M,
                       count);
    rc = scan_i32(n, count, *scr_off, nscr, s);
    phg::dfree(count);
    if (rc) {
        phg::dfree(*scr_off);
        *scr_off = nullptr;
    }
    return rc;
}

int check_problem(const phg_hip_space *sp, const phg_hip_problem *pb)
{
    if (sp == nullptr || pb == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "null space or problem");
    if (pb->dim != 1 && pb->dim != 3)
        return phg::set_error(PHG_HIP_E_ARG, "dim must be 1 or 3 (got %d)",
                              pb->dim);
    if (pb->dim == 3 && (pb->use_kappa || pb->mass_coef != 0.0))
        return phg::set_error(PHG_HIP_E_ARG,
                              "kappa/mass terms are scalar-only");
    // rules up to order 6 for P1-P3, up to order 8 for P4 (scalar, without
    // the quadrature cache)
    const int maxq = sp->order == 4 ? 8 : 6;
    if (sp->order == 4 && (pb->dim != 1 || pb->qcache))
        return phg::set_error(PHG_HIP_E_ARG, "P4: scalar problems without qcache only");
    if (pb->rhs_order > maxq)
        return phg::set_error(PHG_HIP_E_ARG,
                              "load quadrature order %d > %d unsupported",
                              pb->rhs_order, maxq);
    if (pb->mass_coef != 0.0 && 2 * sp->order > maxq)
        return phg::set_error(PHG_HIP_E_ARG, "mass order too high");
    if (pb->qcache && (pb->dim != 1 || pb->use_kappa || pb->mass_coef != 0.0 ||
                       (pb->rhs_order >= 0 && pb->rhs_order != 2 * sp->order)))
        return phg::set_error(PHG_HIP_E_ARG,
                              "qcache: the Laplacian of examples/poisson.c only (no kappa, "
                              "no mass term, load on the order-2p rule)");
    return PHG_HIP_OK;
}

// element matrices into device buffers K (nelem*M*M) and B (nelem*M)
// Incidence lists: for every row (scalar DOF * dim + component) the codes
// e * M + a of the (element, local index) pairs that touch it, in element
// order -- the order in which PHG's element loop adds to the row.  cnt[r]
// holds the list lengths on return.
int build_incidence(const phg_hip_space *sp, int dim, int32_t **cnt_out,
                    int64_t **off_out, int32_t **inc_out)
{
    hipStream_t s = phg::ctx()->stream;
    const int N = sp->nbas, M = N * dim;
    const int64_t n = sp->ndof * dim, ne = sp->nelem;
    int32_t *cnt = nullptr, *inc = nullptr;
    int64_t *off = nullptr, ninc = 0;
    int rc;
    if ((rc = phg::dmalloc((void **)&cnt, (size_t)(n + 1) * sizeof(int32_t))) ||
        (rc = phg::dmalloc((void **)&off, (size_t)(n + 1) * sizeof(int64_t))) ||
        (rc = phg::dmalloc((void **)&inc, (size_t)(ne * M + 1) * sizeof(int32_t))))
        goto fail;
    if (hipMemsetAsync(cnt, 0, (size_t)(n + 1) * sizeof(int32_t), s) != hipSuccess) {
        rc = phg::set_error(PHG_HIP_E_HIP, "memset failed");
        goto fail;
    }
    hipLaunchKernelGGL(k_inc_count, dim3(nblk(ne * M, 256)), dim3(256), 0, s,
                       ne, N, dim, (const int32_t *)sp->d_map, cnt);
    if ((rc = scan_i32(n, cnt, off, &ninc, s)))
        goto fail;
    if (hipMemsetAsync(cnt, 0, (size_t)(n + 1) * sizeof(int32_t), s) != hipSuccess) {
        rc = phg::set_error(PHG_HIP_E_HIP, "memset failed");
        goto fail;
    }
    hipLaunchKernelGGL(k_inc_fill, dim3(nblk(ne * M, 256)), dim3(256), 0, s,
                       ne, N, dim, (const int32_t *)sp->d_map,
                       (const int64_t *)off, cnt, inc);
    hipLaunchKernelGGL(k_inc_sort, dim3(nblk(n, 256)), dim3(256), 0, s, n,
                       (const int64_t *)off, inc);
    *cnt_out = cnt;
    *off_out = off;
    *inc_out = inc;
    return PHG_HIP_OK;
fail:
    phg::dfree(cnt);
    phg::dfree(off);
    phg::dfree(inc);
    return rc;
}

}  // namespace

// scratch buffers of the assembly workspace
enum { WS_K, WS_B, WS_F, WS_KAPPA, WS_G, WS_COUNT, WS_SCOLS, WS_SVALS, WS_ROWPTR, WS_OVF, WS_N };
struct AsmWorkspace {
    int dim = 0;                      // problem dimension of the incidence (0: none)
    int32_t *cnt = nullptr, *inc = nullptr;
    int64_t *inc_off = nullptr;
    int64_t *scr_off = nullptr;       // scratch segment of each row (incidence * M)
    int64_t nscr = 0;
    void *buf[WS_N] = {};
    size_t bytes[WS_N] = {};
};

void free_asm_workspace(AsmWorkspace *w)
{
    if (w == nullptr)
        return;
    phg::dfree(w->cnt);
    phg::dfree(w->inc);
    phg::dfree(w->inc_off);
    phg::dfree(w->scr_off);
    for (int i = 0; i < WS_N; i++)
        phg::dfree(w->buf[i]);
    delete w;
}

namespace {
AsmWorkspace *workspace(phg_hip_space *sp)
{
    if (sp->asm_ws == nullptr)
        sp->asm_ws = new AsmWorkspace();
    return sp->asm_ws;
}

// a scratch buffer of at least `bytes`, kept by the space for later calls
int ws_buffer(phg_hip_space *sp, int slot, size_t bytes, void **p)
{
    AsmWorkspace *w = workspace(sp);
    if (w->bytes[slot] < bytes || w->buf[slot] == nullptr) {
        phg::dfree(w->buf[slot]);
        w->buf[slot] = nullptr;
        w->bytes[slot] = 0;
        int rc = phg::dmalloc(&w->buf[slot], bytes ? bytes : 1);
        if (rc)
            return rc;
        w->bytes[slot] = bytes;
    }
    *p = w->buf[slot];
    return PHG_HIP_OK;
}

// the row incidence of a problem dimension and its scratch segments, built
// on the first call for that dimension
int space_incidence(phg_hip_space *sp, int dim, AsmWorkspace **out)
{
    AsmWorkspace *w = workspace(sp);
    *out = w;
    if (w->dim == dim)
        return PHG_HIP_OK;
    phg::dfree(w->cnt);
    phg::dfree(w->inc);
    phg::dfree(w->inc_off);
    phg::dfree(w->scr_off);
    w->cnt = w->inc = nullptr;
    w->inc_off = w->scr_off = nullptr;
    w->dim = 0;
    int rc;
    if ((rc = build_incidence(sp, dim, &w->cnt, &w->inc_off, &w->inc)) ||
        (rc = scratch_offsets(sp->ndof * dim, w->cnt, sp->nbas * dim, &w->scr_off, &w->nscr)))
        return rc;
    w->dim = dim;
    return PHG_HIP_OK;
}

int element_pass(phg_hip_space *sp, const phg_hip_problem *pb,
                 const double *kappa_h, const double *f_h, double **dK,
                 double **dB)
{
    hipStream_t s = phg::ctx()->stream;
    int N = sp->nbas, M = N * pb->dim, rc;
    double *dk = nullptr, *df = nullptr;
    *dK = *dB = nullptr;
    if (pb->use_kappa && kappa_h == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "use_kappa set but kappa is NULL");
    if (f_h == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "load vector f is NULL");
    // element matrices and loads in the space's workspace
    size_t ksz = (size_t)sp->nelem * M * M * sizeof(double);
    if ((rc = ws_buffer(sp, WS_K, ksz, (void **)dK)) ||
        (rc = ws_buffer(sp, WS_B, (size_t)sp->nelem * M * sizeof(double), (void **)dB)) ||
        (rc = ws_buffer(sp, WS_F, (size_t)sp->ndof * pb->dim * sizeof(double),
                        (void **)&df)) ||
        (rc = phg::upload(df, f_h, (size_t)sp->ndof * pb->dim * sizeof(double))))
        return rc;
    if (pb->use_kappa) {
        if ((rc = ws_buffer(sp, WS_KAPPA, (size_t)sp->nelem * sizeof(double), (void **)&dk)) ||
            (rc = phg::upload(dk, kappa_h, (size_t)sp->nelem * sizeof(double))))
            return rc;
    }
    ElemArgs a;
    a.nelem = sp->nelem;
    a.verts = sp->d_verts;
    a.elem_verts = sp->d_elem_verts;
    a.map = sp->d_map;
    a.kappa = dk;
    a.f = df;
    a.use_kappa = pb->use_kappa;
    a.mass_coef = pb->mass_coef;
    a.rhs_order = pb->rhs_order;
    a.E = pb->E;
    a.nu = pb->nu;
    a.K = *dK;
    a.B = *dB;
    if ((rc = launch_element(sp, pb, a, s)))
        return rc;
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return phg::hip_fail(e, "element kernel", __FILE__, __LINE__);
    return PHG_HIP_OK;
}

// Element coefficients of the stiffness factorisation used by the
// matrix-free operator: K_e = sum_m c_m(e) T^m with, for the pairs
// m = (a, b), 1 <= a <= b <= 3, of independent barycentric coordinates,
//   c_m(e) = kappa_e vol_e (grad lambda_a . grad lambda_b)
// (grad lambda_a = row a of the Jacobian, geom.c:103; grad lambda_0 =
// -(grad lambda_1 + grad lambda_2 + grad lambda_3) is eliminated, which
// keeps the 3 x 3 metric nonsingular and the sum free of the cancellation
// a 4 x 4 Gram matrix of rank 3 would add).  Element i of the output is
// mesh element eperm[i].
__global__ void k_eop_coef(int64_t ne, const double *verts, const int32_t *ev,
                           const double *kappa, const int32_t *eperm, int nc,
                           double mass, float *coef)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne)
        return;
    const int64_t e = eperm[i];
    double x[4][3];
    for (int v = 0; v < 4; v++) {
        const int64_t g = ev[4 * e + v];
        for (int c = 0; c < 3; c++)
            x[v][c] = verts[3 * g + c];
    }
    double J[16];
    jacobian(x, J);
    const double sc = (kappa ? kappa[e] : 1.0) * volume(x);
    int m = 0;
    for (int a = 1; a < 4; a++)
        for (int b = a; b < 4; b++, m++)
            coef[nc * i + m] = (float)(sc * (J[4 * a] * J[4 * b] + J[4 * a + 1] * J[4 * b + 1] +
                                             J[4 * a + 2] * J[4 * b + 2]));
    if (nc > 6)
        coef[nc * i + 6] = (float)(mass * volume(x));
}

}  // namespace

namespace phg {

// T^m_ij = sum_q w_q D_a phi_i D_b phi_j (+ the (b, a) term for a < b),
// D_a = d / d lambda_a - d / d lambda_0 (a = 1..3), on the rule of order
// 2(p-1) (exact for the integrand; quad.c rules, lagrange.c gradients), so
// that sum_m c_m(e) T^m_ij = kappa_e vol_e sum_q w_q grad phi_i . grad phi_j,
// the element stiffness of quad.c:5260 phgQuadGradBasAGradBas.
// T is [7][nb][nb], pairs m in the order (1,1) (1,2) (1,3) (2,2) (2,3) (3,3),
// then T^6_ij = sum_q w_q phi_i phi_j on the rule of order 2p, so that
// vol_e T^6 is the element mass matrix of quad.c:5582 phgQuadBasDotBas.
void eop_reference_tensors(int p, double *T)
{
    const int nb = p == 1 ? 4 : (p == 2 ? 10 : 20);
    Rule q;
    make_rule(2 * (p - 1), q);
    std::vector<double> S((size_t)9 * nb * nb, 0.0);   // S[a-1][b-1][i][j]
    double g[4 * NB], d[3 * NB];
    for (int k = 0; k < q.np; k++) {
        bas_grad_lambda(p, q.pts[k], g);
        for (int i = 0; i < nb; i++)
            for (int a = 1; a < 4; a++)
                d[3 * i + a - 1] = g[4 * i + a] - g[4 * i];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                for (int i = 0; i < nb; i++)
                    for (int j = 0; j < nb; j++)
                        S[(((size_t)(3 * a + b) * nb) + i) * nb + j] +=
                            q.w[k] * d[3 * i + a] * d[3 * j + b];
    }
    int m = 0;
    for (int a = 0; a < 3; a++)
        for (int b = a; b < 3; b++, m++)
            for (int i = 0; i < nb; i++)
                for (int j = 0; j < nb; j++) {
                    double t = S[(((size_t)(3 * a + b) * nb) + i) * nb + j];
                    if (a != b)
                        t += S[(((size_t)(3 * b + a) * nb) + i) * nb + j];
                    T[((size_t)m * nb + i) * nb + j] = t;
                }
    make_rule(2 * p, q);
    double *M = T + (size_t)6 * nb * nb;
    for (int k = 0; k < nb * nb; k++)
        M[k] = 0.0;
    for (int k = 0; k < q.np; k++) {
        double b[NB];
        bas_values(p, q.pts[k], b);
        for (int i = 0; i < nb; i++)
            for (int j = 0; j < nb; j++)
                M[(size_t)i * nb + j] += q.w[k] * (b[j] * b[i]);
    }
}

int eop_rule_gradients(int p, double *w, double *D)
{
    const int nb = p == 1 ? 4 : (p == 2 ? 10 : 20);
    Rule q;
    make_rule(2 * (p - 1), q);
    if (q.np > NPMAX_EOP)
        return -1;
    double g[4 * NB];
    for (int k = 0; k < q.np; k++) {
        w[k] = q.w[k];
        bas_grad_lambda(p, q.pts[k], g);
        for (int i = 0; i < nb; i++)
            for (int a = 1; a < 4; a++)
                D[((size_t)k * nb + i) * 3 + a - 1] = g[4 * i + a] - g[4 * i];
    }
    return q.np;
}

int eop_coefficients(const phg_hip_space *sp, const double *d_kappa,
                     const int32_t *d_eperm, double mass_coef, float *d_coef,
                     hipStream_t s)
{
    if (sp->nelem == 0)
        return PHG_HIP_OK;
    hipLaunchKernelGGL(k_eop_coef, dim3(nblk(sp->nelem, 256)), dim3(256), 0, s,
                       sp->nelem, (const double *)sp->d_verts,
                       (const int32_t *)sp->d_elem_verts, d_kappa, d_eperm,
                       mass_coef != 0.0 ? 7 : 6, mass_coef, d_coef);
    PHG_HIP_CHECK(hipGetLastError());
    return PHG_HIP_OK;
}

}  // namespace phg

//============================================================== C ABI ====

extern "C" int phg_hip_space_create(int32_t order, int64_t nvert,
                                    const double *verts, int64_t nelem,
                                    const int32_t *elem_verts,
                                    const int32_t *map, int64_t ndof,
                                    const uint8_t *bdry, phg_hip_space **out)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (out == nullptr || order < 1 || order > 4 || nvert <= 0 || nelem < 0 ||
        ndof <= 0 || verts == nullptr || elem_verts == nullptr ||
        map == nullptr || bdry == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_space_create: bad args");
    int nb = order == 1 ? 4 : (order == 2 ? 10 : (order == 3 ? 20 : 35));
    // host-side bounds checks so that no kernel indexes out of range
    for (int64_t i = 0; i < nelem * 4; i++)
        if (elem_verts[i] < 0 || elem_verts[i] >= nvert)
            return phg::set_error(PHG_HIP_E_ARG, "element vertex out of range");
    for (int64_t i = 0; i < nelem * nb; i++)
        if (map[i] < 0 || map[i] >= ndof)
            return phg::set_error(PHG_HIP_E_ARG, "DOF index out of range");
    if (nelem * (int64_t)nb * 3 >= (int64_t)INT32_MAX)
        return phg::set_error(PHG_HIP_E_ARG, "mesh too large for int32 codes");
    static std::atomic<uint64_t> last_uid{0};
    phg_hip_space *sp = new phg_hip_space();
    sp->uid = ++last_uid;
    sp->order = order;
    sp->nbas = nb;
    sp->nvert = nvert;
    sp->nelem = nelem;
    sp->ndof = ndof;
    // bounding box of the vertices: every Lagrange node is a convex
    // combination of its element's vertices, so it bounds the nodes too
    for (int c = 0; c < 3; c++)
        sp->lo[c] = sp->hi[c] = verts[c];
    for (int64_t v = 0; v < nvert; v++)
        for (int c = 0; c < 3; c++) {
            sp->lo[c] = std::min(sp->lo[c], verts[3 * v + c]);
            sp->hi[c] = std::max(sp->hi[c], verts[3 * v + c]);
        }
    if ((rc = phg::dmalloc((void **)&sp->d_verts, (size_t)nvert * 3 * sizeof(double))) ||
        (rc = phg::dmalloc((void **)&sp->d_elem_verts, (size_t)nelem * 4 * sizeof(int32_t))) ||
        (rc = phg::dmalloc((void **)&sp->d_map, (size_t)nelem * nb * sizeof(int32_t))) ||
        (rc = phg::dmalloc((void **)&sp->d_bdry, (size_t)ndof)) ||
        (rc = phg::upload(sp->d_verts, verts, (size_t)nvert * 3 * sizeof(double))) ||
        (rc = phg::upload(sp->d_elem_verts, elem_verts, (size_t)nelem * 4 * sizeof(int32_t))) ||
        (rc = phg::upload(sp->d_map, map, (size_t)nelem * nb * sizeof(int32_t))) ||
        (rc = phg::upload(sp->d_bdry, bdry, (size_t)ndof))) {
        phg_hip_space_destroy(sp);
        return rc;
    }
    *out = sp;
    return PHG_HIP_OK;
}

extern "C" int phg_hip_space_destroy(phg_hip_space *sp)
{
    if (sp == nullptr)
        return PHG_HIP_OK;
    if (sp->batch_refs > 0)
        return phg::set_error(PHG_HIP_E_STATE, "phg_hip_space_destroy: the space is in a "
                                               "batch; destroy the batch first");
    phg::dfree(sp->d_verts);
    phg::dfree(sp->d_elem_verts);
    phg::dfree(sp->d_map);
    phg::dfree(sp->d_bdry);
    free_asm_workspace(sp->asm_ws);
    delete sp;
    return PHG_HIP_OK;
}

extern "C" int phg_hip_element_matrices(phg_hip_space *sp,
                                        const phg_hip_problem *pb,
                                        const double *kappa, const double *f,
                                        double *K, double *b)
{
    int rc;
    if ((rc = phg::require_ctx()) || (rc = check_problem(sp, pb)))
        return rc;
    double *dK, *dB;
    if ((rc = element_pass(sp, pb, kappa, f, &dK, &dB)))
        return rc;
    int M = sp->nbas * pb->dim;
    if (K)
        rc = phg::download(K, dK, (size_t)sp->nelem * M * M * sizeof(double));
    if (!rc && b)
        rc = phg::download(b, dB, (size_t)sp->nelem * M * sizeof(double));
    return rc;
}

extern "C" int phg_hip_assemble(phg_hip_space *sp, const phg_hip_problem *pb,
                                const double *kappa, const double *f,
                                const double *g, phg_hip_mat **Aout)
{
    int rc;
    if ((rc = phg::require_ctx()) || (rc = check_problem(sp, pb)))
        return rc;
    if (Aout == nullptr || g == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_assemble: bad args");
    hipStream_t s = phg::ctx()->stream;
    const int N = sp->nbas, dim = pb->dim, M = N * dim;
    const int64_t n = sp->ndof * dim, ne = sp->nelem;
    // scratch (element matrices, incidence, row segments) from the space's
    // workspace; the matrix owns its CSR, right-hand side and row flags
    double *dK = nullptr, *dB = nullptr, *dg = nullptr;
    int32_t *count = nullptr;
    int64_t *rowptr64 = nullptr;
    int32_t *scols = nullptr;
    double *svals = nullptr;
    uint8_t *bdry_row = nullptr;
    AsmWorkspace *w = nullptr;
    int64_t nnz = 0;
    phg_hip_mat *A = nullptr;
    phg::PhaseTrace tr("phg_hip_assemble");

    if ((rc = element_pass(sp, pb, kappa, f, &dK, &dB)))
        return rc;
    tr.mark("element pass (uploads f, kappa)");

    // per-row Dirichlet flags (each node's flag on its dim rows) and values
    if ((rc = phg::dmalloc((void **)&bdry_row, (size_t)n)))
        goto done;
    hipLaunchKernelGGL(k_row_flags, dim3(nblk(n, 256)), dim3(256), 0, s, n, dim,
                       (const uint8_t *)sp->d_bdry, bdry_row);
    if ((rc = ws_buffer(sp, WS_G, (size_t)n * sizeof(double), (void **)&dg)) ||
        (rc = phg::upload(dg, g, (size_t)n * sizeof(double))))
        goto done;

    tr.mark("row flags, upload g");
    // incidence
    if ((rc = space_incidence(sp, dim, &w)))
        goto done;
    tr.mark("incidence");
    // scratch segments: incidence * M entries per row (upper bound)
    if ((rc = ws_buffer(sp, WS_COUNT, (size_t)(n + 1) * sizeof(int32_t), (void **)&count)) ||
        (rc = ws_buffer(sp, WS_SCOLS, (size_t)(w->nscr + 1) * sizeof(int32_t),
                        (void **)&scols)) ||
        (rc = ws_buffer(sp, WS_SVALS, (size_t)(w->nscr + 1) * sizeof(double),
                        (void **)&svals)))
        goto done;

    tr.mark("scratch segments");
    A = new phg_hip_mat();
    A->n = n;
    A->dim = dim;
    A->ndof = sp->ndof;
    A->space_uid = sp->uid;
    for (int c = 0; c < 3; c++) {
        A->xyz_lo[c] = sp->lo[c];
        A->xyz_hi[c] = sp->hi[c];
    }
    A->xyz_box = 1;
    if ((rc = phg::dmalloc((void **)&A->d_rhs, (size_t)n * sizeof(double))) ||
        (rc = phg::dmalloc((void **)&A->d_xyz, (size_t)sp->ndof * 3 * sizeof(double))))
        goto done;
    // nodal coordinates (identical values whichever element writes a node
    // for P1/P2 nodes; P3 face nodes may differ in the last bit, which only
    // matters for the ordering key)
    hipLaunchKernelGGL(k_dof_coords, dim3(nblk(ne * N, 256)), dim3(256), 0, s,
                       ne, (int)sp->order, N, (const double *)sp->d_verts,
                       (const int32_t *)sp->d_elem_verts,
                       (const int32_t *)sp->d_map, A->d_xyz);
    {
        WalkArgs wk;
        wk.n = n;
        wk.N = N;
        wk.dim = dim;
        wk.map = sp->d_map;
        wk.inc_off = w->inc_off;
        wk.inc = w->inc;
        wk.K = dK;
        wk.B = dB;
        wk.bdry = bdry_row;
        wk.g = dg;
        wk.scr_off = w->scr_off;
        wk.scols = scols;
        wk.svals = svals;
        wk.count = count;
        wk.rhs = A->d_rhs;
        if ((rc = ws_buffer(sp, WS_OVF, (size_t)(n + 1) * sizeof(int32_t),
                            (void **)&wk.ovf)))
            goto done;
        PHG_HIP_CHECK(hipMemsetAsync(wk.ovf, 0, sizeof(int32_t), s));
        if ((rc = launch_row_walk(wk, s)))
            goto done;
    }
    tr.mark("row walk");
    if ((rc = ws_buffer(sp, WS_ROWPTR, (size_t)(n + 1) * sizeof(int64_t),
                        (void **)&rowptr64)) ||
        (rc = scan_i32(n, count, rowptr64, &nnz, s)))
        goto done;
    if (nnz >= (int64_t)INT32_MAX) {
        rc = phg::set_error(PHG_HIP_E_ARG, "nnz %lld exceeds int32 CSR",
                            (long long)nnz);
        goto done;
    }
    A->nnz = nnz;
    if ((rc = phg::dmalloc((void **)&A->d_rowptr, (size_t)(n + 1) * sizeof(int32_t))) ||
        (rc = phg::dmalloc((void **)&A->d_cols, (size_t)(nnz + 1) * sizeof(int32_t))) ||
        (rc = phg::dmalloc((void **)&A->d_vals, (size_t)(nnz + 1) * sizeof(double))))
        goto done;
    hipLaunchKernelGGL(k_compact, dim3(nblk(n, RPB)), dim3(256), 0, s, n,
                       (const int64_t *)w->scr_off, (const int32_t *)scols,
                       (const double *)svals, (const int64_t *)rowptr64,
                       A->d_cols, A->d_vals);
    hipLaunchKernelGGL(k_to_int32, dim3(nblk(n + 1, 256)), dim3(256), 0, s, n,
                       (const int64_t *)rowptr64, A->d_rowptr);
    PHG_HIP_CHECK(hipGetLastError());
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    tr.mark("row pointers, compaction");
    A->d_bdry_row = bdry_row;
    bdry_row = nullptr;
    *Aout = A;
    A = nullptr;
    rc = PHG_HIP_OK;
done:
    if (A)
        phg_hip_mat_destroy(A);
    phg::dfree(bdry_row);
    return rc;
}

extern "C" int phg_hip_mat_eliminate(phg_hip_mat *A)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (A == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "null matrix");
    if (A->eliminated)
        return PHG_HIP_OK;
    hipStream_t s = phg::ctx()->stream;
    phg::PhaseTrace tr("phg_hip_mat_eliminate");
    int64_t n = A->n, nnz = 0;
    int32_t *count = nullptr, *ncols = nullptr, *nrow32 = nullptr;
    int64_t *nrow = nullptr;
    double *nvals = nullptr, *nrhs = nullptr, *diag = nullptr;
    if ((rc = phg::dmalloc((void **)&count, (size_t)(n + 1) * sizeof(int32_t))) ||
        (rc = phg::dmalloc((void **)&nrow, (size_t)(n + 1) * sizeof(int64_t))))
        goto fail;
    hipLaunchKernelGGL(k_elim_count, dim3(nblk(n, RPB)), dim3(256), 0, s, n,
                       (const int32_t *)A->d_rowptr, (const int32_t *)A->d_cols,
                       (const uint8_t *)A->d_bdry_row, count);
    if ((rc = scan_i32(n, count, nrow, &nnz, s)))
        goto fail;
    if ((rc = phg::dmalloc((void **)&ncols, (size_t)(nnz + 1) * sizeof(int32_t))) ||
        (rc = phg::dmalloc((void **)&nvals, (size_t)(nnz + 1) * sizeof(double))) ||
        (rc = phg::dmalloc((void **)&nrhs, (size_t)n * sizeof(double))) ||
        (rc = phg::dmalloc((void **)&diag, (size_t)n * sizeof(double))) ||
        (rc = phg::dmalloc((void **)&nrow32, (size_t)(n + 1) * sizeof(int32_t))))
        goto fail;
    hipLaunchKernelGGL(k_elim_fill, dim3(nblk(n, RPB)), dim3(256), 0, s, n,
                       (const int32_t *)A->d_rowptr, (const int32_t *)A->d_cols,
                       (const double *)A->d_vals, (const double *)A->d_rhs,
                       (const uint8_t *)A->d_bdry_row, (const int64_t *)nrow,
                       ncols, nvals, nrhs, diag);
    hipLaunchKernelGGL(k_to_int32, dim3(nblk(n + 1, 256)), dim3(256), 0, s, n,
                       (const int64_t *)nrow, nrow32);
    PHG_HIP_CHECK(hipGetLastError());
    PHG_HIP_CHECK(hipStreamSynchronize(s));
    phg::dfree(A->d_rowptr);
    phg::dfree(A->d_cols);
    phg::dfree(A->d_vals);
    phg::dfree(A->d_rhs);
    phg::dfree(A->d_diag);
    phg::free_sell(A);
    phg::free_solver_copy(A);
    A->d_rowptr = nrow32;
    A->d_cols = ncols;
    A->d_vals = nvals;
    A->d_rhs = nrhs;
    A->d_diag = diag;
    A->nnz = nnz;
    A->eliminated = 1;
    phg::dfree(count);
    phg::dfree(nrow);
    tr.mark("boundary columns, lifting");
    // the row blocks of the PHG-order CSR are split again at its first
    // product
    phg::dfree(A->d_rblk);
    A->d_rblk = nullptr;
    A->nrblk = 0;
    rc = phg::build_solver_copy(A);
    tr.mark("solver numbering and copy");
    return rc;
fail:
    phg::dfree(count);
    phg::dfree(nrow);
    phg::dfree(ncols);
    phg::dfree(nvals);
    phg::dfree(nrhs);
    phg::dfree(diag);
    phg::dfree(nrow32);
    return rc;
}

extern "C" int phg_hip_mat_info(const phg_hip_mat *A, int64_t *nrows,
                                int64_t *nnz)
{
    if (A == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "null matrix");
    if (nrows)
        *nrows = A->n;
    if (nnz)
        *nnz = A->nnz;
    return PHG_HIP_OK;
}

extern "C" int phg_hip_mat_download(const phg_hip_mat *A, int64_t *rowptr,
                                    int32_t *cols, double *vals, double *rhs)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (A == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "null matrix");
    if (rowptr) {
        std::vector<int32_t> r32((size_t)A->n + 1);
        if ((rc = phg::download(r32.data(), A->d_rowptr,
                                (size_t)(A->n + 1) * sizeof(int32_t))))
            return rc;
        for (int64_t i = 0; i <= A->n; i++)
            rowptr[i] = r32[(size_t)i];
    }
    if (cols && (rc = phg::download(cols, A->d_cols, (size_t)A->nnz * sizeof(int32_t))))
        return rc;
    if (vals && (rc = phg::download(vals, A->d_vals, (size_t)A->nnz * sizeof(double))))
        return rc;
    if (rhs && (rc = phg::download(rhs, A->d_rhs, (size_t)A->n * sizeof(double))))
        return rc;
    return PHG_HIP_OK;
}

// matvec.c:6134 phgMatDumpCSR: the size, the n + 1 row pointers (0-based),
// the column indices and the values ("%0.17lg"), one number per line, of
// the rows in storage order; like PHG it assembles (eliminates) first
extern "C" int phg_hip_mat_dump_csr(phg_hip_mat *A, const char *fn)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (A == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "null matrix");
    if (fn == nullptr)
        return PHG_HIP_OK;      // matvec.c:6185: nothing to do
    if (!A->eliminated && (rc = phg_hip_mat_eliminate(A)))
        return rc;
    std::vector<int32_t> rp((size_t)A->n + 1), ci((size_t)A->nnz);
    std::vector<double> va((size_t)A->nnz);
    if ((rc = phg::download(rp.data(), A->d_rowptr, rp.size() * sizeof(int32_t))) ||
        (rc = phg::download(ci.data(), A->d_cols, ci.size() * sizeof(int32_t))) ||
        (rc = phg::download(va.data(), A->d_vals, va.size() * sizeof(double))))
        return rc;
    FILE *fp = fopen(fn, "w+t");
    if (fp == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "cannot open file \"%s\"", fn);
    fprintf(fp, "%lld\n", (long long)A->n);
    for (int64_t i = 0; i <= A->n; i++)
        fprintf(fp, "%d\n", rp[(size_t)i]);
    for (int64_t k = 0; k < A->nnz; k++)
        fprintf(fp, "%d\n", ci[(size_t)k]);
    for (int64_t k = 0; k < A->nnz; k++)
        fprintf(fp, "%0.17lg\n", va[(size_t)k]);
    const bool bad = ferror(fp) != 0;
    if (fclose(fp) != 0 || bad)
        return phg::set_error(PHG_HIP_E_ARG, "write error on \"%s\"", fn);
    return PHG_HIP_OK;
}

// matvec.c:6245 phgVecDump: the size, then one value ("%0.17lg") a line
extern "C" int phg_hip_vec_dump(const phg_hip_vec *v, const char *fn)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (v == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "null vector");
    if (fn == nullptr)
        return PHG_HIP_OK;
    std::vector<double> h((size_t)v->n);
    if ((rc = phg::download(h.data(), v->d, h.size() * sizeof(double))))
        return rc;
    FILE *fp = fopen(fn, "w+t");
    if (fp == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "cannot open file \"%s\"", fn);
    fprintf(fp, "%lld\n", (long long)v->n);
    for (int64_t i = 0; i < v->n; i++)
        fprintf(fp, "%0.17lg\n", h[(size_t)i]);
    const bool bad = ferror(fp) != 0;
    if (fclose(fp) != 0 || bad)
        return phg::set_error(PHG_HIP_E_ARG, "write error on \"%s\"", fn);
    return PHG_HIP_OK;
}

extern "C" int phg_hip_mat_destroy(phg_hip_mat *A)
{
    if (A == nullptr)
        return PHG_HIP_OK;
    if (A->batch_refs > 0)
        return phg::set_error(PHG_HIP_E_STATE,
                              "phg_hip_mat_destroy: the matrix is in a batch; destroy the "
                              "batch first");
    phg::dfree(A->d_rowptr);
    phg::dfree(A->d_cols);
    phg::dfree(A->d_vals);
    phg::dfree(A->d_rhs);
    phg::dfree(A->d_bdry_row);
    phg::dfree(A->d_diag);
    phg::dfree(A->d_rblk);
    phg::dfree(A->d_xyz);
    phg::free_sell(A);
    phg::free_solver_copy(A);
    phg::free_workspace(A);
    delete A;
    return PHG_HIP_OK;
}

//----------------------------------------------------------------- adjoint
extern "C" int phg_hip_adjoint_logkappa(phg_hip_space *sp,
                                        const phg_hip_problem *pb,
                                        const double *kappa,
                                        const phg_hip_mat *A,
                                        const phg_hip_vec *u,
                                        const phg_hip_vec *ud, int32_t maxit,
                                        double rtol, double *J, double *grad,
                                        phg_hip_vec *lambda, int32_t *nits)
{
    int rc;
    if ((rc = phg::require_ctx()) || (rc = check_problem(sp, pb)))
        return rc;
    if (pb->dim != 1 || !pb->use_kappa || kappa == nullptr || A == nullptr ||
        u == nullptr || ud == nullptr || grad == nullptr)
        return phg::set_error(PHG_HIP_E_ARG,
                              "phg_hip_adjoint_logkappa: scalar problem with "
                              "kappa, matrix, u, ud and grad required");
    const int64_t n = sp->ndof, ne = sp->nelem;
    const int N = sp->nbas;
    if (A->n != n || u->n != n || ud->n != n || (lambda && lambda->n != n))
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_adjoint_logkappa: size mismatch");
    hipStream_t s = phg::ctx()->stream;
    double *dk = nullptr, *Me = nullptr, *Jel = nullptr, *part = nullptr,
           *dgrad = nullptr;
    AsmWorkspace *wsp = nullptr;     // the space's incidence (scalar rows)
    phg_hip_vec *rhs = nullptr, *lam = nullptr;
    const int nparts = 256;
    std::vector<double> hp(nparts);
    AdjArgs a;
    if ((rc = phg::dmalloc((void **)&dk, (size_t)ne * sizeof(double))) ||
        (rc = phg::upload(dk, kappa, (size_t)ne * sizeof(double))) ||
        (rc = phg::dmalloc((void **)&Me, (size_t)ne * N * sizeof(double))) ||
        (rc = phg::dmalloc((void **)&Jel, (size_t)ne * sizeof(double))) ||
        (rc = phg::dmalloc((void **)&dgrad, (size_t)ne * sizeof(double))) ||
        (rc = phg::dmalloc((void **)&part, (size_t)nparts * sizeof(double))) ||
        (rc = space_incidence(sp, 1, &wsp)) ||
        (rc = phg_hip_vec_create(n, &rhs)))
        goto done;
    if (lambda == nullptr && (rc = phg_hip_vec_create(n, &lam)))
        goto done;
    a.nelem = ne;
    a.verts = sp->d_verts;
    a.elem_verts = sp->d_elem_verts;
    a.map = sp->d_map;
    a.kappa = dk;
    a.u = u->d;
    a.ud = ud->d;
    a.lam = lambda ? lambda->d : lam->d;
    a.Me = Me;
    a.Jel = Jel;
    a.grad = dgrad;
    {
        dim3 g(nblk(ne, 128)), b(128);
        if (sp->order == 1)
            hipLaunchKernelGGL(k_adj_mass<1>, g, b, 0, s, a);
        else if (sp->order == 2)
            hipLaunchKernelGGL(k_adj_mass<2>, g, b, 0, s, a);
        else if (sp->order == 3)
            hipLaunchKernelGGL(k_adj_mass<3>, g, b, 0, s, a);
        else
            hipLaunchKernelGGL(k_adj_mass<4>, g, b, 0, s, a);
    }
    hipLaunchKernelGGL(k_sum_blocks, dim3(nparts), dim3(256), 0, s, ne,
                       (const double *)Jel, part);
    hipLaunchKernelGGL(k_adj_rows, dim3(nblk(n, 256)), dim3(256), 0, s, n, N,
                       (const int64_t *)wsp->inc_off, (const int32_t *)wsp->inc,
                       (const uint8_t *)A->d_bdry_row, (const double *)Me,
                       rhs->d);
    PHG_HIP_CHECK(hipGetLastError());
    if ((rc = phg::download(hp.data(), part, nparts * sizeof(double))))
        goto done;
    if (J) {
        double t = 0.0;
        for (int i = 0; i < nparts; i++)
            t += hp[(size_t)i];
        *J = t;
    }
    {
        // the transposed system: A is symmetric (Dirichlet rows e_i, interior
        // block with the boundary columns removed), so A^T = A
        phg_hip_vec *l = lambda ? lambda : lam;
        PHG_HIP_CHECK(hipMemsetAsync(l->d, 0, (size_t)n * sizeof(double), s));
        double res = 0.0;
        int it = phg_hip_pcg(A, rhs, l, maxit, rtol, rtol, 0.0, &res);
        if (it < 0) {
            rc = it;
            goto done;
        }
        if (nits)
            *nits = it;
    }
    {
        dim3 g(nblk(ne, 128)), b(128);
        if (sp->order == 1)
            hipLaunchKernelGGL(k_adj_grad<1>, g, b, 0, s, a);
        else if (sp->order == 2)
            hipLaunchKernelGGL(k_adj_grad<2>, g, b, 0, s, a);
        else if (sp->order == 3)
            hipLaunchKernelGGL(k_adj_grad<3>, g, b, 0, s, a);
        else
            hipLaunchKernelGGL(k_adj_grad<4>, g, b, 0, s, a);
    }
    PHG_HIP_CHECK(hipGetLastError());
    rc = phg::download(grad, dgrad, (size_t)ne * sizeof(double));
done:
    phg::dfree(dk);
    phg::dfree(Me);
    phg::dfree(Jel);
    phg::dfree(dgrad);
    phg::dfree(part);
    if (rhs)
        phg_hip_vec_destroy(rhs);
    if (lam)
        phg_hip_vec_destroy(lam);
    return rc;
}

//------------------------------------------------------------- DOF norms
namespace {
constexpr int NORM_PARTS = 256;   // partial sums of one fixed-order reduction

// fixed-order sum of v[0..n): NORM_PARTS per-block partials in part (device),
// summed on the host in index order
int sum_device(const double *v, int64_t n, double *part, double *out)
{
    hipStream_t s = phg::ctx()->stream;
    hipLaunchKernelGGL(k_sum_blocks, dim3(NORM_PARTS), dim3(256), 0, s, n, v, part);
    PHG_HIP_CHECK(hipGetLastError());
    double hp[NORM_PARTS];
    int rc = phg::download(hp, part, sizeof(hp));
    if (rc)
        return rc;
    double t = 0.0;
    for (int i = 0; i < NORM_PARTS; i++)
        t += hp[i];
    *out = t;
    return PHG_HIP_OK;
}

int check_norm_args(const phg_hip_space *sp, const phg_hip_vec *v)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (sp == nullptr || v == nullptr || v->n != sp->ndof)
        return phg::set_error(PHG_HIP_E_ARG,
                              "phg_hip_dof_norms: a scalar DOF of the space required");
    return PHG_HIP_OK;
}

// per-element squared norms into dl / dh [nelem], their fixed-order sums
// through part [NORM_PARTS]
int dof_norms_run(const phg_hip_space *sp, const phg_hip_vec *v, double *dl,
                  double *dh, double *part, double *l2, double *h1)
{
    hipStream_t s = phg::ctx()->stream;
    const int64_t ne = sp->nelem;
    dim3 g(nblk(ne, 128)), b(128);
    if (sp->order == 1)
        hipLaunchKernelGGL(k_dof_norms<1>, g, b, 0, s, ne,
                           (const double *)sp->d_verts,
                           (const int32_t *)sp->d_elem_verts,
                           (const int32_t *)sp->d_map, (const double *)v->d, dl, dh);
    else if (sp->order == 2)
        hipLaunchKernelGGL(k_dof_norms<2>, g, b, 0, s, ne,
                           (const double *)sp->d_verts,
                           (const int32_t *)sp->d_elem_verts,
                           (const int32_t *)sp->d_map, (const double *)v->d, dl, dh);
    else if (sp->order == 3)
        hipLaunchKernelGGL(k_dof_norms<3>, g, b, 0, s, ne,
                           (const double *)sp->d_verts,
                           (const int32_t *)sp->d_elem_verts,
                           (const int32_t *)sp->d_map, (const double *)v->d, dl, dh);
    else
        hipLaunchKernelGGL(k_dof_norms<4>, g, b, 0, s, ne,
                           (const double *)sp->d_verts,
                           (const int32_t *)sp->d_elem_verts,
                           (const int32_t *)sp->d_map, (const double *)v->d, dl, dh);
    PHG_HIP_CHECK(hipGetLastError());
    double a = 0.0, c = 0.0;
    int rc;
    if ((rc = sum_device(dl, ne, part, &a)) || (rc = sum_device(dh, ne, part, &c)))
        return rc;
    if (l2)
        *l2 = sqrt(a);
    if (h1)
        *h1 = sqrt(c);
    return PHG_HIP_OK;
}

// Scratch of phg_hip_dof_norms_noalloc: the per-element values of both norms
// and the partial sums, for meshes of up to NOALLOC_MAX_ELEM elements
// (32 MiB + 2 KiB, held by every GPU the library runs on until the process
// ends).  Host code never passes the array's name to a kernel: the address
// of each device's copy is taken once per device at phg_hip_init.
constexpr int64_t NOALLOC_MAX_ELEM = int64_t(1) << 21;
constexpr int64_t NOALLOC_DOUBLES = 2 * NOALLOC_MAX_ELEM + NORM_PARTS;
constexpr int NOALLOC_MAX_DEVICES = 64;
double *g_scratch[NOALLOC_MAX_DEVICES] = {};
}  // namespace

static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace phg {
// called by phg_hip_init with `device` current
int noalloc_init(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return set_error(PHG_HIP_E_ARG, "device %d beyond the scratch table", device);
    if (g_scratch[device] != nullptr)
        return PHG_HIP_OK;
    void *ptr = nullptr;
    PHG_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[device] = (double *)ptr;
    return PHG_HIP_OK;
}
}  // namespace phg

extern "C" int phg_hip_dof_norms(phg_hip_space *sp, const phg_hip_vec *v,
                                 double *l2, double *h1)
{
    int rc;
    if ((rc = check_norm_args(sp, v)))
        return rc;
    const int64_t ne = sp->nelem;
    double *buf = nullptr;
    if ((rc = phg::dmalloc((void **)&buf, (size_t)(2 * ne + NORM_PARTS) * sizeof(double))))
        return rc;
    rc = dof_norms_run(sp, v, buf, buf + ne, buf + 2 * ne, l2, h1);
    phg::dfree(buf);
    return rc;
}

extern "C" int phg_hip_dof_norms_noalloc(phg_hip_space *sp, const phg_hip_vec *v,
                                         double *l2, double *h1)
{
    int rc;
    if ((rc = check_norm_args(sp, v)))
        return rc;
    const int64_t ne = sp->nelem;
    if (ne > NOALLOC_MAX_ELEM)
        return phg::set_error(PHG_HIP_E_ARG,
                              "phg_hip_dof_norms_noalloc: %lld elements, at most %lld",
                              (long long)ne, (long long)NOALLOC_MAX_ELEM);
    double *buf = g_scratch[phg::ctx()->device];
    if (buf == nullptr)
        return phg::set_error(PHG_HIP_E_STATE, "no scratch address for device %d",
                              phg::ctx()->device);
    return dof_norms_run(sp, v, buf, buf + ne, buf + 2 * ne, l2, h1);
}

//------------------------------------------------------- CSR from the host
extern "C" int phg_hip_mat_from_csr(int64_t n, const int64_t *rowptr,
                                    const int32_t *cols, const double *vals,
                                    const double *rhs, const uint8_t *bdry,
                                    int32_t eliminated, phg_hip_mat **Aout)
{
    int rc;
    if ((rc = phg::require_ctx()))
        return rc;
    if (n <= 0 || rowptr == nullptr || cols == nullptr || vals == nullptr ||
        rhs == nullptr || bdry == nullptr || Aout == nullptr)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_mat_from_csr: bad args");
    const int64_t nnz = rowptr[n];
    if (rowptr[0] != 0 || nnz < 0 || nnz >= (int64_t)INT32_MAX)
        return phg::set_error(PHG_HIP_E_ARG, "phg_hip_mat_from_csr: bad row pointers");
    std::vector<int32_t> rp32((size_t)n + 1);
    for (int64_t i = 0; i <= n; i++) {
        if (i > 0 && rowptr[i] < rowptr[i - 1])
            return phg::set_error(PHG_HIP_E_ARG, "row pointers decrease at %lld",
                                  (long long)i);
        rp32[(size_t)i] = (int32_t)rowptr[i];
    }
    for (int64_t j = 0; j < nnz; j++)
        if (cols[j] < 0 || cols[j] >= n)
            return phg::set_error(PHG_HIP_E_ARG, "column %d out of range at %lld",
                                  cols[j], (long long)j);
    hipStream_t s = phg::ctx()->stream;
    phg_hip_mat *A = new phg_hip_mat();
    A->n = n;
    A->nnz = nnz;
    A->dim = 1;
    A->ndof = n;
    if ((rc = phg::dmalloc((void **)&A->d_rowptr, (size_t)(n + 1) * 4)) ||
        (rc = phg::dmalloc((void **)&A->d_cols, (size_t)(nnz + 1) * 4)) ||
        (rc = phg::dmalloc((void **)&A->d_vals, (size_t)(nnz + 1) * 8)) ||
        (rc = phg::dmalloc((void **)&A->d_rhs, (size_t)n * 8)) ||
        (rc = phg::dmalloc((void **)&A->d_bdry_row, (size_t)n)) ||
        (rc = phg::upload(A->d_rowptr, rp32.data(), (size_t)(n + 1) * 4)) ||
        (rc = phg::upload(A->d_cols, cols, (size_t)nnz * 4)) ||
        (rc = phg::upload(A->d_vals, vals, (size_t)nnz * 8)) ||
        (rc = phg::upload(A->d_rhs, rhs, (size_t)n * 8)) ||
        (rc = phg::upload(A->d_bdry_row, bdry, (size_t)n)))
        goto fail;
    if (eliminated) {
        if ((rc = phg::dmalloc((void **)&A->d_diag, (size_t)n * 8)))
            goto fail;
        hipLaunchKernelGGL(k_diag, dim3(nblk(n, 256)), dim3(256), 0, s, n,
                           (const int32_t *)A->d_rowptr, (const int32_t *)A->d_cols,
                           (const double *)A->d_vals, A->d_diag);
        if (hipGetLastError() != hipSuccess) {
            rc = phg::set_error(PHG_HIP_E_HIP, "diagonal kernel");
            goto fail;
        }
        A->eliminated = 1;
        if ((rc = phg::build_row_blocks(A)) || (rc = phg::build_solver_copy(A)))
            goto fail;
    }
    *Aout = A;
    return PHG_HIP_OK;
fail:
    phg_hip_mat_destroy(A);
    return rc;
}

//================================================================ refill
// Re-assembly of an eliminated matrix for new coefficient values (kappa, the
// load, the Dirichlet values): the kernels of phg_hip_assemble and
// phg_hip_mat_eliminate -- element pass, PHG's row insertion, scans,
// compaction, elimination -- run into buffers the plan owns and into the
// matrix's PHG-order arrays, which the plan re-allocates to a capacity that
// bounds every pattern (each row's incidence times the element size).  The
// pattern may change with the values (PHG drops exact zeros, matvec.c:5460,
// and whether a cancelling sum rounds to zero depends on kappa), so the
// scans run on the device and nothing waits for their totals:
// refill_enqueue only launches kernels -- no allocation, no host
// synchronisation, no memset -- so a HIP graph can capture it.  The new
// entry count is left at nnz_dev.
namespace {

__global__ void k_refill_reset(int32_t *ovf)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
        ovf[0] = 0;
}

// out[0..n] = exclusive scan of in[0..n-1] with the total at out[n] (the
// kernels of scan_i32 on a caller-provided block-sum buffer of nb + 1)
void scan_i32_enqueue(int64_t n, const int32_t *in, int64_t *out, int64_t *bsum,
                      hipStream_t s)
{
    const int64_t nb = (n + 1 + 1023) / 1024;
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(1024), 0, s, n, in, out, bsum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, s, nb, bsum, bsum + nb);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(1024), 0, s, n, out,
                       (const int64_t *)bsum, (const int64_t *)(bsum + nb));
}

}  // namespace

namespace phg {

struct RefillPlan {
    phg_hip_mat *A = nullptr;
    const phg_hip_space *sp = nullptr;
    phg_hip_problem pb;
    ElemArgs ea;
    WalkArgs wk;
    int64_t n = 0, cap = 0;
    // owned device buffers: element matrices and loads, incidence, row-walk
    // scratch, the matrix before elimination (row pointers in both widths,
    // columns, values, right-hand side), the eliminated row counts and
    // pointers, and the scans' block sums
    double *K = nullptr, *B = nullptr;
    int32_t *cnt = nullptr, *inc = nullptr;
    int64_t *inc_off = nullptr, *scr_off = nullptr;
    int32_t *scols = nullptr, *count = nullptr, *ovf = nullptr, *ecount = nullptr;
    double *svals = nullptr;
    int64_t *rp64 = nullptr, *nrp64 = nullptr, *bsum = nullptr;
    int32_t *rp32 = nullptr, *cols = nullptr;
    double *vals = nullptr, *rhs = nullptr;
    ~RefillPlan()
    {
        dfree(K); dfree(B); dfree(cnt); dfree(inc); dfree(inc_off); dfree(scr_off);
        dfree(scols); dfree(count); dfree(ovf); dfree(ecount); dfree(svals); dfree(rp64);
        dfree(nrp64); dfree(bsum); dfree(rp32); dfree(cols); dfree(vals); dfree(rhs);
    }
};

void refill_plan_destroy(RefillPlan *p) { delete p; }

const int64_t *refill_nnz(const RefillPlan *p) { return p->nrp64 + p->n; }

int refill_enqueue(const RefillPlan *P, hipStream_t s)
{
    const int64_t n = P->n;
    phg_hip_mat *A = P->A;
    int rc;
    if ((rc = launch_element(P->sp, &P->pb, P->ea, s)))
        return rc;
    hipLaunchKernelGGL(k_refill_reset, dim3(1), dim3(64), 0, s, P->ovf);
    if ((rc = launch_row_walk(P->wk, s)))
        return rc;
    scan_i32_enqueue(n, P->count, P->rp64, P->bsum, s);
    hipLaunchKernelGGL(k_compact, dim3(nblk(n, RPB)), dim3(256), 0, s, n,
                       (const int64_t *)P->scr_off, (const int32_t *)P->scols,
                       (const double *)P->svals, (const int64_t *)P->rp64, P->cols, P->vals);
    hipLaunchKernelGGL(k_to_int32, dim3(nblk(n + 1, 256)), dim3(256), 0, s, n,
                       (const int64_t *)P->rp64, P->rp32);
    hipLaunchKernelGGL(k_elim_count, dim3(nblk(n, RPB)), dim3(256), 0, s, n,
                       (const int32_t *)P->rp32, (const int32_t *)P->cols,
                       (const uint8_t *)A->d_bdry_row, P->ecount);
    scan_i32_enqueue(n, P->ecount, P->nrp64, P->bsum, s);
    hipLaunchKernelGGL(k_elim_fill, dim3(nblk(n, RPB)), dim3(256), 0, s, n,
                       (const int32_t *)P->rp32, (const int32_t *)P->cols,
                       (const double *)P->vals, (const double *)P->rhs,
                       (const uint8_t *)A->d_bdry_row, (const int64_t *)P->nrp64, A->d_cols,
                       A->d_vals, A->d_rhs, A->d_diag);
    hipLaunchKernelGGL(k_to_int32, dim3(nblk(n + 1, 256)), dim3(256), 0, s, n,
                       (const int64_t *)P->nrp64, A->d_rowptr);
    PHG_HIP_CHECK(hipGetLastError());
    return PHG_HIP_OK;
}

int refill_plan_create(phg_hip_space *sp, const phg_hip_problem *pb, phg_hip_mat *A,
                       const double *d_kappa, const double *d_f, const double *d_g,
                       RefillPlan **out)
{
    int rc;
    *out = nullptr;
    if ((rc = check_problem(sp, pb)))
        return rc;
    if (A == nullptr || !A->eliminated || A->space_uid != sp->uid || A->dim != pb->dim ||
        A->n != sp->ndof * pb->dim || A->d_bdry_row == nullptr)
        return set_error(PHG_HIP_E_ARG, "refill: the matrix must be the eliminated matrix "
                                        "phg_hip_assemble made on this space for this problem");
    if ((pb->use_kappa && d_kappa == nullptr) || d_f == nullptr || d_g == nullptr)
        return set_error(PHG_HIP_E_ARG, "refill: kappa (when used), f and g are required");
    hipStream_t s = ctx()->stream;
    const int N = sp->nbas, dim = pb->dim, M = N * dim;
    const int64_t n = A->n, ne = sp->nelem;
    std::unique_ptr<RefillPlan> P(new RefillPlan());
    P->A = A;
    P->sp = sp;
    P->pb = *pb;
    P->n = n;
    int64_t nscr = 0;
    const int64_t nb = (n + 1 + 1023) / 1024;
    if ((rc = dmalloc((void **)&P->K, (size_t)ne * M * M * sizeof(double) + 8)) ||
        (rc = dmalloc((void **)&P->B, (size_t)ne * M * sizeof(double) + 8)) ||
        (rc = build_incidence(sp, dim, &P->cnt, &P->inc_off, &P->inc)) ||
        (rc = scratch_offsets(n, P->cnt, M, &P->scr_off, &nscr)) ||
        (rc = dmalloc((void **)&P->scols, (size_t)(nscr + 1) * sizeof(int32_t))) ||
        (rc = dmalloc((void **)&P->svals, (size_t)(nscr + 1) * sizeof(double))) ||
        (rc = dmalloc((void **)&P->count, (size_t)(n + 1) * sizeof(int32_t))) ||
        (rc = dmalloc((void **)&P->ecount, (size_t)(n + 1) * sizeof(int32_t))) ||
        (rc = dmalloc((void **)&P->ovf, (size_t)(n + 1) * sizeof(int32_t))) ||
        (rc = dmalloc((void **)&P->rhs, (size_t)(n + 1) * sizeof(double))) ||
        (rc = dmalloc((void **)&P->rp64, (size_t)(n + 1) * sizeof(int64_t))) ||
        (rc = dmalloc((void **)&P->nrp64, (size_t)(n + 1) * sizeof(int64_t))) ||
        (rc = dmalloc((void **)&P->bsum, (size_t)(nb + 1) * sizeof(int64_t))) ||
        (rc = dmalloc((void **)&P->rp32, (size_t)(n + 1) * sizeof(int32_t))) ||
        (rc = dmalloc((void **)&P->cols, (size_t)(nscr + 1) * sizeof(int32_t))) ||
        (rc = dmalloc((void **)&P->vals, (size_t)(nscr + 1) * sizeof(double))))
        return rc;
    if (nscr >= (int64_t)INT32_MAX)
        return set_error(PHG_HIP_E_ARG, "refill: %lld entries bound exceeds int32 CSR",
                         (long long)nscr);
    // the matrix's PHG-order columns and values at the capacity
    if (A->nnz > nscr)
        return set_error(PHG_HIP_E_STATE, "refill: the matrix exceeds its pattern bound");
    int32_t *ncols = nullptr;
    double *nvals = nullptr;
    if ((rc = dmalloc((void **)&ncols, (size_t)(nscr + 1) * sizeof(int32_t))) ||
        (rc = dmalloc((void **)&nvals, (size_t)(nscr + 1) * sizeof(double)))) {
        dfree(ncols);
        return rc;
    }
    if (hipMemcpyAsync(ncols, A->d_cols, (size_t)A->nnz * sizeof(int32_t),
                       hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpyAsync(nvals, A->d_vals, (size_t)A->nnz * sizeof(double),
                       hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        dfree(ncols);
        dfree(nvals);
        return hip_fail(hipGetLastError(), "refill: matrix copy", __FILE__, __LINE__);
    }
    P->cap = nscr;
    ElemArgs &a = P->ea;
    a.nelem = ne;
    a.verts = sp->d_verts;
    a.elem_verts = sp->d_elem_verts;
    a.map = sp->d_map;
    a.kappa = pb->use_kappa ? d_kappa : nullptr;
    a.f = d_f;
    a.use_kappa = pb->use_kappa;
    a.mass_coef = pb->mass_coef;
    a.rhs_order = pb->rhs_order;
    a.E = pb->E;
    a.nu = pb->nu;
    a.K = P->K;
    a.B = P->B;
    WalkArgs &w = P->wk;
    w.n = n;
    w.N = N;
    w.dim = dim;
    w.map = sp->d_map;
    w.inc_off = P->inc_off;
    w.inc = P->inc;
    w.K = P->K;
    w.B = P->B;
    w.bdry = A->d_bdry_row;
    w.g = d_g;
    w.scr_off = P->scr_off;
    w.scols = P->scols;
    w.svals = P->svals;
    w.count = P->count;
    w.rhs = P->rhs;
    w.ovf = P->ovf;
    // hand the capacity arrays to the matrix (the solver copy aliases the
    // PHG-order arrays when the matrix is not renumbered: it is dropped and
    // rebuilt when next needed)
    free_solver_copy(A);
    free_sell(A);
    A->q_stale = 1;
    dfree(A->d_cols);
    dfree(A->d_vals);
    A->d_cols = ncols;
    A->d_vals = nvals;
    dfree(A->d_rblk);
    A->d_rblk = nullptr;
    A->nrblk = 0;
    *out = P.release();
    return PHG_HIP_OK;
}

}  // namespace phg
