/* Oracle: CPU restatement of PHG's finite-element hot path (test
 * infrastructure only; see phg_oracle.h).  Built with -ffp-contract=off. */
#include "phg_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void *xmalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (p == NULL) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}

/*------------------------------ quadrature --------------------------------*/
/* quad.c:1799-2000 static 3D rules, permutations of
 * include/phg/quad-permu.h:83-120 in the same point order. */

#define NQMAX 46
typedef struct {
    int np;
    double pts[NQMAX][4];
    double wts[NQMAX];
} qrule;

static qrule rules[9];
static int rules_ready = 0;

static void add31(qrule *q, double a, double w)
{
    double b = 1.0 - 3.0 * a;
    double P[4][4] = {{b, a, a, a}, {a, b, a, a}, {a, a, b, a}, {a, a, a, b}};
    for (int i = 0; i < 4; i++) {
        memcpy(q->pts[q->np], P[i], sizeof(P[i]));
        q->wts[q->np++] = w;
    }
}

static void add4(qrule *q, double w)
{
    q->pts[q->np][0] = q->pts[q->np][1] = q->pts[q->np][2] = q->pts[q->np][3] = .25;
    q->wts[q->np++] = w;
}

static void add22(qrule *q, double a, double w)
{
    double h = 0.5 - a;
    double P[6][4] = {{a, a, h, h}, {a, h, a, h}, {a, h, h, a},
                      {h, a, h, a}, {h, a, a, h}, {h, h, a, a}};
    for (int i = 0; i < 6; i++) {
        memcpy(q->pts[q->np], P[i], sizeof(P[i]));
        q->wts[q->np++] = w;
    }
}

static void add211(qrule *q, double a, double b, double w)
{
    double c = 1.0 - a - a - b;
    double P[12][4] = {{a, a, b, c}, {a, a, c, b}, {a, b, a, c}, {a, b, c, a},
                       {a, c, a, b}, {a, c, b, a}, {b, a, a, c}, {b, a, c, a},
                       {b, c, a, a}, {c, a, a, b}, {c, a, b, a}, {c, b, a, a}};
    for (int i = 0; i < 12; i++) {
        memcpy(q->pts[q->np], P[i], sizeof(P[i]));
        q->wts[q->np++] = w;
    }
}

static void init_rules(void)
{
    qrule *q;
    if (rules_ready)
        return;
    memset(rules, 0, sizeof(rules));
    /* order 1 (quad.c:1809): one point, weight 1 */
    q = &rules[1];
    q->pts[0][0] = q->pts[0][1] = q->pts[0][2] = q->pts[0][3] = .25;
    q->wts[0] = 1.;
    q->np = 1;
    /* order 2 (quad.c:1827) */
    add31(&rules[2], .13819660112501051517954131656343619, .25);
    /* order 3 (quad.c:1858) */
    add31(&rules[3], .32805469671142664733580581998119743,
          .13852796651186214232361769837564129);
    add31(&rules[3], .10695227393293068277170204157061650,
          .11147203348813785767638230162435871);
    /* order 4 (quad.c:1904) */
    add31(&rules[4], .09273525031089122628655892066032137,
          .07349304311636194934358694586367885);
    add31(&rules[4], .31088591926330060975814749494040332,
          .11268792571801585036501492847638892);
    add22(&rules[4], .04550370412564965000000000000000000,
          .04254602077708146686093208377328816);
    /* order 5 (quad.c:1955) */
    add31(&rules[5], .31088591926330060979734573376345783,
          .11268792571801585079918565233328633);
    add31(&rules[5], .09273525031089122640232391373703061,
          .07349304311636194954371020548632750);
    add22(&rules[5], .04550370412564964949188052627933943,
          .04254602077708146643806942812025744);
    /* order 6 (quad.c:1983) */
    add31(&rules[6], .21460287125915202928883921938628499,
          .03992275025816749209969062755747998);
    add31(&rules[6], .04067395853461135311557944895641006,
          .01007721105532064294801323744593686);
    add31(&rules[6], .32233789014227551034399447076249213,
          .05535718154365472209515327785372602);
    add211(&rules[6], .06366100187501752529923552760572698,
           .60300566479164914136743113906093969, 27. / 560.);
    /* order 7 (quad.c:1990) */
    add4(&rules[7], .09548528946413084886057843611722638);
    add31(&rules[7], .31570114977820279942342999959331149,
          .04232958120996702907628617079854674);
    add22(&rules[7], .05048982259839636876305382298656247,
          .03189692783285757993427482408294246);
    add211(&rules[7], .18883383102600104773643110385458576,
           .57517163758700002348324157702230752, .03720713072833462136961556119148112);
    add211(&rules[7], .02126547254148324598883610149981994,
           .81083024109854856111810537984823239, .00811077082990334156610343349109654);
    /* order 8 (quad.c:2017) */
    add31(&rules[8], .03967542307038990126507132953938949,
          .00639714777990232132145142033517302);
    add31(&rules[8], .31448780069809631378416056269714830,
          .04019044802096617248816115847981783);
    add31(&rules[8], .10198669306270330000000000000000000,
          .02430797550477032117486910877192260);
    add31(&rules[8], .18420369694919151227594641734890918,
          .05485889241369744046692412399039144);
    add22(&rules[8], .06343628775453989240514123870189827,
          .03571961223409918246495096899661762);
    add211(&rules[8], .02169016206772800480266248262493018,
           .71993192203946593588943495335273478, .00718319069785253940945110521980376);
    add211(&rules[8], .20448008063679571424133557487274534,
           .58057719012880922417539817139062041, .01637218194531911754093813975611913);
    /* order 0 falls to the order-1 rule (quad.c:2807 init_table) */
    rules[0] = rules[1];
    rules_ready = 1;
}

int orc_quad3d(int order, const double **pts, const double **wts)
{
    init_rules();
    if (order < 0)
        order = 0;
    if (order > 8)
        return -1;
    *pts = &rules[order].pts[0][0];
    *wts = rules[order].wts;
    return rules[order].np;
}

/*------------------------------- bases ------------------------------------*/

static const int EV[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
static const int FV[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};

void orc_bas(int p, const double *lam, double *val)
{
    double L0 = lam[0], L1 = lam[1], L2 = lam[2], L3 = lam[3];
    if (p == 1) {                         /* lagrange.c:80 */
        memcpy(val, lam, 4 * sizeof(double));
    }
    else if (p == 2) {                    /* lagrange.c:160 full block */
        double *q = val;
        *(q++) = L0 * (L0 + L0 - 1.);
        *(q++) = L1 * (L1 + L1 - 1.);
        *(q++) = L2 * (L2 + L2 - 1.);
        *(q++) = L3 * (L3 + L3 - 1.);
        *(q++) = 4. * L0 * L1;
        *(q++) = 4. * L0 * L2;
        *(q++) = 4. * L0 * L3;
        *(q++) = 4. * L1 * L2;
        *(q++) = 4. * L1 * L3;
        *q = 4. * L2 * L3;
    }
    else if (p == 4) {                    /* lagrange.c:757 full block */
        double L0T4M1 = 4. * L0 - 1., L1T4M1 = 4. * L1 - 1., L2T4M1 = 4. * L2 - 1.,
               L3T4M1 = 4. * L3 - 1.;
        double t0, t1, t2, t3, d01, d02, d03, d12, d13, d23, d123, d023, d013, d012;
        double *q = val;
        *(q++) = L0T4M1 * ((2. / 3.) * L0 - (1. / 3.)) * (L0T4M1 - 2.) * L0;
        *(q++) = L1T4M1 * ((2. / 3.) * L1 - (1. / 3.)) * (L1T4M1 - 2.) * L1;
        *(q++) = L2T4M1 * ((2. / 3.) * L2 - (1. / 3.)) * (L2T4M1 - 2.) * L2;
        *(q++) = L3T4M1 * ((2. / 3.) * L3 - (1. / 3.)) * (L3T4M1 - 2.) * L3;
        t0 = L0T4M1 * ((32. / 3.) * L0 - (16. / 3.));
        t1 = L1T4M1 * ((32. / 3.) * L1 - (16. / 3.));
        t2 = L2T4M1 * ((32. / 3.) * L2 - (16. / 3.));
        t3 = L3T4M1 * ((32. / 3.) * L3 - (16. / 3.));
        d01 = L0 * L1;
        d02 = L0 * L2;
        d03 = L0 * L3;
        d12 = L1 * L2;
        d13 = L1 * L3;
        d23 = L2 * L3;
        /* edges 0-1, 0-2, 0-3, 1-2, 1-3, 2-3: near the first vertex, middle,
         * near the second */
        *(q++) = t0 * d01; *(q++) = 4. * L0T4M1 * L1T4M1 * d01; *(q++) = t1 * d01;
        *(q++) = t0 * d02; *(q++) = 4. * L0T4M1 * L2T4M1 * d02; *(q++) = t2 * d02;
        *(q++) = t0 * d03; *(q++) = 4. * L0T4M1 * L3T4M1 * d03; *(q++) = t3 * d03;
        *(q++) = t1 * d12; *(q++) = 4. * L1T4M1 * L2T4M1 * d12; *(q++) = t2 * d12;
        *(q++) = t1 * d13; *(q++) = 4. * L1T4M1 * L3T4M1 * d13; *(q++) = t3 * d13;
        *(q++) = t2 * d23; *(q++) = 4. * L2T4M1 * L3T4M1 * d23; *(q++) = t3 * d23;
        L0T4M1 *= 32.;
        L1T4M1 *= 32.;
        L2T4M1 *= 32.;
        L3T4M1 *= 32.;
        d123 = L1 * d23;
        d023 = L0 * d23;
        d013 = L0 * d13;
        d012 = L0 * d12;
        /* faces 0..3, one node near each face vertex in face-vertex order */
        *(q++) = L1T4M1 * d123; *(q++) = L2T4M1 * d123; *(q++) = L3T4M1 * d123;
        *(q++) = L0T4M1 * d023; *(q++) = L2T4M1 * d023; *(q++) = L3T4M1 * d023;
        *(q++) = L0T4M1 * d013; *(q++) = L1T4M1 * d013; *(q++) = L3T4M1 * d013;
        *(q++) = L0T4M1 * d012; *(q++) = L1T4M1 * d012; *(q++) = L2T4M1 * d012;
        *q = 256. * L0 * d123;            /* element */
    }
    else {                                /* lagrange.c:392 full block */
        double L03 = 3. * L0, L13 = 3. * L1, L23 = 3. * L2, L33 = 3. * L3, d;
        double *q = val;
        *(q++) = .5 * (L03 - 1.) * (L03 - 2.) * L0;
        *(q++) = .5 * (L13 - 1.) * (L13 - 2.) * L1;
        *(q++) = .5 * (L23 - 1.) * (L23 - 2.) * L2;
        *(q++) = .5 * (L33 - 1.) * (L33 - 2.) * L3;
        *(q++) = 4.5 * (L03 - 1.) * (d = L0 * L1);
        *(q++) = 4.5 * (L13 - 1.) * d;
        *(q++) = 4.5 * (L03 - 1.) * (d = L0 * L2);
        *(q++) = 4.5 * (L23 - 1.) * d;
        *(q++) = 4.5 * (L03 - 1.) * (d = L0 * L3);
        *(q++) = 4.5 * (L33 - 1.) * d;
        *(q++) = 4.5 * (L13 - 1.) * (d = L1 * L2);
        *(q++) = 4.5 * (L23 - 1.) * d;
        *(q++) = 4.5 * (L13 - 1.) * (d = L1 * L3);
        *(q++) = 4.5 * (L33 - 1.) * d;
        *(q++) = 4.5 * (L23 - 1.) * (d = L2 * L3);
        *(q++) = 4.5 * (L33 - 1.) * d;
        *(q++) = 27. * L1 * L2 * L3;
        *(q++) = 27. * L0 * L2 * L3;
        *(q++) = 27. * L0 * L1 * L3;
        *q = 27. * L0 * L1 * L2;
    }
}

int orc_nbas(int p)
{
    return p == 1 ? 4 : (p == 2 ? 10 : (p == 3 ? 20 : 35));
}

void orc_grad_lambda(int p, const double *lam, double *grad)
{
    int nb = orc_nbas(p);
    memset(grad, 0, (size_t)nb * 4 * sizeof(double));
    if (p == 1) {                         /* lagrange.c:111 */
        for (int k = 0; k < 4; k++)
            grad[4 * k + k] = 1.0;
        return;
    }
    if (p == 2) {                         /* lagrange.c:248 */
        for (int k = 0; k < 4; k++)
            grad[4 * k + k] = 4. * lam[k] - 1.;
        for (int k = 0; k < 6; k++) {
            int a = EV[k][0], b = EV[k][1];
            grad[4 * (4 + k) + a] = 4. * lam[b];
            grad[4 * (4 + k) + b] = 4. * lam[a];
        }
        return;
    }
    if (p == 4) {                         /* lagrange.c:1085 full block */
        for (int k = 0; k < 4; k++) {
            double L0 = lam[k];
            grad[4 * k + k] = ((8. / 3.) * L0 - 1.) * (L0 * (16. * L0 - 12.) + 1.);
        }
        for (int k = 0; k < 6; k++) {
            /* edge (a, b): near a (TEMP0 a, b), middle (TEMP1), near b
             * (TEMP0 b, a) */
            int ab[3][2] = {{EV[k][0], EV[k][1]}, {EV[k][0], EV[k][1]},
                            {EV[k][1], EV[k][0]}};
            for (int t = 0; t < 3; t++) {
                int v0 = ab[t][0], v1 = ab[t][1];
                double L0 = lam[v0], L1 = lam[v1], *g = grad + 4 * (4 + 3 * k + t);
                if (t == 1) {
                    g[v0] = (32. * L0 - 4.) * (4. * L1 - 1.) * L1;
                    g[v1] = L0 * (4. * L0 - 1.) * (32. * L1 - 4.);
                }
                else {
                    g[v0] = (L0 * ((384. / 3.) * L0 - 64.) + (16. / 3.)) * L1;
                    g[v1] = L0 * ((64. / 3.) * L0 - (16. / 3.)) * (2. * L0 - 1.);
                }
            }
        }
        for (int f = 0; f < 4; f++) {
            /* face (a, b, c): nodes (a, b, c), (b, c, a), (c, a, b) */
            int a = FV[f][0], b = FV[f][1], c = FV[f][2];
            int cyc[3][3] = {{a, b, c}, {b, c, a}, {c, a, b}};
            for (int t = 0; t < 3; t++) {
                int v1 = cyc[t][0], v2 = cyc[t][1], v3 = cyc[t][2];
                double L1 = lam[v1], L2 = lam[v2], L3 = lam[v3];
                double *g = grad + 4 * (22 + 3 * f + t);
                g[v1] = (256. * L1 - 32.) * L2 * L3;
                g[v2] = (128. * L1 - 32.) * L1 * L3;
                g[v3] = (128. * L1 - 32.) * L1 * L2;
            }
        }
        grad[4 * 34 + 0] = 256. * lam[1] * lam[2] * lam[3];
        grad[4 * 34 + 1] = 256. * lam[0] * lam[2] * lam[3];
        grad[4 * 34 + 2] = 256. * lam[0] * lam[1] * lam[3];
        grad[4 * 34 + 3] = 256. * lam[0] * lam[1] * lam[2];
        return;
    }
    /* lagrange.c:538 */
    for (int k = 0; k < 4; k++)
        grad[4 * k + k] = (13.5 * lam[k] - 9.) * lam[k] + 1.;
    for (int k = 0; k < 6; k++) {
        int a = EV[k][0], b = EV[k][1];
        double *g0 = grad + 4 * (4 + 2 * k), *g1 = g0 + 4;
        g0[a] = (27. * lam[a] - 4.5) * lam[b];
        g0[b] = 4.5 * lam[a] * (3. * lam[a] - 1.);
        g1[a] = 4.5 * lam[b] * (3. * lam[b] - 1.);
        g1[b] = (27. * lam[b] - 4.5) * lam[a];
    }
    for (int k = 0; k < 4; k++) {
        int a = FV[k][0], b = FV[k][1], c = FV[k][2];
        double *g = grad + 4 * (16 + k);
        g[a] = 27. * lam[b] * lam[c];
        g[b] = 27. * lam[a] * lam[c];
        g[c] = 27. * lam[a] * lam[b];
    }
}

/*------------------------------- geometry ---------------------------------*/

double orc_volume(const double x[4][3])
{
    /* geom.c:46 _phg_compute_elem_data */
    static const double oned6 = 1.0 / (double)6.0;
    double x0 = x[0][0], y0 = x[0][1], z0 = x[0][2];
    double x1 = x[1][0] - x0, y1 = x[1][1] - y0, z1 = x[1][2] - z0;
    double x2 = x[2][0] - x0, y2 = x[2][1] - y0, z2 = x[2][2] - z0;
    double x3 = x[3][0] - x0, y3 = x[3][1] - y0, z3 = x[3][2] - z0;
    return fabs(x1 * y2 * z3 + x2 * y3 * z1 + y1 * z2 * x3 -
                (z1 * y2 * x3 + y1 * x2 * z3 + z2 * y3 * x1)) * oned6;
}

void orc_jacobian(const double x[4][3], double *J)
{
    /* geom.c:103 compute_jacobian: Gauss-Jordan on [x y z 1 | I] */
    const int n = 4;
    double d, t, a[4][8];
    int i, j, k;
    for (j = 0; j < n; j++)
        for (i = 0; i < n; i++)
            a[j][n + i] = (i == j) ? 1.0 : 0.0;
    for (j = 0; j < n; j++) {
        for (i = 0; i < 3; i++)
            a[j][i] = x[j][i];
        a[j][3] = 1.0;
    }
    for (j = 0; j < n; j++) {
        d = fabs(a[j][j]);
        k = j;
        for (i = j + 1; i < n; i++)
            if (d < (t = fabs(a[i][j]))) {
                d = t;
                k = i;
            }
        if (k != j)
            for (i = j; i < n + n; i++) {
                t = a[j][i];
                a[j][i] = a[k][i];
                a[k][i] = t;
            }
        d = 1.0 / ((t = a[j][j]) == 0. ? 1. : t);
        for (i = j + 1; i < n + n; i++)
            a[j][i] *= d;
        for (i = j + 1; i < n; i++) {
            d = a[i][j];
            for (k = j + 1; k < n + n; k++)
                a[i][k] -= d * a[j][k];
        }
    }
    for (j = n - 2; j >= 0; j--)
        for (i = j + 1; i < n; i++) {
            d = a[j][i];
            for (k = n; k < n + n; k++)
                a[j][k] -= d * a[i][k];
        }
    for (j = 0; j < n; j++)
        for (i = 0; i < n; i++)
            *(J++) = a[i][j + n];
}

/*--------------------------- element integrals ----------------------------*/

/* quad.c:3314 phgQuadGetBasisGradient for basis k at point q:
 * P1 copies Jacobian row k; vertex bases of P2/P3 scale row k by the
 * diagonal lambda derivative; other bases take the 4-term product. */
static void xgrad(int p, int k, const double *G /* [4] lambda grad of k */,
                  const double *J, double *out)
{
    if (p == 1) {
        out[0] = J[4 * k + 0];
        out[1] = J[4 * k + 1];
        out[2] = J[4 * k + 2];
        return;
    }
    if (k < 4) {
        double d = G[k];
        out[0] = d * J[4 * k + 0];
        out[1] = d * J[4 * k + 1];
        out[2] = d * J[4 * k + 2];
        return;
    }
    for (int c = 0; c < 3; c++)
        out[c] = G[0] * J[0 * 4 + c] + G[1] * J[1 * 4 + c] +
                 G[2] * J[2 * 4 + c] + G[3] * J[3 * 4 + c];
}

#define NBMAX 35

void orc_element(const orc_problem *pb, const double x[4][3], double kappa,
                 const double *f_loc, double *K, double *b)
{
    int p = pb->p, N = orc_nbas(p), dim = pb->dim;
    double J[16], vol;
    const double *qp, *qw;
    int nq;
    static double gx[NQMAX][NBMAX][3];   /* x-gradients per point/basis */
    static double bv[NQMAX][NBMAX];      /* basis values per point */
    double glam[NBMAX * 4];

    orc_jacobian(x, J);
    vol = orc_volume(x);

    if (dim == 1) {
        /* stiffness: order (p-1)+(p-1) [+0 for a P0 coefficient]
         * (quad.c:5271-5279).  With the quadrature cache
         * (examples/poisson.c:126 phgQCSetQuad) one rule of order 2p serves
         * stiffness and load; quad-cache.c:1016 phgQCIntegrate_ sums, per
         * point, d = sum_k g_j[k] g_i[k] (k = 0..2, from 0) and then
         * val += d w, times the volume -- the operation order below.  Its
         * x-gradients (quad-cache.c:85 grd) take the 4-term product for
         * every basis; for P1 and vertex bases the other lambda derivatives
         * are exact zeros, so xgrad's shortcuts give the same values. */
        nq = orc_quad3d(pb->qcache ? 2 * p : 2 * (p - 1), &qp, &qw);
        for (int q = 0; q < nq; q++) {
            orc_grad_lambda(p, qp + 4 * q, glam);
            for (int k = 0; k < N; k++)
                xgrad(p, k, glam + 4 * k, J, gx[q][k]);
        }
        for (int i = 0; i < N; i++)
            for (int j = 0; j <= i; j++) {
                /* phgQuadGradBasAGradBas(e, u, j, A, u, i): g1 = grad j */
                double d = 0., d0;
                for (int q = 0; q < nq; q++) {
                    const double *g1 = gx[q][j], *g2 = gx[q][i];
                    d0 = 0.;
                    d0 += g1[0] * g2[0] + g1[1] * g2[1] + g1[2] * g2[2];
                    if (pb->use_kappa)
                        d += d0 * qw[q] * kappa;
                    else
                        d += d0 * qw[q];
                }
                K[i * N + j] = K[j * N + i] = d * vol;
            }
        if (pb->mass_coef != 0.0) {
            /* simplest.c:70 a * phgQuadBasDotBas, order p+p (quad.c:5582) */
            nq = orc_quad3d(2 * p, &qp, &qw);
            for (int q = 0; q < nq; q++)
                orc_bas(p, qp + 4 * q, bv[q]);
            for (int i = 0; i < N; i++)
                for (int j = 0; j <= i; j++) {
                    double d = 0., d0;
                    for (int q = 0; q < nq; q++) {
                        d0 = 0.;
                        d0 += bv[q][j] * bv[q][i];
                        d += d0 * qw[q];
                    }
                    K[i * N + j] = K[j * N + i] =
                        K[i * N + j] + pb->mass_coef * (d * vol);
                }
        }
        /* load: quad.c:5481 phgQuadDofTimesBas with f_h at the points
         * evaluated by dof-utils.c:59 phgDofEval_ (sum in basis order) */
        nq = orc_quad3d(pb->rhs_order < 0 ? 2 * p : pb->rhs_order, &qp, &qw);
        {
            double fq[NQMAX];
            for (int q = 0; q < nq; q++) {
                orc_bas(p, qp + 4 * q, bv[q]);
                fq[q] = 0.0;
                for (int k = 0; k < N; k++)
                    fq[q] += bv[q][k] * f_loc[k];
            }
            for (int i = 0; i < N; i++) {
                double d = 0.0;
                if (pb->qcache) {
                    /* quad-cache.c:1016 phgQCIntegrate_(Q_BAS i, Q_f):
                     * both data non-constant, M = N = K = 1 */
                    for (int q = 0; q < nq; q++)
                        d += bv[q][i] * fq[q] * qw[q];
                }
                else {
                    for (int q = 0; q < nq; q++)
                        d += bv[q][i] * qw[q] * fq[q];
                }
                b[i] = d * vol;
            }
        }
        return;
    }

    /* linear elasticity, elastic.c:90 stiffness_matrix with
     * quad = phgQuadGetQuad3D((p-1)*2) (elastic.c:141) */
    {
        int M = 3 * N;
        double E = pb->E, nu = pb->nu;
        double a = nu / (1.0 - nu);
        double bb = (1. - 2. * nu) / (2. * (1. - nu));
        double c = E * (1. - nu) / ((1. + nu) * (1. - 2. * nu)) * vol;
        nq = orc_quad3d((p - 1) * 2, &qp, &qw);
        for (int q = 0; q < nq; q++) {
            orc_grad_lambda(p, qp + 4 * q, glam);
            for (int k = 0; k < N; k++)
                xgrad(p, k, glam + 4 * k, J, gx[q][k]);
        }
        for (int i = 0; i < N; i++)
            for (int j = 0; j <= i; j++) {
                double a00, a01, a02, a10, a11, a12, a20, a21, a22;
                a00 = a01 = a02 = a10 = a11 = a12 = a20 = a21 = a22 = 0.;
                for (int q = 0; q < nq; q++) {
                    const double *pm = gx[q][i], *pn = gx[q][j];
                    double w = qw[q];
                    a00 += (pm[0] * pn[0] + bb * (pm[1] * pn[1] + pm[2] * pn[2])) * w;
                    a01 += (a * pm[0] * pn[1] + bb * pm[1] * pn[0]) * w;
                    a02 += (a * pm[0] * pn[2] + bb * pm[2] * pn[0]) * w;
                    a10 += (a * pm[1] * pn[0] + bb * pm[0] * pn[1]) * w;
                    a11 += (pm[1] * pn[1] + bb * (pm[0] * pn[0] + pm[2] * pn[2])) * w;
                    a12 += (a * pm[1] * pn[2] + bb * pm[2] * pn[1]) * w;
                    a20 += (a * pm[2] * pn[0] + bb * pm[0] * pn[2]) * w;
                    a21 += (a * pm[2] * pn[1] + bb * pm[1] * pn[2]) * w;
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
        /* load: phgQuadDofTimesBas, vector path, then negated
         * (elastic.c:208-212) */
        nq = orc_quad3d(pb->rhs_order < 0 ? 2 * p : pb->rhs_order, &qp, &qw);
        {
            double fq[NQMAX][3];
            for (int q = 0; q < nq; q++) {
                orc_bas(p, qp + 4 * q, bv[q]);
                fq[q][0] = fq[q][1] = fq[q][2] = 0.0;
                for (int k = 0; k < N; k++)
                    for (int cc = 0; cc < 3; cc++)
                        fq[q][cc] += bv[q][k] * f_loc[3 * k + cc];
            }
            for (int i = 0; i < N; i++) {
                double r[3] = {0, 0, 0}, d;
                for (int q = 0; q < nq; q++) {
                    d = bv[q][i] * qw[q];
                    for (int cc = 0; cc < 3; cc++)
                        r[cc] += d * fq[q][cc];
                }
                for (int cc = 0; cc < 3; cc++) {
                    r[cc] *= vol;
                    b[3 * i + cc] = -r[cc];
                }
            }
        }
    }
}

/*------------------------- PHG row-list insertion -------------------------*/

typedef struct {
    int32_t ncols, alloc;
    int32_t *cols;
    double *data;
} mrow;

#define TRUNK 8

/* matvec.c:3045 add_columns_to_row, SORT_COLUMNS == 0: linear search,
 * add to an existing column or append */
static void row_add(mrow *r, int32_t col, double v)
{
    int32_t i;
    for (i = 0; i < r->ncols; i++)
        if (r->cols[i] == col)
            break;
    if (i < r->ncols) {
        r->data[i] += v;
        return;
    }
    if (r->ncols >= r->alloc) {
        r->alloc += TRUNK;
        r->cols = realloc(r->cols, (size_t)r->alloc * sizeof(int32_t));
        r->data = realloc(r->data, (size_t)r->alloc * sizeof(double));
    }
    r->cols[r->ncols] = col;
    r->data[r->ncols++] = v;
}

orc_csr *orc_assemble(const orc_problem *pb, int64_t nelem, const double *verts,
                      const int32_t *elem_verts, const int32_t *map,
                      int64_t ndof, const uint8_t *bdry, const double *g,
                      const double *f, const double *kappa)
{
    int p = pb->p, N = orc_nbas(p), dim = pb->dim;
    int M = N * dim;
    int64_t n = ndof * dim;
    mrow *rows = calloc((size_t)n, sizeof(mrow));
    double *rhs = calloc((size_t)n, sizeof(double));
    double *K = xmalloc((size_t)M * M * sizeof(double));
    double *B = xmalloc((size_t)M * sizeof(double));
    double *floc = xmalloc((size_t)M * sizeof(double));
    int32_t *I = xmalloc((size_t)M * sizeof(int32_t));

    for (int64_t e = 0; e < nelem; e++) {
        double x[4][3];
        const int32_t *m = map + (int64_t)N * e;
        for (int a = 0; a < 4; a++)
            for (int c = 0; c < 3; c++)
                x[a][c] = verts[3 * (int64_t)elem_verts[4 * e + a] + c];
        for (int i = 0; i < N; i++)
            for (int c = 0; c < dim; c++) {
                I[i * dim + c] = (int32_t)((int64_t)m[i] * dim + c);
                floc[i * dim + c] = f[(int64_t)m[i] * dim + c];
            }
        orc_element(pb, x, kappa ? kappa[e] : 0.0, floc, K, B);
        for (int i = 0; i < N; i++) {
            int is_b = bdry[m[i]] != 0;
            for (int c = 0; c < dim; c++) {
                int a = i * dim + c;
                mrow *r = rows + I[a];
                if (is_b) {
                    /* Dirichlet row e_i, rhs g (dof-utils.c:1369-1392);
                     * zeros dropped (matvec.c:5460) */
                    row_add(r, I[a], 1.0);
                    B[a] = g[(int64_t)m[i] * dim + c];
                }
                else {
                    for (int b2 = 0; b2 < M; b2++) {
                        double v = K[a * M + b2];
                        if (v == 0.0)
                            continue;
                        row_add(r, I[b2], v);
                    }
                }
            }
        }
        /* phgVecAddEntries (matvec.c:5661): zeros skipped */
        for (int a = 0; a < M; a++)
            if (B[a] != 0.0)
                rhs[I[a]] += B[a];
    }

    orc_csr *A = xmalloc(sizeof(*A));
    A->n = n;
    A->rowptr = xmalloc((size_t)(n + 1) * sizeof(int64_t));
    A->rowptr[0] = 0;
    for (int64_t i = 0; i < n; i++)
        A->rowptr[i + 1] = A->rowptr[i] + rows[i].ncols;
    A->nnz = A->rowptr[n];
    A->cols = xmalloc((size_t)A->nnz * sizeof(int32_t));
    A->vals = xmalloc((size_t)A->nnz * sizeof(double));
    for (int64_t i = 0; i < n; i++) {
        memcpy(A->cols + A->rowptr[i], rows[i].cols, (size_t)rows[i].ncols * sizeof(int32_t));
        memcpy(A->vals + A->rowptr[i], rows[i].data, (size_t)rows[i].ncols * sizeof(double));
        free(rows[i].cols);
        free(rows[i].data);
    }
    A->rhs = rhs;
    free(rows);
    free(K);
    free(B);
    free(floc);
    free(I);
    return A;
}

void orc_eliminate(orc_csr *A, const uint8_t *bdry)
{
    int64_t n = A->n, w = 0;
    int64_t *bstart = xmalloc((size_t)(n + 1) * sizeof(int64_t));
    int32_t *bcols = xmalloc((size_t)A->nnz * sizeof(int32_t));
    double *bvals = xmalloc((size_t)A->nnz * sizeof(double));
    int64_t nb = 0;

    /* save_bdry_columns(mat, TRUE): keep interior columns and the diagonal
     * in the row, move every boundary column to bdry_rows */
    for (int64_t i = 0; i < n; i++) {
        int64_t s = A->rowptr[i], t = A->rowptr[i + 1];
        bstart[i] = nb;
        A->rowptr[i] = w;
        for (int64_t j = s; j < t; j++) {
            int32_t c = A->cols[j];
            if (c == i || !bdry[c]) {
                A->cols[w] = c;
                A->vals[w++] = A->vals[j];
            }
            if (bdry[c]) {
                bcols[nb] = c;
                bvals[nb++] = A->vals[j];
            }
        }
    }
    A->rowptr[n] = w;
    bstart[n] = nb;
    A->nnz = w;

    /* update_rhs (solver.c:714-724): rhs_i -= coeff*rhs_k*a_ik / d_k */
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = bstart[i]; j < bstart[i + 1]; j++) {
            int32_t k = bcols[j];
            if (k == i)
                continue;
            double dk = bvals[bstart[k]];
            A->rhs[i] -= 1.0 * A->rhs[k] * bvals[j] / dk;
        }
    free(bstart);
    free(bcols);
    free(bvals);
}

void orc_matvec(const orc_csr *A, const double *x, double *y)
{
    for (int64_t i = 0; i < A->n; i++) {
        int64_t s = A->rowptr[i], t = A->rowptr[i + 1];
        double beta;
        y[i] = 0.0;
        if (t == s)
            continue;
        beta = A->vals[s] * x[A->cols[s]];
        for (int64_t j = s + 1; j < t; j++)
            beta += A->vals[j] * x[A->cols[j]];
        y[i] += beta;
    }
}

static double norm2(int64_t n, const double *v)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i++)
        s += v[i] * v[i];
    return sqrt(s);
}

static double dot(int64_t n, const double *x, const double *y)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i++)
        s += x[i] * y[i];
    return s;
}

int orc_pcg(const orc_csr *A, const double *b, double *x, int maxit,
            double rtol, double btol, double atol, double *residual)
{
    int64_t n = A->n;
    double *diag = xmalloc((size_t)n * sizeof(double));
    double *r = xmalloc((size_t)n * sizeof(double));
    double *z = xmalloc((size_t)n * sizeof(double));
    double *p = xmalloc((size_t)n * sizeof(double));
    double *Ap = xmalloc((size_t)n * sizeof(double));
    double rb, r0, res, tol, rz1 = 0.0, rz2, alpha, x1 = 0.0;
    int nits = 0;

    /* phgMatSetupDiagonal (matvec.c:500): first diagonal entry in the row */
    for (int64_t i = 0; i < n; i++) {
        diag[i] = 0.0;
        for (int64_t j = A->rowptr[i]; j < A->rowptr[i + 1]; j++)
            if (A->cols[j] == i) {
                diag[i] = A->vals[j];
                break;
            }
    }
    rb = norm2(n, b);
    /* IsZero (utils.h:451): |a| < 10 * FLOAT_MIN */
    if (fabs(rb) < 10.0 * DBL_MIN) {
        /* solver-phg.c:497-501: x = 0, *res left as the caller set it */
        memset(x, 0, (size_t)n * sizeof(double));
        free(diag);
        free(r);
        free(z);
        free(p);
        free(Ap);
        return 0;
    }
    for (int64_t i = 0; i < n; i++)
        x1 += fabs(x[i]);
    if (!(fabs(x1 / (double)n) < 10.0 * DBL_MIN)) {
        orc_matvec(A, x, Ap);
        for (int64_t i = 0; i < n; i++)
            r[i] = -Ap[i] + b[i];
    }
    else {
        memcpy(r, b, (size_t)n * sizeof(double));
    }
    r0 = res = norm2(n, r);
    tol = r0 * rtol;
    if (tol < rb * btol)
        tol = rb * btol;
    if (tol < atol)
        tol = atol;

    while (res > tol && nits < maxit) {
        for (int64_t i = 0; i < n; i++) {
            if (diag[i] == 0.0) {
                nits = -1;
                goto done;
            }
            z[i] = r[i] / diag[i];
        }
        nits++;
        if (nits == 1) {
            memcpy(p, z, (size_t)n * sizeof(double));
            rz1 = dot(n, r, z);
        }
        else {
            rz2 = rz1;
            rz1 = dot(n, r, z);
            if (rz2 == 0.0)
                break;
            double beta = rz1 / rz2;
            for (int64_t i = 0; i < n; i++)
                p[i] = beta * p[i];
            for (int64_t i = 0; i < n; i++)
                p[i] += z[i];
        }
        orc_matvec(A, p, Ap);
        if ((alpha = dot(n, p, Ap)) == 0.0)
            break;
        alpha = rz1 / alpha;
        for (int64_t i = 0; i < n; i++)
            x[i] += alpha * p[i];
        double ma = -1.0 * alpha;
        for (int64_t i = 0; i < n; i++)
            r[i] += ma * Ap[i];
        res = norm2(n, r);
    }
done:
    if (residual)
        *residual = res;
    free(diag);
    free(r);
    free(z);
    free(p);
    free(Ap);
    return nits;
}

/* matvec.c:1237 phgVecAXPBY: y = b y (skipped for b = 1), then y += a x
 * (skipped for a = 0); the special cases a, b = +-1 of the reference give
 * the same values */
static void axpby(int64_t n, double a, const double *x, double b, double *y)
{
    if (a == 0.0 && b == 0.0) {
        memset(y, 0, (size_t)n * sizeof(double));
        return;
    }
    if (b != 1.0)
        for (int64_t i = 0; i < n; i++)
            y[i] = b * y[i];
    if (a == 0.0)
        return;
    for (int64_t i = 0; i < n; i++)
        y[i] += a * x[i];
}

/* solver-phg.c:207 jacobi_pc; returns -1 on a zero diagonal */
static int jacobi(int64_t n, const double *diag, const double *r, double *z)
{
    for (int64_t i = 0; i < n; i++) {
        if (diag[i] == 0.0)
            return -1;
        z[i] = r[i] / diag[i];
    }
    return 0;
}

int orc_bicgstab(const orc_csr *A, const double *b, double *x, int maxit,
                 double rtol, double btol, double atol, double *residual)
{
    /* solver-phg.c:607-782 solver_BiCGstab with the Jacobi PC */
    int64_t n = A->n;
    size_t sz = (size_t)n * sizeof(double);
    double *diag = xmalloc(sz), *r1 = xmalloc(sz), *rg = xmalloc(sz);
    double *rh = xmalloc(sz), *pg = xmalloc(sz), *ph = xmalloc(sz);
    double *sg = xmalloc(sz), *sh = xmalloc(sz), *tg = xmalloc(sz);
    double *vg = xmalloc(sz);
    double rb, r0, res = 0.0, tol, x1 = 0.0;
    double rho0 = 0.0, rho1, alpha = 0.0, beta, omega = 0.0;
    int nits = 0;

    for (int64_t i = 0; i < n; i++) {
        diag[i] = 0.0;
        for (int64_t j = A->rowptr[i]; j < A->rowptr[i + 1]; j++)
            if (A->cols[j] == i) {
                diag[i] = A->vals[j];
                break;
            }
    }
    rb = norm2(n, b);
    if (fabs(rb) < 10.0 * DBL_MIN) {
        /* solver-phg.c:631: x = 0, the residual is left as the caller set it */
        memset(x, 0, sz);
        res = 1.0;
        goto done;
    }
    for (int64_t i = 0; i < n; i++)
        x1 += fabs(x[i]);
    if (!(fabs(x1 / (double)n) < 10.0 * DBL_MIN)) {
        orc_matvec(A, x, r1);
        memcpy(rg, r1, sz);
        axpby(n, 1.0, b, -1.0, rg);
    }
    else
        memcpy(rg, b, sz);
    r0 = res = norm2(n, rg);
    tol = r0 * rtol;
    if (tol < rb * btol)
        tol = rb * btol;
    if (tol < atol)
        tol = atol;
    memcpy(rh, rg, sz);
    memset(sh, 0, sz);
    memset(ph, 0, sz);
    (void)r0;

    while (res > tol && nits < maxit) {
        nits++;
        rho1 = dot(n, rg, rh);
        if (rho1 == 0.0)
            break;                          /* "method failed" */
        if (nits == 1)
            memcpy(pg, rg, sz);
        else {
            beta = (rho1 * alpha) / (rho0 * omega);
            memcpy(sg, pg, sz);
            axpby(n, -omega, vg, 1.0, sg);
            memcpy(pg, rg, sz);
            axpby(n, beta, sg, 1.0, pg);
        }
        rho0 = rho1;
        if (jacobi(n, diag, pg, ph)) {
            nits = -1;
            goto done;
        }
        orc_matvec(A, ph, vg);
        alpha = rho1 / dot(n, rh, vg);
        memcpy(sg, rg, sz);
        axpby(n, -alpha, vg, 1.0, sg);
        omega = norm2(n, sg);
        if (omega <= 1e-60) {
            axpby(n, alpha, ph, 1.0, x);
            orc_matvec(A, x, r1);
            memcpy(rg, b, sz);
            axpby(n, -1.0, r1, 1.0, rg);
            res = norm2(n, rg);
            break;
        }
        if (jacobi(n, diag, sg, sh)) {
            nits = -1;
            goto done;
        }
        orc_matvec(A, sh, tg);
        omega = dot(n, tg, sg) / dot(n, tg, tg);
        axpby(n, omega, sh, alpha, ph);
        axpby(n, 1.0, ph, 1.0, x);
        memcpy(rg, sg, sz);
        axpby(n, -omega, tg, 1.0, rg);
        res = norm2(n, rg);
        if (res <= tol)
            break;
    }
done:
    if (residual)
        *residual = res;
    free(diag); free(r1); free(rg); free(rh); free(pg); free(ph);
    free(sg); free(sh); free(tg); free(vg);
    return nits;
}

/* matvec.c:552 phgMatVec with A == NULL (the identity): y := alpha x +
 * beta y with the reference's case split (matvec.c:606-737) */
static void vaxpby(int64_t n, double alpha, const double *x, double beta,
                   double *y)
{
    if (alpha == 0.0) {
        if (beta == 0.0)
            memset(y, 0, (size_t)n * sizeof(double));
        else if (beta == -1.0)
            for (int64_t i = 0; i < n; i++)
                y[i] = -y[i];
        else if (beta != 1.0)
            for (int64_t i = 0; i < n; i++)
                y[i] *= beta;
        return;
    }
    for (int64_t i = 0; i < n; i++) {
        double ax = alpha == 1.0 ? x[i] : (alpha == -1.0 ? -x[i] : alpha * x[i]);
        if (beta == 0.0)
            y[i] = ax;
        else if (beta == -1.0)
            y[i] = ax - y[i];
        else if (beta == 1.0)
            y[i] += ax;
        else
            y[i] = ax + beta * y[i];
    }
}

/* phgMatVec with a matrix: y := beta y, then y_i += row sum of
 * a_ij (alpha x_j) in storage order (matvec.c:606-633, :842-930) */
static void amatvec(const orc_csr *A, double alpha, const double *x,
                    double beta, double *y)
{
    int64_t n = A->n;
    double *data = xmalloc((size_t)n * sizeof(double));
    vaxpby(n, 0.0, NULL, beta, y);
    for (int64_t i = 0; i < n; i++)
        data[i] = alpha == 1.0 ? x[i] : (alpha == -1.0 ? -x[i] : alpha * x[i]);
    for (int64_t i = 0; i < n; i++) {
        int64_t j0 = A->rowptr[i], j1 = A->rowptr[i + 1];
        if (j1 == j0)
            continue;
        double s = A->vals[j0] * data[A->cols[j0]];
        for (int64_t j = j0 + 1; j < j1; j++)
            s += A->vals[j] * data[A->cols[j]];
        y[i] += s;
    }
    free(data);
}

/* The GMRES family of solver-phg.c:59-77 (gmres_names / gmres_funcs) with
 * jacobi_pc (solver-phg.c:207), one restatement for the six variants:
 *   type 0 gmres_l   solver-phg.c:783   left preconditioned
 *   type 1 gmres_r   solver-phg.c:1051  right preconditioned
 *   type 2 fgmres_r  solver-phg.c:1263  right, flexible (keeps M^-1 v_i)
 *   type 3 lgmres_l  solver-phg.c:1481  left, augmented
 *   type 4 lgmres_r  solver-phg.c:1801  right, augmented
 *   type 5 flgmres_r solver-phg.c:2063  right, flexible, augmented
 * The augmented ("L") variants add the corrections of the last augk
 * cycles to the Krylov space (cycle c has restart + min(c, augk) steps).
 * Left variants run on z = M^-1 (b - A x); when the Givens estimate |z|
 * reaches the tolerance they check the true residual and, if it is still
 * above it, continue the cycle until the estimate has dropped by the ratio
 * of the two (false convergence, solver-phg.c:904-908, :979-1003); they
 * report the last true residual.  Right variants report the Givens
 * estimate once it reaches the tolerance.  Either reports |r0| when no
 * tolerance check was made. */
enum { GM_L = 0, GM_R, GM_FR, GM_LL, GM_LR, GM_FLR };

int orc_gmres(int type, const orc_csr *A, const double *b, double *x, int restart,
              int augk, int maxit, double rtol, double btol, double atol,
              double *residual)
{
    int64_t n = A->n;
    size_t sz = (size_t)n * sizeof(double);
    const int left = type == GM_L || type == GM_LL;
    const int flex = type == GM_FR || type == GM_FLR;
    const int aug = type >= GM_LL;
    int nits = 0, i, j, k, itr_out = 0, mk, m, rs;
    int break_flag = 0, false_conv = 0;
    double res, r_norm, b_norm, tol, tmp, h1, h2, x1 = 0.0, s0 = 1.0;
    if (type < GM_L || type > GM_FLR)
        return -2;
    if (restart <= 0)
        restart = maxit + 1;
    if (!aug)
        augk = 0;                   /* Unused(augk) in the plain variants */
    else if (augk <= 0)
        augk = 2;
    b_norm = norm2(n, b);
    if (fabs(b_norm) < 10.0 * DBL_MIN) {
        memset(x, 0, sz);
        if (residual)
            *residual = 1.0;        /* GMRES_Solve's initial value */
        return 0;
    }
    double *diag = xmalloc(sz), *z = xmalloc(sz), *t = xmalloc(sz);
    for (int64_t q = 0; q < n; q++) {
        diag[q] = 0.0;
        for (int64_t jj = A->rowptr[q]; jj < A->rowptr[q + 1]; jj++)
            if (A->cols[jj] == q) {
                diag[q] = A->vals[jj];
                break;
            }
    }
    mk = restart;
    m = restart + augk;             /* largest subspace */
    double **v = xmalloc((size_t)(m + 2) * sizeof(double *));
    double **flv = xmalloc((size_t)(m + 2) * sizeof(double *));
    double **zp = xmalloc((size_t)(augk + 1) * sizeof(double *));
    for (i = 0; i < m + 2; i++) {
        v[i] = xmalloc(sz);
        flv[i] = flex ? xmalloc(sz) : NULL;
    }
    for (i = 0; i < augk; i++)
        zp[i] = xmalloc(sz);
    double *H = xmalloc((size_t)(m + 2) * (m + 2) * sizeof(double));
#define HH(a, c) H[(size_t)(a) * (m + 2) + (c)]
    /* y persists over cycles and starts at zero: the augmented update
     * reads y[mk+1..mk+ak] even when the cycle stopped before step mk+ak
     * (solver-phg.c:1679-1683), i.e. values of an earlier cycle */
    double *y = xmalloc((size_t)(m + 2) * 4 * sizeof(double));
    memset(y, 0, (size_t)(m + 2) * 4 * sizeof(double));
    double *sv = y + m + 2, *c1 = y + 2 * (m + 2), *c2 = y + 3 * (m + 2);

    /* r (left: the vector t, right: z) = b - A x, skipped for x = 0 */
    double *r0v = left ? t : z;
    memcpy(r0v, b, sz);
    for (int64_t q = 0; q < n; q++)
        x1 += fabs(x[q]);
    if (x1 != 0.0)
        amatvec(A, -1.0, x, 1.0, r0v);
    r_norm = res = norm2(n, r0v);
    tol = r_norm * rtol;
    if (tol < b_norm * btol)
        tol = b_norm * btol;
    if (tol < atol)
        tol = atol;
    if (left && jacobi(n, diag, t, z)) {
        nits = -1;
        goto done;
    }
    sv[0] = 1.0;
    while (nits < maxit) {
        rs = itr_out <= augk ? mk + itr_out : mk + augk;
        sv[1] = r_norm = norm2(n, z);
        vaxpby(n, 1.0 / r_norm, z, 0.0, v[1]);
        i = 0;
        memset(H, 0, (size_t)(m + 2) * (m + 2) * sizeof(double));
        while (i < rs && (r_norm > tol || false_conv) && nits < maxit) {
            i++, nits++;
            double *src = i <= mk ? v[i] : zp[i - mk - 1], *w = v[i + 1];
            if (left) {                         /* w = M^-1 A src */
                amatvec(A, 1.0, src, 0.0, t);
                if (jacobi(n, diag, t, w)) {
                    nits = -1;
                    goto done;
                }
            }
            else {                              /* w = A M^-1 src */
                if (jacobi(n, diag, src, t)) {
                    nits = -1;
                    goto done;
                }
                amatvec(A, 1.0, t, 0.0, w);
                if (flex)
                    memcpy(flv[i], t, sz);
            }
            for (k = 1; k <= i; k++) {          /* modified Gram-Schmidt */
                HH(i, k) = dot(n, w, v[k]);
                vaxpby(n, -HH(i, k), v[k], 1.0, w);
            }
            HH(i, i + 1) = norm2(n, w);
            if (fabs(HH(i, i + 1)) < 10.0 * DBL_MIN) {
                i--;                            /* happy breakdown */
                break_flag = 1;
                break;
            }
            vaxpby(n, 0.0, NULL, 1.0 / HH(i, i + 1), w);
            for (j = 1; j < i; j++) {           /* previous rotations */
                h1 = c1[j] * HH(i, j) + c2[j] * HH(i, j + 1);
                h2 = -c2[j] * HH(i, j) + c1[j] * HH(i, j + 1);
                HH(i, j) = h1;
                HH(i, j + 1) = h2;
            }
            tmp = sqrt(HH(i, i) * HH(i, i) + HH(i, i + 1) * HH(i, i + 1));
            if (fabs(tmp) < 10.0 * DBL_MIN)
                tmp = 1.e-16;
            c1[i] = HH(i, i) / tmp;
            c2[i] = HH(i, i + 1) / tmp;
            sv[i + 1] = -c2[i] * sv[i];
            sv[i] = c1[i] * sv[i];
            HH(i, i) = c1[i] * HH(i, i) + c2[i] * HH(i, i + 1);
            r_norm = fabs(sv[i + 1]);
            if (left) {
                if (false_conv && r_norm * s0 <= tol)
                    false_conv = 0;
            }
            else if (r_norm <= tol)
                break;
        }
        for (j = i; j > 0; j--) {               /* H y = s */
            y[j] = sv[j] / HH(j, j);
            for (k = j - 1; k > 0; k--)
                sv[k] -= HH(j, k) * y[j];
        }
        /* flexible: x += sum y_j flv_j (x += z after a cycle of no step) */
        if (flex) {
            if (i > 0) {
                vaxpby(n, y[1], flv[1], 0.0, z);
                for (j = 2; j <= i; j++)
                    vaxpby(n, y[j], flv[j], 1.0, z);
            }
            vaxpby(n, 1.0, z, 1.0, x);
        }
        /* z = V y plus the augmentation part: the correction (left) or its
         * preimage under M^-1 (right), kept as augmentation vector */
        if ((!flex || aug) && i > 0) {
            vaxpby(n, y[1], v[1], 0.0, z);
            if (i <= mk) {
                for (j = 2; j <= i; j++)
                    vaxpby(n, y[j], v[j], 1.0, z);
            }
            else {
                int ak = itr_out < augk ? itr_out : augk;
                for (j = 2; j <= mk; j++)
                    vaxpby(n, y[j], v[j], 1.0, z);
                for (j = 1; j <= ak; j++)
                    vaxpby(n, y[j + mk], zp[j - 1], 1.0, z);
            }
        }
        if (left)
            vaxpby(n, 1.0, z, 1.0, x);
        else if (!flex) {
            if (jacobi(n, diag, z, t)) {
                nits = -1;
                goto done;
            }
            vaxpby(n, 1.0, t, 1.0, x);
        }
        if (aug)
            memcpy(zp[itr_out % augk], z, sz);
        if (!left) {
            if (r_norm <= tol) {
                res = r_norm;
                break;
            }
            memcpy(z, b, sz);
            amatvec(A, -1.0, x, 1.0, z);
        }
        else {
            false_conv = 0;
            memcpy(t, b, sz);
            amatvec(A, -1.0, x, 1.0, t);
            if (r_norm <= tol) {
                res = norm2(n, t);
                if (res <= tol || break_flag)
                    break;
                if (i < rs) {
                    false_conv = 1;
                    s0 = res / r_norm;
                }
            }
            if (jacobi(n, diag, t, z)) {
                nits = -1;
                goto done;
            }
        }
        itr_out++;
    }
done:
#undef HH
    if (residual)
        *residual = res;
    for (i = 0; i < m + 2; i++) {
        free(v[i]);
        free(flv[i]);
    }
    for (i = 0; i < augk; i++)
        free(zp[i]);
    free(v); free(flv); free(zp); free(H); free(y);
    free(diag); free(z); free(t);
    return nits;
}

void orc_csr_free(orc_csr *A)
{
    if (!A)
        return;
    free(A->rowptr);
    free(A->cols);
    free(A->vals);
    free(A->rhs);
    free(A);
}

/*------------------------------------------------------------ adjoint
 * No reference counterpart (PHG has no adjoint).  Restated from the
 * definitions: J = 1/2 (u - ud)^T M (u - ud) with M the Pk mass matrix
 * (quad.c:5582 phgQuadBasDotBas, order 2p), the adjoint right-hand side
 * M (u - ud) accumulated element by element in index order, and
 * dJ/dlog(kappa_e) = -kappa_e lambda_e^T K1_e u_e with K1_e the element
 * stiffness for kappa = 1 (quad.c:5260). */
double orc_mass_apply(int p, int64_t nelem, const double *verts,
                      const int32_t *elem_verts, const int32_t *map,
                      int64_t ndof, const double *e, double *Me)
{
    int N = orc_nbas(p);
    orc_problem mp = {p, 1, 1, 1.0, -1, 0.0, 0.0, 0};
    double *K = xmalloc((size_t)N * N * sizeof(double));
    double *B = xmalloc((size_t)N * sizeof(double));
    double *fz = calloc((size_t)N, sizeof(double));
    double J = 0.0;
    for (int64_t i = 0; i < ndof; i++)
        Me[i] = 0.0;
    for (int64_t el = 0; el < nelem; el++) {
        double x[4][3];
        const int32_t *m = map + (int64_t)N * el;
        for (int a = 0; a < 4; a++)
            for (int c = 0; c < 3; c++)
                x[a][c] = verts[3 * (int64_t)elem_verts[4 * el + a] + c];
        /* kappa = 0 removes the stiffness term: K = M_e */
        orc_element(&mp, x, 0.0, fz, K, B);
        for (int i = 0; i < N; i++) {
            double s = 0.0;
            for (int j = 0; j < N; j++)
                s += K[i * N + j] * e[m[j]];
            Me[m[i]] += s;
            J += 0.5 * e[m[i]] * s;
        }
    }
    free(K);
    free(B);
    free(fz);
    return J;
}

void orc_adjoint_grad(int p, int64_t nelem, const double *verts,
                      const int32_t *elem_verts, const int32_t *map,
                      const double *kappa, const double *u,
                      const double *lam, double *grad)
{
    int N = orc_nbas(p);
    orc_problem sp = {p, 1, 1, 0.0, -1, 0.0, 0.0, 0};
    double *K = xmalloc((size_t)N * N * sizeof(double));
    double *B = xmalloc((size_t)N * sizeof(double));
    double *fz = calloc((size_t)N, sizeof(double));
    for (int64_t el = 0; el < nelem; el++) {
        double x[4][3];
        const int32_t *m = map + (int64_t)N * el;
        for (int a = 0; a < 4; a++)
            for (int c = 0; c < 3; c++)
                x[a][c] = verts[3 * (int64_t)elem_verts[4 * el + a] + c];
        orc_element(&sp, x, 1.0, fz, K, B);
        double s = 0.0;
        for (int i = 0; i < N; i++) {
            double t = 0.0;
            for (int j = 0; j < N; j++)
                t += K[i * N + j] * u[m[j]];
            s += lam[m[i]] * t;
        }
        grad[el] = -kappa[el] * s;
    }
    free(K);
    free(B);
    free(fz);
}
