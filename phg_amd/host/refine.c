/* Local bisection refinement of Kuhn meshes toward a point (host).
 *
 * PHG refines by bisection with conforming closure (refine.c:1230
 * phgRefineMarkedElements).  This restatement uses the bisection rule of
 * Maubach for simplices ordered along a lattice path: an element
 * (x0, x1, x2, x3) with tag d is cut at the midpoint z of edge (x0, xd) into
 *     (x0, ..., x(d-1), z, x(d+1), ..., x3)  and
 *     (x1, ..., xd, z, x(d+1), ..., x3),
 * both with tag d - 1 (3 after 1).  The Kuhn tetrahedra of a structured cube
 * mesh, ordered by increasing coordinate sum and tagged 3, form a
 * compatible initial mesh for this rule, so refining an element after
 * first refining every neighbour that shares its refinement edge with a
 * different refinement edge keeps the mesh conforming.  The element order
 * and vertex numbering differ from PHG's own refinement; the result is a
 * plain conforming mesh handed to phg_grid_number.
 */
#include "phg_grid.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int32_t v[4];
    int8_t tag;
    int8_t alive;
} relem;

/* open-addressing map (a, b) -> int32 value, a < b */
typedef struct {
    int64_t cap;
    int64_t *key;
    int32_t *val;
} pmap;

static void *xalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p)
        abort();
    return p;
}

static uint64_t mix(uint64_t k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

static void pmap_init(pmap *m, int64_t cap)
{
    int64_t c = 1024;
    while (c < 2 * cap)
        c <<= 1;
    m->cap = c;
    m->key = xalloc((size_t)c * sizeof(int64_t));
    m->val = xalloc((size_t)c * sizeof(int32_t));
    for (int64_t i = 0; i < c; i++)
        m->key[i] = -1;
}

static void pmap_free(pmap *m)
{
    free(m->key);
    free(m->val);
}

static int64_t pkey(int32_t a, int32_t b)
{
    if (a > b) {
        int32_t t = a;
        a = b;
        b = t;
    }
    return ((int64_t)a << 32) | (uint32_t)b;
}

/* slot of key k (existing or empty) */
static int64_t pmap_slot(pmap *m, int64_t k)
{
    int64_t i = (int64_t)(mix((uint64_t)k) & (uint64_t)(m->cap - 1));
    while (m->key[i] != -1 && m->key[i] != k)
        i = (i + 1) & (m->cap - 1);
    return i;
}

static void pmap_grow(pmap *m)
{
    pmap n;
    pmap_init(&n, m->cap);
    for (int64_t i = 0; i < m->cap; i++)
        if (m->key[i] != -1) {
            int64_t j = pmap_slot(&n, m->key[i]);
            n.key[j] = m->key[i];
            n.val[j] = m->val[i];
        }
    pmap_free(m);
    *m = n;
}

/* refinement state */
typedef struct {
    double *xyz;
    int64_t nv, cv;
    relem *el;
    int64_t ne, ce;
    pmap mid;               /* edge -> midpoint vertex */
    pmap head;              /* edge -> first node of its element list */
    int64_t nmid, nhead;    /* entries in the two maps */
    int32_t *lnext, *lelem; /* element lists (linked) */
    int64_t nl, cl;
} rstate;

static int32_t new_vertex(rstate *s, const double *x)
{
    if (s->nv == s->cv) {
        s->cv *= 2;
        s->xyz = realloc(s->xyz, (size_t)s->cv * 3 * sizeof(double));
        if (!s->xyz)
            abort();
    }
    memcpy(s->xyz + 3 * s->nv, x, 3 * sizeof(double));
    return (int32_t)s->nv++;
}

static void list_add(rstate *s, int32_t a, int32_t b, int32_t e)
{
    if (4 * s->nhead > 3 * s->head.cap)
        pmap_grow(&s->head);
    int64_t k = pkey(a, b), i = pmap_slot(&s->head, k);
    if (s->nl == s->cl) {
        s->cl *= 2;
        s->lnext = realloc(s->lnext, (size_t)s->cl * sizeof(int32_t));
        s->lelem = realloc(s->lelem, (size_t)s->cl * sizeof(int32_t));
        if (!s->lnext || !s->lelem)
            abort();
    }
    s->lelem[s->nl] = e;
    if (s->head.key[i] == -1) {
        s->head.key[i] = k;
        s->lnext[s->nl] = -1;
        s->nhead++;
    }
    else
        s->lnext[s->nl] = s->head.val[i];
    s->head.val[i] = (int32_t)s->nl++;
}

static const int tet_edges[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

static int32_t new_elem(rstate *s, const int32_t v[4], int tag)
{
    if (s->ne == s->ce) {
        s->ce *= 2;
        s->el = realloc(s->el, (size_t)s->ce * sizeof(relem));
        if (!s->el)
            abort();
    }
    int32_t e = (int32_t)s->ne++;
    memcpy(s->el[e].v, v, 4 * sizeof(int32_t));
    s->el[e].tag = (int8_t)tag;
    s->el[e].alive = 1;
    for (int k = 0; k < 6; k++)
        list_add(s, v[tet_edges[k][0]], v[tet_edges[k][1]], e);
    return e;
}

static int32_t midpoint(rstate *s, int32_t a, int32_t b)
{
    if (4 * s->nmid > 3 * s->mid.cap)
        pmap_grow(&s->mid);
    int64_t k = pkey(a, b), i = pmap_slot(&s->mid, k);
    if (s->mid.key[i] == k)
        return s->mid.val[i];
    double x[3];
    for (int c = 0; c < 3; c++)
        x[c] = 0.5 * (s->xyz[3 * (int64_t)a + c] + s->xyz[3 * (int64_t)b + c]);
    int32_t z = new_vertex(s, x);
    i = pmap_slot(&s->mid, k);
    s->mid.key[i] = k;
    s->mid.val[i] = z;
    s->nmid++;
    return z;
}

static void bisect(rstate *s, int32_t e)
{
    relem t = s->el[e];
    int d = t.tag;
    int32_t z = midpoint(s, t.v[0], t.v[d]);
    int32_t c1[4], c2[4];
    int k = 0;
    for (int i = 0; i < 4; i++)
        c1[i] = i == d ? z : t.v[i];
    /* (x1, ..., xd, z, x(d+1), ..., x3) */
    for (int i = 1; i <= d; i++)
        c2[k++] = t.v[i];
    c2[k++] = z;
    for (int i = d + 1; i < 4; i++)
        c2[k++] = t.v[i];
    int nt = d > 1 ? d - 1 : 3;
    s->el[e].alive = 0;
    new_elem(s, c1, nt);
    new_elem(s, c2, nt);
}

static int has_edge(const relem *t, int32_t a, int32_t b)
{
    int fa = 0, fb = 0;
    for (int i = 0; i < 4; i++) {
        fa |= t->v[i] == a;
        fb |= t->v[i] == b;
    }
    return fa && fb;
}

/* refines element e and, first, every element sharing its refinement edge
 * whose own refinement edge differs (recursively) */
static void refine(rstate *s, int32_t e, int depth)
{
    if (!s->el[e].alive || depth > 64)
        return;
    int32_t a = s->el[e].v[0], b = s->el[e].v[s->el[e].tag];
    int64_t k = pkey(a, b);
    for (;;) {
        int changed = 0;
        int64_t i = pmap_slot(&s->head, k);
        for (int32_t n = s->head.key[i] == k ? s->head.val[i] : -1; n != -1;
             n = s->lnext[n]) {
            int32_t f = s->lelem[n];
            if (!s->el[f].alive || f == e)
                continue;
            int32_t fa = s->el[f].v[0], fb = s->el[f].v[s->el[f].tag];
            if (pkey(fa, fb) != k) {
                refine(s, f, depth + 1);
                changed = 1;
                break;    /* the list grew: scan again */
            }
        }
        if (!changed)
            break;
    }
    /* bisect every live element around the edge (all share it as the
     * refinement edge now) */
    int64_t i = pmap_slot(&s->head, k);
    int32_t around[512];
    int na = 0;
    for (int32_t n = s->head.key[i] == k ? s->head.val[i] : -1; n != -1;
         n = s->lnext[n]) {
        int32_t f = s->lelem[n];
        if (s->el[f].alive && has_edge(&s->el[f], a, b) && na < 512)
            around[na++] = f;
    }
    for (int j = 0; j < na; j++)
        if (s->el[around[j]].alive)
            bisect(s, around[j]);
}

static int contains(const rstate *s, const relem *t, const double *p)
{
    /* barycentric test with a small tolerance */
    const double *x0 = s->xyz + 3 * (int64_t)t->v[0];
    double m[3][3], r[3];
    for (int i = 0; i < 3; i++) {
        const double *xi = s->xyz + 3 * (int64_t)t->v[i + 1];
        for (int c = 0; c < 3; c++)
            m[c][i] = xi[c] - x0[c];
        r[i] = p[i] - x0[i];
    }
    double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                 m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                 m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    if (det == 0.0)
        return 0;
    double l[3];
    for (int i = 0; i < 3; i++) {
        double mm[3][3];
        memcpy(mm, m, sizeof(mm));
        for (int c = 0; c < 3; c++)
            mm[c][i] = r[c];
        l[i] = (mm[0][0] * (mm[1][1] * mm[2][2] - mm[1][2] * mm[2][1]) -
                mm[0][1] * (mm[1][0] * mm[2][2] - mm[1][2] * mm[2][0]) +
                mm[0][2] * (mm[1][0] * mm[2][1] - mm[1][1] * mm[2][0])) / det;
    }
    const double eps = -1e-12;
    return l[0] >= eps && l[1] >= eps && l[2] >= eps &&
           1.0 - l[0] - l[1] - l[2] >= eps;
}

phg_grid *phg_grid_refine_toward(const phg_grid *g, const double pt[3],
                                 int levels, double radius, double shrink,
                                 const int bcode[6])
{
    rstate s;
    memset(&s, 0, sizeof(s));
    s.cv = g->nvert * 4 + 1024;
    s.xyz = xalloc((size_t)s.cv * 3 * sizeof(double));
    memcpy(s.xyz, g->verts, (size_t)g->nvert * 3 * sizeof(double));
    s.nv = g->nvert;
    s.ce = (int64_t)g->nelem * 4 + 1024;
    s.el = xalloc((size_t)s.ce * sizeof(relem));
    s.cl = s.ce * 6;
    s.lnext = xalloc((size_t)s.cl * sizeof(int32_t));
    s.lelem = xalloc((size_t)s.cl * sizeof(int32_t));
    pmap_init(&s.mid, (int64_t)g->nelem * 2);
    pmap_init(&s.head, (int64_t)g->nelem * 4);
    for (int64_t e = 0; e < g->nelem; e++) {
        /* order the vertices along the lattice path (increasing coordinate
         * sum); tag 3 */
        int32_t v[4];
        double key[4];
        memcpy(v, g->elem_verts + 4 * e, sizeof(v));
        for (int i = 0; i < 4; i++) {
            const double *x = g->verts + 3 * (int64_t)v[i];
            key[i] = x[0] + x[1] + x[2];
        }
        for (int i = 1; i < 4; i++)
            for (int j = i; j > 0 && key[j - 1] > key[j]; j--) {
                double tk = key[j];
                key[j] = key[j - 1];
                key[j - 1] = tk;
                int32_t tv = v[j];
                v[j] = v[j - 1];
                v[j - 1] = tv;
            }
        new_elem(&s, v, 3);
    }
    double r = radius;
    for (int lev = 0; lev < levels; lev++) {
        int64_t ne0 = s.ne;
        int32_t *marked = xalloc((size_t)ne0 * sizeof(int32_t));
        int64_t nm = 0;
        for (int64_t e = 0; e < ne0; e++) {
            if (!s.el[e].alive)
                continue;
            double c[3] = {0, 0, 0};
            for (int i = 0; i < 4; i++)
                for (int d = 0; d < 3; d++)
                    c[d] += 0.25 * s.xyz[3 * (int64_t)s.el[e].v[i] + d];
            double dist = sqrt((c[0] - pt[0]) * (c[0] - pt[0]) +
                               (c[1] - pt[1]) * (c[1] - pt[1]) +
                               (c[2] - pt[2]) * (c[2] - pt[2]));
            if (dist < r || contains(&s, &s.el[e], pt))
                marked[nm++] = (int32_t)e;
        }
        for (int64_t i = 0; i < nm; i++)
            refine(&s, marked[i], 0);
        free(marked);
        r *= shrink;
    }
    /* the live elements form the new grid */
    int64_t nlive = 0;
    for (int64_t e = 0; e < s.ne; e++)
        nlive += s.el[e].alive;
    phg_grid *o = calloc(1, sizeof(phg_grid));
    if (!o)
        abort();
    o->nvert = (int32_t)s.nv;
    o->nelem = (int32_t)nlive;
    o->verts = s.xyz;
    s.xyz = NULL;
    o->elem_verts = xalloc((size_t)nlive * 4 * sizeof(int32_t));
    o->elem_bound = xalloc((size_t)nlive * 4);
    /* its ALBERT file carries types (all 0), so PHG imports it unpermuted */
    o->elem_type = calloc((size_t)nlive + 1, 1);
    if (!o->elem_type)
        abort();
    int64_t k = 0;
    static const int fv[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    for (int64_t e = 0; e < s.ne; e++) {
        if (!s.el[e].alive)
            continue;
        memcpy(o->elem_verts + 4 * k, s.el[e].v, 4 * sizeof(int32_t));
        for (int f = 0; f < 4; f++) {
            uint8_t b = PHG_BT_INTERIOR;
            for (int d = 0; d < 3; d++) {
                double x0 = o->verts[3 * (int64_t)s.el[e].v[fv[f][0]] + d];
                double x1 = o->verts[3 * (int64_t)s.el[e].v[fv[f][1]] + d];
                double x2 = o->verts[3 * (int64_t)s.el[e].v[fv[f][2]] + d];
                if (x0 == x1 && x0 == x2 && (x0 == 0.0 || x0 == 1.0)) {
                    int code = bcode[2 * d + (x0 == 1.0)];
                    b = code > 0 ? PHG_BT_DIRICHLET
                                 : (code < 0 ? PHG_BT_NEUMANN : PHG_BT_INTERIOR);
                }
            }
            o->elem_bound[4 * k + f] = b;
        }
        k++;
    }
    free(s.el);
    free(s.lnext);
    free(s.lelem);
    pmap_free(&s.mid);
    pmap_free(&s.head);
    phg_grid_number(o);
    return o;
}
