/* PHG's bisection refinement on the host (refine.c, serial run).
 *
 * The grid becomes a forest whose roots are its elements, in its traversal
 * order, and refine.c's algorithm runs on it:
 *   - refine_marked_traverse (refine.c:2052): depth first over the trees,
 *     children[0] before children[1]; a leaf with mark > 0 is refined by
 *     refine_path and the traversal continues into its children;
 *   - refine_path (refine.c:1661): a last-in first-out stack of elements to
 *     bisect; after each bisection update_neighbours pushes the children that
 *     have hanging faces, and the two neighbours sharing the refinement edge
 *     are pushed or matched against their own children;
 *   - divide_simplex (refine.c:749): children's vertex order, boundary types
 *     and bisection types for the types D, F, E, O and M; both children get
 *     the next element index;
 *   - new_vertex / add_vertex (refine.c:1230, :1173): the midpoint of the
 *     refinement edge, created once per edge in creation order;
 *   - update_neighbours (refine.c:1003): the neighbour links of the children.
 * Afterwards update_indices (refine.c:2176) lets the child that holds the
 * refinement edge's smaller vertex keep its parent's element index and the
 * halves of the parent's edges and faces (phgMapP2C, refine.c:518), and new
 * edges and faces are numbered in tree traversal order (utils.c:4282
 * update_edges_callback, utils.c:4469 update_faces_callback1).  In a serial
 * run phgUpdateElementIndices (utils.c:4631) changes nothing, so the element
 * index of a new element is the element count when it was created.
 *
 * Neighbours are node numbers (-1 for none, PHG's NULL); a face is interior
 * when its boundary type is PHG_BT_INTERIOR (HasLocalNeighbour). */
#include "phg_grid.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* bisection types, ALBERT codes (albert.c:36) */
enum { T_D = 0, T_F = 1, T_E = 2, T_M = 3, T_O = 4 };

typedef struct {
    int32_t verts[4], edges[6], faces[4];
    int32_t nb[4];              /* neighbour node or -1 */
    uint8_t bound[4];           /* boundary type of face i */
    int32_t child[2], parent;
    int32_t index, mark;
    uint8_t type;
} node_t;

/* (a, b) -> value, open addressing */
typedef struct {
    int64_t *key;
    int32_t *val;
    int64_t cap, n;
} pmap;

static void *xalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "phg_grid_bisect: out of memory\n");
        abort();
    }
    return p;
}

static void pm_init(pmap *m, int64_t n)
{
    m->cap = 64;
    while (m->cap < 2 * n)
        m->cap *= 2;
    m->key = xalloc((size_t)m->cap * sizeof(int64_t));
    m->val = xalloc((size_t)m->cap * sizeof(int32_t));
    for (int64_t i = 0; i < m->cap; i++)
        m->key[i] = -1;
    m->n = 0;
}

static void pm_free(pmap *m)
{
    free(m->key);
    free(m->val);
}

static int32_t *pm_slot(pmap *m, int32_t a, int32_t b, int *found);

static void pm_grow(pmap *m)
{
    pmap o = *m;
    pm_init(m, o.cap);
    for (int64_t i = 0; i < o.cap; i++)
        if (o.key[i] >= 0) {
            int f;
            *pm_slot(m, (int32_t)(o.key[i] >> 32), (int32_t)(o.key[i] & 0xffffffff), &f) =
                o.val[i];
        }
    pm_free(&o);
}

static int32_t *pm_slot(pmap *m, int32_t a, int32_t b, int *found)
{
    if (2 * (m->n + 1) > m->cap)
        pm_grow(m);
    const int64_t k = (int64_t)a << 32 | (uint32_t)b;
    uint64_t h = (uint64_t)k * 0x9E3779B97F4A7C15ull;
    int64_t i = (int64_t)(h >> 20) & (m->cap - 1);
    while (m->key[i] >= 0 && m->key[i] != k)
        i = (i + 1) & (m->cap - 1);
    *found = m->key[i] == k;
    if (!*found) {
        m->key[i] = k;
        m->n++;
    }
    return &m->val[i];
}

typedef struct {
    node_t *n;
    int64_t nn, cap;
    double *xyz;
    int64_t nv, vcap;
    int32_t nelem;              /* element count: the next new index */
    pmap mid;                   /* refinement edge -> new vertex */
    int32_t *stack;
    int64_t sp, scap;
    int failed;
} bis_t;

#define N(i) (B->n[i])

static int local(const bis_t *B, int32_t e, int v)
{
    return N(e).bound[v] == PHG_BT_INTERIOR;
}

/* GetNeighbour (utils.h:91) */
static int32_t nbr(const bis_t *B, int32_t e, int v)
{
    return local(B, e, v) ? N(e).nb[v] : -1;
}

static int32_t new_node(bis_t *B)
{
    if (B->nn >= B->cap) {
        B->cap = B->cap * 2 + 1024;
        B->n = realloc(B->n, (size_t)B->cap * sizeof(node_t));
        if (!B->n)
            abort();
    }
    node_t *c = &B->n[B->nn];
    for (int i = 0; i < 4; i++) {
        c->verts[i] = -1;
        c->faces[i] = -1;
        c->nb[i] = -1;
        c->bound[i] = PHG_BT_INTERIOR;
    }
    for (int i = 0; i < 6; i++)
        c->edges[i] = -1;
    c->child[0] = c->child[1] = c->parent = -1;
    c->index = -1;
    c->mark = 0;
    c->type = T_D;
    return (int32_t)B->nn++;
}

static void push(bis_t *B, int32_t e)
{
    if (B->sp >= B->scap) {
        B->scap = B->scap * 2 + 64;
        B->stack = realloc(B->stack, (size_t)B->scap * sizeof(int32_t));
        if (!B->stack)
            abort();
    }
    B->stack[B->sp++] = e;
}

/* refine.c:749 divide_simplex */
static void divide(bis_t *B, int32_t e)
{
    const int32_t c0 = new_node(B), c1 = new_node(B);
    node_t *E = &N(e), *a = &N(c0), *b = &N(c1);
    const int32_t *v = E->verts;
    const uint8_t *bt = E->bound, I = PHG_BT_INTERIOR;
    a->index = b->index = B->nelem++;
    E->child[0] = c0;
    E->child[1] = c1;
    a->parent = b->parent = e;
    a->mark = b->mark = E->mark - 1;
#define SET(c, t, v0, v1, v2, b0, b1, b2, b3) do { \
        (c)->type = (t); \
        (c)->verts[0] = (v0); (c)->verts[1] = (v1); (c)->verts[2] = (v2); (c)->verts[3] = -1; \
        (c)->bound[0] = (b0); (c)->bound[1] = (b1); (c)->bound[2] = (b2); (c)->bound[3] = (b3); \
    } while (0)
    switch (E->type) {
    case T_D:
        SET(a, T_F, v[0], v[2], v[3], I, bt[2], bt[3], bt[1]);
        SET(b, T_F, v[1], v[3], v[2], I, bt[3], bt[2], bt[0]);
        break;
    case T_F:
        SET(a, T_E, v[0], v[2], v[3], I, bt[2], bt[3], bt[1]);
        SET(b, T_E, v[1], v[2], v[3], I, bt[2], bt[3], bt[0]);
        break;
    case T_E:
        SET(a, T_D, v[0], v[2], v[3], I, bt[2], bt[3], bt[1]);
        SET(b, T_D, v[1], v[2], v[3], I, bt[2], bt[3], bt[0]);
        break;
    case T_O:
        SET(a, T_F, v[3], v[2], v[0], bt[3], bt[2], I, bt[1]);
        SET(b, T_F, v[3], v[2], v[1], bt[3], bt[2], I, bt[0]);
        break;
    default:    /* T_M */
        SET(a, T_F, v[0], v[2], v[3], I, bt[2], bt[3], bt[1]);
        SET(b, T_F, v[3], v[2], v[1], bt[3], bt[2], I, bt[0]);
        break;
    }
#undef SET
}

/* refine.c:1230 new_vertex with refine.c:1173 add_vertex */
static void new_vertex(bis_t *B, int32_t e)
{
    int32_t a = N(e).verts[0], b = N(e).verts[1], m;
    const int32_t lo = a < b ? a : b, hi = a < b ? b : a;
    int found;
    int32_t *slot = pm_slot(&B->mid, lo, hi, &found);
    if (found) {
        m = *slot;
    }
    else {
        if (B->nv >= B->vcap) {
            B->vcap = B->vcap * 2 + 4096;
            B->xyz = realloc(B->xyz, (size_t)B->vcap * 3 * sizeof(double));
            if (!B->xyz)
                abort();
        }
        m = (int32_t)B->nv++;
        *slot = m;
        const double *v0 = B->xyz + 3 * (int64_t)a, *v1 = B->xyz + 3 * (int64_t)b;
        double *x = B->xyz + 3 * (int64_t)m;
        x[0] = (v0[0] + v1[0]) * .5;
        x[1] = (v0[1] + v1[1]) * .5;
        x[2] = (v0[2] + v1[2]) * .5;
    }
    N(N(e).child[0]).verts[3] = N(N(e).child[1]).verts[3] = m;
}

/* refine.c:131 opposite_vertex: the vertex of e_op not on face v of e */
static int opposite_vertex(const bis_t *B, int32_t e, int v, int32_t e_op)
{
    int32_t u[3];
    for (int i = 0, k = 0; i < 4; i++)
        if (i != v)
            u[k++] = N(e).verts[i];
    for (int i = 0; i < 4; i++) {
        const int32_t w = N(e_op).verts[i];
        if (w != u[0] && w != u[1] && w != u[2])
            return i;
    }
    return 4;
}

/* refine.c:1603 share_refinement_edge: e2 holds the refinement edge of e1 */
static int share_refinement_edge(const bis_t *B, int32_t e1, int32_t e2)
{
    int i;
    for (i = 0; i < 4; i++)
        if (N(e1).verts[0] == N(e2).verts[i])
            break;
    if (i >= 4)
        return 0;
    for (i = 0; i < 4; i++)
        if (N(e1).verts[1] == N(e2).verts[i])
            break;
    return i < 4;
}

/* refine.c:1624 match_neighbour */
static int match_neighbour(bis_t *B, int32_t e0, int32_t e1)
{
    int i, j, k, m0[4] = {-1, -1, -1, -1}, m1[4] = {-1, -1, -1, -1};
    for (k = i = 0; i < 4; i++)
        for (j = 0; j < 4; j++) {
            if (m1[j] >= 0)
                continue;
            if (N(e0).verts[i] == N(e1).verts[j]) {
                k++;
                m0[i] = j;
                m1[j] = i;
                break;
            }
        }
    if (k == 3) {
        i = j = 0;
        for (k = 0; k < 4; k++) {
            if (m0[k] == -1)
                i = k;
            if (m1[k] == -1)
                j = k;
        }
        N(e0).nb[i] = e1;
        N(e1).nb[j] = e0;
        return 1;
    }
    return 0;
}

/* the face of e0 whose neighbour is e (GetNeighbour(e0, i) == e), -1 if none */
static int back_face(const bis_t *B, int32_t e0, int32_t e)
{
    for (int i = 0; i < 4; i++)
        if (nbr(B, e0, i) == e)
            return i;
    return -1;
}

/* refine.c:1003 update_neighbours; returns bit 0 / bit 1 when child 0 / 1
 * has a hanging face toward a neighbour not sharing the refinement edge */
static int update_neighbours(bis_t *B, int32_t e)
{
    const int32_t c0 = N(e).child[0], c1 = N(e).child[1];
    int ret = 0, i;
    int32_t e0;
    const int t = N(e).type;

    switch (t) {
    case T_O:
        N(c0).nb[2] = c1;
        N(c1).nb[2] = c0;
        break;
    case T_M:
        N(c0).nb[0] = c1;
        N(c1).nb[2] = c0;
        break;
    default:
        N(c0).nb[0] = c1;
        N(c1).nb[0] = c0;
    }
    N(c0).nb[3] = N(e).nb[1];
    N(c1).nb[3] = N(e).nb[0];

    /* the two neighbours not sharing the refinement edge */
    if ((e0 = nbr(B, e, 0)) >= 0) {
        if ((i = back_face(B, e0, e)) < 0) {
            B->failed = 1;
            return 0;
        }
        N(e0).nb[i] = c1;
        if (N(e0).child[0] >= 0 && opposite_vertex(B, e, 0, e0) >= 2)
            ret |= 2;
    }
    if ((e0 = nbr(B, e, 1)) >= 0) {
        if ((i = back_face(B, e0, e)) < 0) {
            B->failed = 1;
            return 0;
        }
        N(e0).nb[i] = c0;
        if (N(e0).child[0] >= 0 && opposite_vertex(B, e, 1, e0) >= 2)
            ret |= 1;
    }

    /* the two neighbours sharing the refinement edge */
    for (int f = 2; f < 4; f++) {
        int c0f, c1f;
        if (f == 2) {
            c0f = 1;
            c1f = t == T_D ? 2 : 1;
        }
        else {
            c0f = t == T_O ? 0 : 2;
            c1f = t == T_D ? 1 : ((t == T_O || t == T_M) ? 0 : 2);
        }
        if ((e0 = nbr(B, e, f)) < 0 || N(e0).child[0] < 0)
            continue;
        const int a0 = N(e0).verts[0] == N(e).verts[0] ? 0 : 1, a1 = 1 - a0;
        const int32_t ca[2] = {c0, c1};
        N(ca[0]).nb[c0f] = N(e0).child[a0];
        N(ca[1]).nb[c1f] = N(e0).child[a1];
        int b0, b1;
        const int t0 = N(e0).type;
        if (N(e0).nb[2] == e) {
            b0 = 1;
            b1 = t0 == T_D ? 2 : 1;
        }
        else {
            b0 = t0 == T_O ? 0 : 2;
            b1 = t0 == T_D ? 1 : ((t0 == T_O || t0 == T_M) ? 0 : 2);
        }
        const int32_t d0 = N(e0).child[0], d1 = N(e0).child[1];
        N(d0).nb[b0] = ca[a0];
        if (N(d0).child[0] >= 0 && b0 <= 1) {
            N(N(d0).child[1 - b0]).nb[3] = ca[a0];
            N(ca[a0]).nb[a0 == 0 ? c0f : c1f] = N(d0).child[1 - b0];
        }
        N(d1).nb[b1] = ca[a1];
        if (N(d1).child[0] >= 0 && b1 <= 1) {
            N(N(d1).child[1 - b1]).nb[3] = ca[a1];
            N(ca[a1]).nb[a1 == 0 ? c0f : c1f] = N(d1).child[1 - b1];
        }
    }
    return ret;
}

/* refine.c:1661 refine_path (non-recursive version) */
static void refine_path(bis_t *B, int32_t e)
{
    B->sp = 0;
    push(B, e);
    while (B->sp > 0 && !B->failed) {
        do {
            e = B->stack[--B->sp];
            if (N(e).child[0] < 0)
                break;
            if (B->sp <= 0) {
                e = -1;
                break;
            }
        } while (1);
        if (e < 0)
            break;
        divide(B, e);
        new_vertex(B, e);
        switch (update_neighbours(B, e)) {
        case 1:
            push(B, N(e).child[0]);
            break;
        case 2:
            push(B, N(e).child[1]);
            break;
        case 3:
            push(B, N(e).child[0]);
            push(B, N(e).child[1]);
            break;
        }
        /* the two neighbours sharing the refinement edge */
        for (int i = 2; i < 4; i++) {
            int32_t e0 = nbr(B, e, i);
            if (e0 < 0)
                continue;
            if (N(e0).child[0] < 0) {
                push(B, e0);
                continue;
            }
            int j = opposite_vertex(B, e, i, e0);
            if (j < 2) {
                e0 = N(e0).child[1 - j];
                N(e).nb[i] = e0;
                if (N(e0).child[0] < 0) {
                    push(B, e0);
                    continue;
                }
            }
            for (int k = 0; k < 2; k++) {
                const int32_t e1 = N(e0).child[k];
                if (N(e1).child[0] < 0)
                    continue;
                j = N(e0).verts[0] == N(e).verts[0] ? k : 1 - k;
                const int32_t ee = N(e).child[j];
                if (share_refinement_edge(B, e1, ee)) {
                    push(B, ee);
                    continue;
                }
                for (int l = 0; l < 2; l++) {
                    const int32_t c = N(e1).child[l];
                    if (N(c).child[0] >= 0 && match_neighbour(B, ee, c)) {
                        if (share_refinement_edge(B, c, ee))
                            push(B, ee);
                        break;
                    }
                }
            }
        }
    }
}

/* refine.c:2052 refine_marked_traverse, with its hanging-face repair */
static void refine_marked_traverse(bis_t *B, int32_t e)
{
    int flag = 0;
    if (B->failed)
        return;
    if (N(e).child[0] >= 0) {
        flag = 1;
        refine_marked_traverse(B, N(e).child[0]);
    }
    if (N(e).child[1] >= 0) {
        flag = 1;
        refine_marked_traverse(B, N(e).child[1]);
    }
    if (flag || N(e).mark <= 0) {
        for (int i = 0; i < 4 && !B->failed; i++) {
            if (N(e).child[0] >= 0 || N(e).child[1] >= 0)
                break;
            if (!(local(B, e, i) && N(e).nb[i] < 0))
                continue;
            /* a hanging face: refine the neighbour of the nearest ancestor
             * that has one across the matching face */
            int32_t e0 = e, e1;
            int face = i, child;
            for (;;) {
                e1 = e0;
                e0 = N(e0).parent;
                if (e0 < 0) {
                    B->failed = 1;
                    return;
                }
                switch (N(e0).type) {
                case T_D: case T_F: case T_E:
                    child = N(e1).verts[0] == N(e0).verts[0] ? 0 : 1;
                    switch (face) {
                    case 0:
                        B->failed = 1;
                        return;
                    case 1: case 2:
                        if (child == 1 && N(e0).type == T_D)
                            face = 3 - face;
                        face++;
                        break;
                    case 3:
                        face = 1 - child;
                        break;
                    }
                    break;
                case T_O:
                    child = N(e1).verts[2] == N(e0).verts[0] ? 0 : 1;
                    switch (face) {
                    case 0: case 1:
                        face = 3 - face;
                        break;
                    case 2:
                        B->failed = 1;
                        return;
                    case 3:
                        face = 1 - child;
                        break;
                    }
                    break;
                default:    /* T_M */
                    child = N(e1).verts[0] == N(e0).verts[0] ? 0 : 1;
                    switch (face) {
                    case 0:
                        if (child == 0) {
                            B->failed = 1;
                            return;
                        }
                        face = 3;
                        break;
                    case 1:
                        face = 2;
                        break;
                    case 2:
                        if (child == 1) {
                            B->failed = 1;
                            return;
                        }
                        face = 3;
                        break;
                    case 3:
                        face = 1 - child;
                        break;
                    }
                    break;
                }
                if (!local(B, e0, face)) {
                    B->failed = 1;
                    return;
                }
                if (N(e0).nb[face] >= 0)
                    break;
            }
            e0 = N(e0).nb[face];
            if (N(e0).mark < 1)
                N(e0).mark = 1;
            refine_marked_traverse(B, e0);
        }
        return;
    }
    refine_path(B, e);
    refine_marked_traverse(B, e);
}

/* refine.c:518 phgMapP2C: parent vertex / edge i -> child vertex / edge */
static void map_p2c(int type, int child, int V[4], int Em[6])
{
    static const int v0[4] = {0, -1, 1, 2}, e0[6] = {2, 0, 1, -1, -1, 3};
    static const int vd1[4] = {-1, 0, 2, 1}, ed1[6] = {2, -1, -1, 1, 0, 3};
    static const int vf1[4] = {-1, 0, 1, 2}, ef1[6] = {2, -1, -1, 0, 1, 3};
    static const int vo0[4] = {2, -1, 1, 0}, eo0[6] = {5, 3, 1, -1, -1, 0};
    static const int vo1[4] = {-1, 2, 1, 0}, eo1[6] = {5, -1, -1, 3, 1, 0};
    const int *v, *e;
    if (child == 0) {
        if (type == T_O) {
            v = vo0;
            e = eo0;
        }
        else {
            v = v0;
            e = e0;
        }
    }
    else if (type == T_D) {
        v = vd1;
        e = ed1;
    }
    else if (type == T_F || type == T_E) {
        v = vf1;
        e = ef1;
    }
    else {      /* O and M */
        v = vo1;
        e = eo1;
    }
    memcpy(V, v, 4 * sizeof(int));
    memcpy(Em, e, 6 * sizeof(int));
}

/* refine.c:2176 update_indices */
static void update_indices(bis_t *B, int32_t e)
{
    if (N(e).child[0] < 0)
        return;
    int fm[4], em[6];
    const int32_t c0 = N(e).child[0], c1 = N(e).child[1];
    int flag = N(e).verts[0] < N(e).verts[1];
    map_p2c(N(e).type, 0, fm, em);
    if (flag)
        N(c0).index = N(e).index;
    N(c0).edges[em[1]] = N(e).edges[1];
    N(c0).edges[em[2]] = N(e).edges[2];
    N(c0).edges[em[5]] = N(e).edges[5];
    if (flag)
        N(c0).edges[em[0]] = N(e).edges[0];
    N(c0).faces[3] = N(e).faces[1];
    if (flag) {
        N(c0).faces[fm[2]] = N(e).faces[2];
        N(c0).faces[fm[3]] = N(e).faces[3];
    }
    update_indices(B, c0);
    flag = !flag;
    map_p2c(N(e).type, 1, fm, em);
    if (flag)
        N(c1).index = N(e).index;
    N(c1).edges[em[3]] = N(e).edges[3];
    N(c1).edges[em[4]] = N(e).edges[4];
    N(c1).edges[em[5]] = N(e).edges[5];
    if (flag)
        N(c1).edges[em[0]] = N(e).edges[0];
    N(c1).faces[3] = N(e).faces[0];
    if (flag) {
        N(c1).faces[fm[2]] = N(e).faces[2];
        N(c1).faces[fm[3]] = N(e).faces[3];
    }
    update_indices(B, c1);
}

static const int EV[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

/* leaves in depth-first order (utils.c:2519 phgTraverseElements) */
static void collect_leaves(const bis_t *B, int32_t e, int32_t *out, int64_t *k)
{
    if (N(e).child[0] < 0 && N(e).child[1] < 0) {
        out[(*k)++] = e;
        return;
    }
    if (N(e).child[0] >= 0)
        collect_leaves(B, N(e).child[0], out, k);
    if (N(e).child[1] >= 0)
        collect_leaves(B, N(e).child[1], out, k);
}

phg_grid *phg_grid_bisect(const phg_grid *g, const int32_t *marks)
{
    if (g == NULL || marks == NULL || g->elem_edges == NULL || g->elem_faces == NULL ||
        g->elem_neigh == NULL || g->elem_type == NULL)
        return NULL;
    bis_t S, *B = &S;
    memset(B, 0, sizeof(S));
    const int64_t ne = g->nelem;
    B->nv = B->vcap = g->nvert;
    B->xyz = xalloc((size_t)g->nvert * 3 * sizeof(double));
    memcpy(B->xyz, g->verts, (size_t)g->nvert * 3 * sizeof(double));
    B->nelem = g->nelem;
    pm_init(&B->mid, 1024);
    /* the roots: the grid's elements in its traversal order; root t is
     * node t */
    int32_t *root_of = xalloc((size_t)(ne + 1) * sizeof(int32_t));  /* element -> node */
    for (int64_t t = 0; t < ne; t++) {
        const int32_t e = g->trav ? g->trav[t] : (int32_t)t;
        const int32_t r = new_node(B);
        node_t *R = &N(r);
        memcpy(R->verts, g->elem_verts + 4 * (int64_t)e, sizeof(R->verts));
        memcpy(R->edges, g->elem_edges + 6 * (int64_t)e, sizeof(R->edges));
        memcpy(R->faces, g->elem_faces + 4 * (int64_t)e, sizeof(R->faces));
        memcpy(R->bound, g->elem_bound + 4 * (int64_t)e, sizeof(R->bound));
        R->index = e;
        R->mark = marks[e];
        R->type = g->elem_type[e];
        root_of[e] = r;
    }
    for (int64_t t = 0; t < ne; t++)
        for (int f = 0; f < 4; f++) {
            const int32_t o = g->elem_neigh[4 * (int64_t)N(t).index + f];
            N(t).nb[f] = o >= 0 ? root_of[o] : -1;
        }
    for (int64_t t = 0; t < ne && !B->failed; t++)
        refine_marked_traverse(B, (int32_t)t);
    phg_grid *o = NULL;
    int32_t *leaves = NULL, *by_index = NULL;
    if (B->failed)
        goto done;
    for (int64_t t = 0; t < ne; t++)
        update_indices(B, (int32_t)t);
    /* leaves in traversal order */
    leaves = xalloc((size_t)(B->nn + 1) * sizeof(int32_t));
    int64_t nl = 0;
    for (int64_t t = 0; t < ne; t++)
        collect_leaves(B, (int32_t)t, leaves, &nl);
    if (nl != B->nelem) {
        B->failed = 1;
        goto done;
    }
    /* new edges (utils.c:4282): the first leaf in traversal order to meet
     * an unnumbered edge numbers it for every leaf around it */
    int32_t nedge = g->nedge, nface = g->nface;
    {
        pmap em;
        pm_init(&em, 2 * nl);
        for (int64_t t = 0; t < nl; t++) {
            node_t *L = &N(leaves[t]);
            for (int i = 0; i < 6; i++) {
                if (L->edges[i] != -1)
                    continue;
                int32_t a = L->verts[EV[i][0]], b = L->verts[EV[i][1]], found;
                if (a > b) {
                    const int32_t x = a;
                    a = b;
                    b = x;
                }
                int32_t *slot = pm_slot(&em, a, b, &found);
                if (!found)
                    *slot = nedge++;
                L->edges[i] = *slot;
            }
        }
        pm_free(&em);
    }
    /* new faces (utils.c:4469): a shared face at the visit of the leaf whose
     * opposite vertex has the smaller index, a boundary face at its leaf */
    for (int64_t t = 0; t < nl && !B->failed; t++) {
        const int32_t e = leaves[t];
        for (int i = 0; i < 4; i++) {
            if (N(e).faces[i] != -1)
                continue;
            if (local(B, e, i)) {
                const int32_t e1 = N(e).nb[i];
                if (e1 < 0 || N(e1).child[0] >= 0 || N(e1).child[1] >= 0) {
                    B->failed = 1;
                    break;
                }
                const int j = opposite_vertex(B, e, i, e1);
                if (j > 3) {
                    B->failed = 1;
                    break;
                }
                if (N(e).verts[i] < N(e1).verts[j]) {
                    N(e).faces[i] = N(e1).faces[j] = nface++;
                }
            }
            else {
                N(e).faces[i] = nface++;
            }
        }
    }
    if (B->failed)
        goto done;
    /* the refined grid, elements by index */
    by_index = xalloc((size_t)(nl + 1) * sizeof(int32_t));
    for (int64_t i = 0; i < nl; i++)
        by_index[i] = -1;
    for (int64_t t = 0; t < nl; t++) {
        const int32_t ix = N(leaves[t]).index;
        if (ix < 0 || ix >= nl || by_index[ix] != -1) {
            B->failed = 1;
            goto done;
        }
        by_index[ix] = leaves[t];
    }
    o = calloc(1, sizeof(phg_grid));
    if (!o)
        abort();
    o->nvert = (int32_t)B->nv;
    o->nelem = (int32_t)nl;
    o->nedge = nedge;
    o->nface = nface;
    o->verts = B->xyz;
    B->xyz = NULL;
    o->elem_verts = xalloc((size_t)nl * 4 * sizeof(int32_t));
    o->elem_edges = xalloc((size_t)nl * 6 * sizeof(int32_t));
    o->elem_faces = xalloc((size_t)nl * 4 * sizeof(int32_t));
    o->elem_neigh = xalloc((size_t)nl * 4 * sizeof(int32_t));
    o->elem_bound = xalloc((size_t)nl * 4);
    o->elem_type = xalloc((size_t)nl + 1);
    o->trav = xalloc((size_t)nl * sizeof(int32_t));
    for (int64_t i = 0; i < nl; i++) {
        const node_t *L = &N(by_index[i]);
        memcpy(o->elem_verts + 4 * i, L->verts, sizeof(L->verts));
        memcpy(o->elem_edges + 6 * i, L->edges, sizeof(L->edges));
        memcpy(o->elem_faces + 4 * i, L->faces, sizeof(L->faces));
        memcpy(o->elem_bound + 4 * i, L->bound, sizeof(L->bound));
        o->elem_type[i] = L->type;
        for (int f = 0; f < 4; f++)
            o->elem_neigh[4 * i + f] =
                local(B, by_index[i], f) && L->nb[f] >= 0 ? N(L->nb[f]).index : -1;
    }
    for (int64_t t = 0; t < nl; t++)
        o->trav[t] = N(leaves[t]).index;
    phg_grid_types(o);
done:
    free(leaves);
    free(by_index);
    free(root_of);
    free(B->n);
    free(B->xyz);
    free(B->stack);
    pm_free(&B->mid);
    return o;
}

/*------------------------------------------------------------- marking
 * mark.c:912 phgMarkElements for an element-wise (P0) error indicator,
 * refinement strategies only (this build does not coarsen).  Sums and
 * norms run over the elements in index order, as ForAllElements does. */

/* mark.c:48 part_sum, element data, refine: the sum of the indicators
 * above mx * factor */
static double part_sum(int64_t n, const double *err, double mx, double factor)
{
    const double tmp = mx * factor;
    double sum = 0.0;
    for (int64_t i = 0; i < n; i++)
        if (err[i] > tmp)
            sum += err[i];
    return sum;
}

/* mark.c:190 mark_cal_factor, refine branch: bisection on the factor until
 * the partial sum repeats an end value, meets eta, or the interval is below
 * 1e-4; *first stands for the function's static call counter */
static double cal_factor(int64_t n, const double *err, double mx, double eta, double low,
                         double up, int *first, double *low_sum, double *up_sum)
{
    for (;;) {
        const double sm = (low + up) * 0.5;
        double c2;
        if (!*first) {
            c2 = part_sum(n, err, mx, sm);
        }
        else {
            *low_sum = part_sum(n, err, mx, low);
            c2 = part_sum(n, err, mx, sm);
            *up_sum = part_sum(n, err, mx, up);
            *first = 0;
        }
        if (*low_sum == c2 || c2 == *up_sum || c2 == eta || up - low < 0.0001)
            return sm;
        if (c2 > eta) {
            *low_sum = c2;
            low = sm;
        }
        else {
            *up_sum = c2;
            up = sm;
        }
    }
}

int phg_grid_mark(const phg_grid *g, int strategy, const double *err, double theta,
                  double tol, int depth, int32_t *marks)
{
    if (g == NULL || marks == NULL || depth <= 0 ||
        (strategy != PHG_MARK_ALL && strategy != PHG_MARK_NONE && err == NULL))
        return -1;
    const int64_t n = g->nelem;
    for (int64_t i = 0; i < n; i++)        /* mark.c:330 mark_init */
        marks[i] = 0;
    switch (strategy) {
    case PHG_MARK_ALL:
        for (int64_t i = 0; i < n; i++)
            marks[i] = depth;
        return 0;
    case PHG_MARK_NONE:
        return 0;
    case PHG_MARK_MAX: {                    /* mark.c:684 mark_refine_max_elem */
        double mx = 0.0;
        for (int64_t i = 0; i < n; i++)
            if (fabs(err[i]) > mx)
                mx = fabs(err[i]);
        const double tmp = mx * theta;
        if (tmp <= 0.)
            return 0;
        for (int64_t i = 0; i < n; i++)
            if (err[i] > tmp)
                marks[i] = depth;
        return 0;
    }
    case PHG_MARK_GERS: {                   /* mark.c:406 mark_refine_gers_elem */
        if (theta < 0. || theta > 1.)
            return -1;
        double pe = 0.0, l1 = 0.0;
        for (int64_t i = 0; i < n; i++) {
            if (fabs(err[i]) > pe)
                pe = fabs(err[i]);
            l1 += fabs(err[i]);
        }
        if (pe <= 0.0)
            return 0;
        const double flag = l1 * theta;
        int first = 1;
        double ls = 0.0, us = 0.0;
        const double gama = cal_factor(n, err, pe, flag, 0.0, 1.0, &first, &ls, &us);
        pe *= gama;
        for (int64_t i = 0; i < n; i++)
            if (err[i] > pe)
                marks[i] = depth;
        return 0;
    }
    case PHG_MARK_EQDIST: {                 /* mark.c:646 mark_refine_eqdist */
        if (!(tol > 0.) || theta < 0.)
            return -1;
        const double flag = theta * tol;
        for (int64_t i = 0; i < n; i++)
            if (err[i] > flag)
                marks[i] = depth;
        return 0;
    }
    }
    return -1;
}
