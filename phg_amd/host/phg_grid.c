/* Host grid, PHG import numbering and Lagrange P1-P3 DOF layout.
 * The rules restated here are listed in phg_grid.h. */
#include "phg_grid.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

const int phg_tet_edge_vertex[6][2] = {{0, 1}, {0, 2}, {0, 3},
                                       {1, 2}, {1, 3}, {2, 3}};

/* vertices of face i (opposite vertex i), in increasing local order */
static const int face_verts[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3},
                                     {0, 1, 2}};
/* edges lying on face i (utils.c:962 types_callback) */
static const int face_edges[4][3] = {{3, 4, 5}, {1, 2, 5}, {0, 2, 4},
                                     {0, 1, 3}};

static void *xmalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (p == NULL) {
        fprintf(stderr, "phg_grid: out of memory (%zu bytes)\n", n);
        abort();
    }
    return p;
}

/*--------------------- open-addressing hash of int triples -----------------*/

typedef struct {
    int64_t cap;
    int32_t *keys;   /* [cap][3], keys[0] == -1 marks an empty slot */
    int64_t *vals;
} khash;

static void kh_init(khash *h, int64_t n)
{
    int64_t cap = 16;
    while (cap < 2 * n)
        cap <<= 1;
    h->cap = cap;
    h->keys = xmalloc((size_t)cap * 3 * sizeof(int32_t));
    h->vals = xmalloc((size_t)cap * sizeof(int64_t));
    for (int64_t i = 0; i < cap; i++)
        h->keys[3 * i] = -1;
}

static void kh_free(khash *h)
{
    free(h->keys);
    free(h->vals);
}

static uint64_t kh_hash(int32_t a, int32_t b, int32_t c)
{
    uint64_t x = (uint64_t)(uint32_t)a * 0x9E3779B97F4A7C15ull;
    x ^= (uint64_t)(uint32_t)b * 0xC2B2AE3D27D4EB4Full + (x >> 29);
    x ^= (uint64_t)(uint32_t)c * 0x165667B19E3779F9ull + (x >> 32);
    x ^= x >> 31;
    return x;
}

/* returns a pointer to the value slot; *found tells whether it existed */
static int64_t *kh_get(khash *h, int32_t a, int32_t b, int32_t c, int *found)
{
    uint64_t m = (uint64_t)h->cap - 1, i = kh_hash(a, b, c) & m;
    for (;;) {
        int32_t *k = h->keys + 3 * i;
        if (k[0] == -1) {
            k[0] = a;
            k[1] = b;
            k[2] = c;
            *found = 0;
            return h->vals + i;
        }
        if (k[0] == a && k[1] == b && k[2] == c) {
            *found = 1;
            return h->vals + i;
        }
        i = (i + 1) & m;
    }
}

static void sort3(int32_t *v)
{
    int32_t t;
    if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
    if (v[1] > v[2]) { t = v[1]; v[1] = v[2]; v[2] = t; }
    if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
}

/*------------------------------ construction ------------------------------*/

static phg_grid *grid_alloc(int nvert, int nelem)
{
    phg_grid *g = xmalloc(sizeof(*g));
    memset(g, 0, sizeof(*g));
    g->nvert = nvert;
    g->nelem = nelem;
    g->verts = xmalloc((size_t)nvert * 3 * sizeof(double));
    g->elem_verts = xmalloc((size_t)nelem * 4 * sizeof(int32_t));
    g->elem_bound = xmalloc((size_t)nelem * 4);
    memset(g->elem_bound, 0, (size_t)nelem * 4);
    g->elem_type = xmalloc((size_t)nelem + 1);
    memset(g->elem_type, 0, (size_t)nelem + 1);
    g->trav = NULL;
    return g;
}

void phg_grid_free(phg_grid *g)
{
    if (g == NULL)
        return;
    free(g->verts);
    free(g->elem_verts);
    free(g->elem_edges);
    free(g->elem_faces);
    free(g->elem_neigh);
    free(g->elem_bound);
    free(g->types_vert);
    free(g->types_edge);
    free(g->types_face);
    free(g->elem_type);
    free(g->trav);
    free(g);
}

static uint8_t bdry_from_albert(int n)
{
    /* albert.c:41 BDRY_TYPE_FROM_ALBERT for the codes used here */
    if (n == 0)
        return PHG_BT_INTERIOR;
    if (n < 0)
        return PHG_BT_NEUMANN;
    return PHG_BT_DIRICHLET;
}

static int bdry_to_albert(uint8_t b)
{
    if (b & PHG_BT_DIRICHLET)
        return 1;
    if (b & PHG_BT_NEUMANN)
        return -1;
    return 0;
}

phg_grid *phg_grid_new_kuhn(int n, const int bcode[6])
{
    /* cube corner labels as in test/cube.dat:
     * 0=(0,0,0) 1=(1,0,0) 2=(0,0,1) 3=(1,0,1) 4=(1,1,0) 5=(1,1,1)
     * 6=(0,1,0) 7=(0,1,1) */
    static const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 0, 1},
                                     {1, 0, 1}, {1, 1, 0}, {1, 1, 1},
                                     {0, 1, 0}, {0, 1, 1}};
    static const int tets[6][4] = {{0, 5, 4, 1}, {0, 5, 3, 1}, {0, 5, 3, 2},
                                   {0, 5, 4, 6}, {0, 5, 7, 6}, {0, 5, 7, 2}};
    int n1 = n + 1;
    phg_grid *g = grid_alloc(n1 * n1 * n1, 6 * n * n * n);
    int64_t e = 0;

    for (int k = 0; k <= n; k++)
        for (int j = 0; j <= n; j++)
            for (int i = 0; i <= n; i++) {
                double *x = g->verts + 3 * ((int64_t)i + n1 * ((int64_t)j + n1 * k));
                x[0] = (double)i / (double)n;
                x[1] = (double)j / (double)n;
                x[2] = (double)k / (double)n;
            }

    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++)
                for (int t = 0; t < 6; t++, e++) {
                    int32_t *v = g->elem_verts + 4 * e;
                    int c[4][3];
                    for (int a = 0; a < 4; a++) {
                        const int *o = corner[tets[t][a]];
                        c[a][0] = i + o[0];
                        c[a][1] = j + o[1];
                        c[a][2] = k + o[2];
                        v[a] = c[a][0] + n1 * (c[a][1] + n1 * c[a][2]);
                    }
                    /* face f (opposite vertex f) lies on a cube side iff its
                     * three vertices share that coordinate value */
                    for (int f = 0; f < 4; f++) {
                        const int *fv = face_verts[f];
                        uint8_t b = PHG_BT_INTERIOR;
                        for (int d = 0; d < 3; d++) {
                            int x0 = c[fv[0]][d];
                            if (x0 == c[fv[1]][d] && x0 == c[fv[2]][d]) {
                                if (x0 == 0)
                                    b = bdry_from_albert(bcode[2 * d]);
                                else if (x0 == n)
                                    b = bdry_from_albert(bcode[2 * d + 1]);
                            }
                        }
                        g->elem_bound[4 * e + f] = b;
                    }
                }
    phg_grid_number(g);
    return g;
}

/*------------------------------- numbering --------------------------------*/

void phg_grid_number(phg_grid *g)
{
    int64_t ne = g->nelem;
    khash h;
    int found;

    free(g->elem_neigh);
    free(g->elem_edges);
    free(g->elem_faces);
    g->elem_neigh = xmalloc((size_t)ne * 4 * sizeof(int32_t));
    g->elem_edges = xmalloc((size_t)ne * 6 * sizeof(int32_t));
    g->elem_faces = xmalloc((size_t)ne * 4 * sizeof(int32_t));

    /* neighbours: match the vertex triples of faces */
    kh_init(&h, 4 * ne);
    for (int64_t e = 0; e < ne; e++) {
        const int32_t *v = g->elem_verts + 4 * e;
        for (int f = 0; f < 4; f++) {
            int32_t k[3] = {v[face_verts[f][0]], v[face_verts[f][1]],
                            v[face_verts[f][2]]};
            sort3(k);
            int64_t *slot = kh_get(&h, k[0], k[1], k[2], &found);
            g->elem_neigh[4 * e + f] = -1;
            if (!found) {
                *slot = 4 * e + f;
            }
            else {
                int64_t o = *slot;
                g->elem_neigh[4 * e + f] = (int32_t)(o / 4);
                g->elem_neigh[o] = (int32_t)e;
            }
        }
    }
    kh_free(&h);

    /* edges: first encounter in (element, local edge) order */
    kh_init(&h, 6 * ne / 4 + 16);
    g->nedge = 0;
    for (int64_t e = 0; e < ne; e++) {
        const int32_t *v = g->elem_verts + 4 * e;
        for (int i = 0; i < 6; i++) {
            int32_t a = v[phg_tet_edge_vertex[i][0]], b = v[phg_tet_edge_vertex[i][1]];
            if (a > b) { int32_t t = a; a = b; b = t; }
            int64_t *slot = kh_get(&h, a, b, 0, &found);
            if (!found)
                *slot = g->nedge++;
            g->elem_edges[6 * e + i] = (int32_t)*slot;
        }
    }
    kh_free(&h);

    /* faces: a shared face is numbered at the visit of the element whose
     * opposite vertex has the smaller global index */
    for (int64_t i = 0; i < 4 * ne; i++)
        g->elem_faces[i] = -1;
    g->nface = 0;
    for (int64_t e = 0; e < ne; e++) {
        const int32_t *v = g->elem_verts + 4 * e;
        for (int f = 0; f < 4; f++) {
            if (g->elem_faces[4 * e + f] != -1)
                continue;
            int32_t e1 = g->elem_neigh[4 * e + f];
            if (e1 >= 0) {
                const int32_t *v1 = g->elem_verts + 4 * (int64_t)e1;
                int j;
                for (j = 0; j < 4; j++)
                    if (g->elem_neigh[4 * (int64_t)e1 + j] == (int32_t)e) {
                        /* the face of e1 shared with e: its opposite vertex
                         * is the one not among e's face vertices */
                        int32_t w = v1[j];
                        if (w != v[face_verts[f][0]] && w != v[face_verts[f][1]]
                            && w != v[face_verts[f][2]])
                            break;
                    }
                if (j >= 4) {
                    fprintf(stderr, "phg_grid: inconsistent neighbours\n");
                    abort();
                }
                if (v[f] < v1[j]) {
                    g->elem_faces[4 * e + f] = g->nface;
                    g->elem_faces[4 * (int64_t)e1 + j] = g->nface;
                    g->nface++;
                }
            }
            else {
                g->elem_faces[4 * e + f] = g->nface++;
            }
        }
    }

    phg_grid_types(g);
}

/* boundary types OR-ed from faces onto vertices and edges (utils.c:962) */
void phg_grid_types(phg_grid *g)
{
    const int64_t ne = g->nelem;
    free(g->types_vert);
    free(g->types_edge);
    free(g->types_face);
    g->types_vert = xmalloc((size_t)g->nvert);
    g->types_edge = xmalloc((size_t)g->nedge);
    g->types_face = xmalloc((size_t)g->nface);
    memset(g->types_vert, 0, (size_t)g->nvert);
    memset(g->types_edge, 0, (size_t)g->nedge);
    memset(g->types_face, 0, (size_t)g->nface);
    for (int64_t e = 0; e < ne; e++) {
        const int32_t *v = g->elem_verts + 4 * e;
        for (int f = 0; f < 4; f++) {
            uint8_t b = g->elem_bound[4 * e + f];
            if (g->elem_neigh[4 * e + f] >= 0)
                b = PHG_BT_INTERIOR;
            if (b == PHG_BT_INTERIOR)
                continue;
            for (int a = 0; a < 3; a++) {
                g->types_vert[v[face_verts[f][a]]] |= b;
                g->types_edge[g->elem_edges[6 * e + face_edges[f][a]]] |= b;
            }
            g->types_face[g->elem_faces[4 * e + f]] |= b;
        }
    }
}

/*-------------------------------- ALBERT ----------------------------------*/

int phg_grid_write_albert(const phg_grid *g, const char *path)
{
    FILE *f = fopen(path, "w");
    if (f == NULL)
        return -1;
    fprintf(f, "DIM: 3\nDIM_OF_WORLD: 3\n");
    fprintf(f, "number of vertices: %d\nnumber of elements: %d\n\n",
            g->nvert, g->nelem);
    fprintf(f, "vertex coordinates:\n");
    for (int64_t i = 0; i < g->nvert; i++)
        fprintf(f, "%.17g %.17g %.17g\n", g->verts[3 * i], g->verts[3 * i + 1],
                g->verts[3 * i + 2]);
    /* elements in traversal order, as albert.c:537 phgExportALBERT writes
     * them */
    fprintf(f, "\nelement vertices:\n");
    for (int64_t t = 0; t < g->nelem; t++) {
        const int64_t e = g->trav ? g->trav[t] : t;
        const int32_t *v = g->elem_verts + 4 * e;
        fprintf(f, "%d %d %d %d\n", v[0], v[1], v[2], v[3]);
    }
    fprintf(f, "\nelement boundaries:\n");
    for (int64_t t = 0; t < g->nelem; t++) {
        const int64_t e = g->trav ? g->trav[t] : t;
        const uint8_t *b = g->elem_bound + 4 * e;
        fprintf(f, "%d %d %d %d\n", bdry_to_albert(b[0]), bdry_to_albert(b[1]),
                bdry_to_albert(b[2]), bdry_to_albert(b[3]));
    }
    fprintf(f, "\nelement type:\n");
    for (int64_t t = 0; t < g->nelem; t++)
        fprintf(f, "%d\n", g->elem_type[g->trav ? g->trav[t] : t]);
    return fclose(f) == 0 ? 0 : -1;
}

static int read_line(FILE *f, char *buf, int n)
{
    return fgets(buf, n, f) != NULL;
}

/* refine.c:1829 permute_vertices: local vertices v0 and v1 of element e
 * trade places together with their faces (face i lies opposite vertex i),
 * neighbours and boundary types; each local edge takes the global edge of
 * the vertex pair it now joins (the two edge pairs of the switch at
 * refine.c:1862). */
static void permute_vertices(phg_grid *g, int64_t e, int v0, int v1)
{
    if (v0 == v1)
        return;
    int32_t *v = g->elem_verts + 4 * e, *ed = g->elem_edges + 6 * e;
    int32_t *fa = g->elem_faces + 4 * e, *nb = g->elem_neigh + 4 * e;
    uint8_t *bt = g->elem_bound + 4 * e;
    int sigma[4] = {0, 1, 2, 3}, t;
    int32_t old[6], x;
    sigma[v0] = v1;
    sigma[v1] = v0;
    x = v[v0]; v[v0] = v[v1]; v[v1] = x;
    x = fa[v0]; fa[v0] = fa[v1]; fa[v1] = x;
    x = nb[v0]; nb[v0] = nb[v1]; nb[v1] = x;
    t = bt[v0]; bt[v0] = bt[v1]; bt[v1] = (uint8_t)t;
    memcpy(old, ed, sizeof(old));
    for (int k = 0; k < 6; k++) {
        int a = sigma[phg_tet_edge_vertex[k][0]], b = sigma[phg_tet_edge_vertex[k][1]];
        for (int m = 0; m < 6; m++)
            if ((phg_tet_edge_vertex[m][0] == a && phg_tet_edge_vertex[m][1] == b) ||
                (phg_tet_edge_vertex[m][0] == b && phg_tet_edge_vertex[m][1] == a))
                ed[m] = old[k];
    }
}

/* refine.c:1814 comp_edge: squared length, then global edge index */
static int comp_edge(const double *len, const int32_t *ed, int i0, int i1)
{
    const int32_t e0 = ed[i0], e1 = ed[i1];
    const double d = len[e0] - len[e1];
    if (d == 0)
        return e0 > e1 ? 1 : (e0 < e1 ? -1 : 0);
    return d < 0 ? -1 : 1;
}

/* refine.c:1916 phgRefineInit with do_marking (Liu/Arnold initial marking),
 * run by PHG on meshes imported without refinement types (utils.c:5617,
 * after the edges and faces are numbered): per element the longest edge
 * becomes edge 0-1, then the marked edges of faces 0 and 1 pick one of the
 * permutations of refine.c:1987-2021.  Global numbers do not change; the
 * elements' local order does. */
static void refine_init_marking(phg_grid *g)
{
    double *len = xmalloc((size_t)(g->nedge > 0 ? g->nedge : 1) * sizeof(double));
    for (int64_t i = 0; i < g->nedge; i++)
        len[i] = -1.0;
    for (int64_t e = 0; e < g->nelem; e++)
        for (int k = 0; k < 6; k++) {
            const int32_t id = g->elem_edges[6 * e + k];
            if (len[id] >= 0)
                continue;
            const double *a = g->verts + 3 * g->elem_verts[4 * e + phg_tet_edge_vertex[k][0]];
            const double *b = g->verts + 3 * g->elem_verts[4 * e + phg_tet_edge_vertex[k][1]];
            const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
            len[id] = dx * dx + dy * dy + dz * dz;
        }
    for (int64_t e = 0; e < g->nelem; e++) {
        const int32_t *ed = g->elem_edges + 6 * e;
        int k = 0, j;
        for (j = 1; j < 6; j++)
            if (comp_edge(len, ed, k, j) < 0)
                k = j;
        permute_vertices(g, e, 0, phg_tet_edge_vertex[k][0]);
        permute_vertices(g, e, 1, phg_tet_edge_vertex[k][1]);
        j = 3;                          /* marked edge of face 0 */
        if (comp_edge(len, ed, j, 4) < 0)
            j = 4;
        if (comp_edge(len, ed, j, 5) < 0)
            j = 5;
        k = 1;                          /* marked edge of face 1 */
        if (comp_edge(len, ed, k, 2) < 0)
            k = 2;
        if (comp_edge(len, ed, k, 5) < 0)
            k = 5;
        /* types as ALBERT codes: D 0, F 1, M 3, O 4 */
        switch (j * 10 + k) {
        case 31:
            g->elem_type[e] = 1;
            break;
        case 32:
            g->elem_type[e] = 0;
            permute_vertices(g, e, 2, 3);
            break;
        case 35:
            g->elem_type[e] = 3;
            permute_vertices(g, e, 0, 1);
            break;
        case 41:
            g->elem_type[e] = 0;
            break;
        case 42:
            g->elem_type[e] = 1;
            permute_vertices(g, e, 2, 3);
            break;
        case 45:
            g->elem_type[e] = 3;
            permute_vertices(g, e, 2, 3);
            permute_vertices(g, e, 0, 1);
            break;
        case 51:
            g->elem_type[e] = 3;
            break;
        case 52:
            g->elem_type[e] = 3;
            permute_vertices(g, e, 2, 3);
            break;
        default:                        /* 55 */
            g->elem_type[e] = 4;
            break;
        }
    }
    free(len);
}

phg_grid *phg_grid_read_albert(const char *path)
{
    FILE *f = fopen(path, "r");
    char line[512];
    int nv = -1, ne = -1, typed = 0;
    phg_grid *g = NULL;

    if (f == NULL)
        return NULL;
    while (read_line(f, line, sizeof(line))) {
        char *p = line, *q;
        while (*p == ' ' || *p == '\t')
            p++;
        if (*p == '\n' || *p == '\0' || *p == '#')
            continue;
        if ((q = strchr(p, ':')) == NULL)
            continue;
        *q = '\0';
        if (!strcmp(p, "number of vertices"))
            nv = atoi(q + 1);
        else if (!strcmp(p, "number of elements"))
            ne = atoi(q + 1);
        else if (!strcmp(p, "vertex coordinates") ||
                 !strcmp(p, "element vertices") ||
                 !strcmp(p, "element boundaries") ||
                 !strcmp(p, "element neighbours") ||
                 !strcmp(p, "element type")) {
            if (nv < 0 || ne < 0)
                goto fail;
            if (g == NULL)
                g = grid_alloc(nv, ne);
            int what = !strcmp(p, "vertex coordinates") ? 0 :
                       !strcmp(p, "element vertices") ? 1 :
                       !strcmp(p, "element boundaries") ? 2 :
                       !strcmp(p, "element type") ? 4 : 3;
            if (!strcmp(p, "element type"))
                typed = 1;
            int64_t cnt = what == 0 ? nv : ne;
            for (int64_t i = 0; i < cnt; i++) {
                if (!read_line(f, line, sizeof(line)))
                    goto fail;
                if (what == 0) {
                    double *x = g->verts + 3 * i;
                    if (sscanf(line, "%lf %lf %lf", x, x + 1, x + 2) != 3)
                        goto fail;
                }
                else if (what == 1) {
                    int32_t *v = g->elem_verts + 4 * i;
                    if (sscanf(line, "%d %d %d %d", v, v + 1, v + 2, v + 3) != 4)
                        goto fail;
                    for (int a = 0; a < 4; a++)
                        if (v[a] < 0 || v[a] >= nv)
                            goto fail;
                }
                else if (what == 2) {
                    int b[4];
                    if (sscanf(line, "%d %d %d %d", b, b + 1, b + 2, b + 3) != 4)
                        goto fail;
                    for (int a = 0; a < 4; a++)
                        g->elem_bound[4 * i + a] = bdry_from_albert(b[a]);
                }
                else if (what == 4) {
                    int t;
                    if (sscanf(line, "%d", &t) != 1 || t < 0 || t > 4)
                        goto fail;
                    g->elem_type[i] = (uint8_t)t;
                }
            }
        }
    }
    fclose(f);
    if (g == NULL)
        return NULL;
    phg_grid_number(g);
    /* albert.c:461 returns whether the file had element types; without them
     * phgImport_ runs the initial marking (utils.c:5617 phgRefineInit(g,
     * !flag)), with them the file's vertex order stands */
    if (!typed)
        refine_init_marking(g);
    return g;
fail:
    fclose(f);
    phg_grid_free(g);
    return NULL;
}

/*------------------------------ DOF layout --------------------------------*/

int phg_dof_nbas(int order)
{
    switch (order) {
    case 1: return 4;
    case 2: return 10;
    case 3: return 20;
    case 4: return 35;
    }
    return -1;
}

int64_t phg_dof_count(const phg_grid *g, int order)
{
    switch (order) {
    case 1: return g->nvert;
    case 2: return (int64_t)g->nvert + g->nedge;
    case 3: return (int64_t)g->nvert + 2 * (int64_t)g->nedge + g->nface;
    case 4: return (int64_t)g->nvert + 3 * (int64_t)g->nedge + 3 * (int64_t)g->nface + g->nelem;
    }
    return -1;
}

void phg_dof_map(const phg_grid *g, int order, int32_t *map)
{
    int nb = phg_dof_nbas(order);
    int64_t v0 = g->nvert, f0 = (int64_t)g->nvert + 2 * (int64_t)g->nedge;

    for (int64_t e = 0; e < g->nelem; e++) {
        const int32_t *v = g->elem_verts + 4 * e;
        const int32_t *ed = g->elem_edges + 6 * e;
        int32_t *m = map + (int64_t)nb * e;
        for (int k = 0; k < 4; k++)
            m[k] = v[k];
        if (order == 2) {
            for (int k = 0; k < 6; k++)
                m[4 + k] = (int32_t)(v0 + ed[k]);
        }
        else if (order == 3) {
            for (int k = 0; k < 6; k++) {
                /* phgDofMapEdgeData: reversed pair when the edge runs from
                 * the larger to the smaller global vertex */
                int rev = v[phg_tet_edge_vertex[k][0]] > v[phg_tet_edge_vertex[k][1]];
                m[4 + 2 * k + 0] = (int32_t)(v0 + 2 * (int64_t)ed[k] + (rev ? 1 : 0));
                m[4 + 2 * k + 1] = (int32_t)(v0 + 2 * (int64_t)ed[k] + (rev ? 0 : 1));
            }
            for (int k = 0; k < 4; k++)
                m[16 + k] = (int32_t)(f0 + g->elem_faces[4 * e + k]);
        }
        else if (order == 4) {
            /* dof.c:4821 phgDofMapE2D_: vertex, edge, face, element blocks */
            const int64_t f4 = (int64_t)g->nvert + 3 * (int64_t)g->nedge,
                          e4 = f4 + 3 * (int64_t)g->nface;
            for (int k = 0; k < 6; k++) {
                /* dof.c:978 phgDofMapEdgeData: reversed when the edge runs
                 * from the larger to the smaller global vertex */
                int rev = v[phg_tet_edge_vertex[k][0]] > v[phg_tet_edge_vertex[k][1]];
                for (int t = 0; t < 3; t++)
                    m[4 + 3 * k + t] = (int32_t)(v0 + 3 * (int64_t)ed[k] + (rev ? 2 - t : t));
            }
            for (int k = 0; k < 4; k++) {
                /* lagrange.c:1571 P4_map: the face's node t (near its
                 * vertex t) goes to the slot of that vertex's rank among the
                 * face's global vertex indices */
                int32_t a = v[face_verts[k][0]], b = v[face_verts[k][1]],
                        c = v[face_verts[k][2]];
                int M[3] = {0, 0, 0};
                if (a > b) M[0]++; else M[1]++;
                if (a > c) M[0]++; else M[2]++;
                if (b > c) M[1]++; else M[2]++;
                for (int t = 0; t < 3; t++)
                    m[22 + 3 * k + t] = (int32_t)(f4 + 3 * (int64_t)g->elem_faces[4 * e + k] + M[t]);
            }
            m[34] = (int32_t)(e4 + e);
        }
    }
}

/* barycentric coordinates of local node `b` of Pk (dof.c:4925
 * phgDofGetElementBasisInfo_ with the points tables of lagrange.c:1646) */
static void node_lambda(int order, int b, double lam[4])
{
    static const int fv[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    lam[0] = lam[1] = lam[2] = lam[3] = 0.0;
    if (b < 4) {
        lam[b] = 1.0;
        return;
    }
    b -= 4;
    if (order == 4) {
        /* lagrange.c:1671 P4_points: edge nodes from the first vertex,
         * face node t near face vertex t, the centroid */
        if (b < 18) {
            int k = b / 3, t = b % 3;
            lam[phg_tet_edge_vertex[k][0]] = 0.75 - 0.25 * t;
            lam[phg_tet_edge_vertex[k][1]] = 0.25 + 0.25 * t;
            return;
        }
        b -= 18;
        if (b < 12) {
            int f = b / 3, t = b % 3;
            for (int i = 0; i < 3; i++)
                lam[fv[f][i]] = i == t ? 0.5 : 0.25;
            return;
        }
        lam[0] = lam[1] = lam[2] = lam[3] = 0.25;
        return;
    }
    if (order == 2) {
        lam[phg_tet_edge_vertex[b][0]] = 0.5;
        lam[phg_tet_edge_vertex[b][1]] = 0.5;
        return;
    }
    if (b < 12) {
        int k = b / 2, l = b % 2;
        lam[phg_tet_edge_vertex[k][0]] = l == 0 ? 2. / 3. : 1. / 3.;
        lam[phg_tet_edge_vertex[k][1]] = l == 0 ? 1. / 3. : 2. / 3.;
        return;
    }
    b -= 12;
    lam[fv[b][0]] = 1. / 3.;
    lam[fv[b][1]] = 1. / 3.;
    lam[fv[b][2]] = 1. / 3.;
}

void phg_dof_coords(const phg_grid *g, int order, double *xyz)
{
    int nb = phg_dof_nbas(order);
    int64_t nd = phg_dof_count(g, order);
    int32_t *map = xmalloc((size_t)g->nelem * nb * sizeof(int32_t));
    uint8_t *done = xmalloc((size_t)nd);

    memset(done, 0, (size_t)nd);
    phg_dof_map(g, order, map);
    for (int64_t e = 0; e < g->nelem; e++) {
        const int32_t *v = g->elem_verts + 4 * e;
        for (int b = 0; b < nb; b++) {
            int32_t d = map[(int64_t)nb * e + b];
            double lam[4], x = 0., y = 0., z = 0.;
            if (done[d])
                continue;
            done[d] = 1;
            node_lambda(order, b, lam);
            /* geom.c:1259 phgGeomLambda2XYZ */
            for (int i = 0; i < 4; i++) {
                const double *c = g->verts + 3 * (int64_t)v[i];
                x += c[0] * lam[i];
                y += c[1] * lam[i];
                z += c[2] * lam[i];
            }
            xyz[3 * (int64_t)d] = x;
            xyz[3 * (int64_t)d + 1] = y;
            xyz[3 * (int64_t)d + 2] = z;
        }
    }
    free(done);
    free(map);
}

void phg_dof_bdry(const phg_grid *g, int order, unsigned mask, uint8_t *flags)
{
    int64_t nd = phg_dof_count(g, order);
    int64_t v0 = g->nvert, f0 = (int64_t)g->nvert + 2 * (int64_t)g->nedge;

    for (int64_t i = 0; i < g->nvert; i++)
        flags[i] = (g->types_vert[i] & mask) != 0;
    if (order == 2) {
        for (int64_t i = 0; i < g->nedge; i++)
            flags[v0 + i] = (g->types_edge[i] & mask) != 0;
    }
    else if (order == 3) {
        for (int64_t i = 0; i < g->nedge; i++)
            flags[v0 + 2 * i] = flags[v0 + 2 * i + 1] =
                (g->types_edge[i] & mask) != 0;
        for (int64_t i = 0; i < g->nface; i++)
            flags[f0 + i] = (g->types_face[i] & mask) != 0;
    }
    else if (order == 4) {
        /* element-interior nodes are never boundary DOFs */
        const int64_t f4 = v0 + 3 * (int64_t)g->nedge, e4 = f4 + 3 * (int64_t)g->nface;
        for (int64_t i = 0; i < g->nedge; i++)
            flags[v0 + 3 * i] = flags[v0 + 3 * i + 1] = flags[v0 + 3 * i + 2] =
                (g->types_edge[i] & mask) != 0;
        for (int64_t i = 0; i < g->nface; i++)
            flags[f4 + 3 * i] = flags[f4 + 3 * i + 1] = flags[f4 + 3 * i + 2] =
                (g->types_face[i] & mask) != 0;
        for (int64_t i = 0; i < g->nelem; i++)
            flags[e4 + i] = 0;
    }
    (void)nd;
}
