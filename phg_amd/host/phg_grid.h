/* Host-side tetrahedral grid and Lagrange DOF layout with PHG's numbering.
 *
 * This is the host prerequisite of the GPU path: it produces the element ->
 * DOF maps, nodal coordinates and Dirichlet flags that the HIP assembly
 * consumes.  The numbering restates PHG's import pipeline for a flat mesh:
 *   - vertices keep file order (albert.c:351 phgImportALBERT),
 *   - edges are numbered on first encounter while elements are traversed in
 *     index order and local edges 0..5 (utils.c:4282 update_edges_callback),
 *   - a face shared by two elements is numbered when the element whose
 *     opposite vertex has the smaller global index is visited; boundary faces
 *     when their element is visited (utils.c:4499 update_faces_callback1),
 *   - boundary types of faces are OR-ed onto their vertices and edges
 *     (utils.c:962 types_callback),
 *   - DOF data are laid out vertex | edge | face | element with P3 edge pairs
 *     ordered by global vertex index (dof.c:4821 phgDofMapE2D_,
 *     dof.c:978 phgDofMapEdgeData).
 * Meshes read from ALBERT files that carry an "element type" section are
 * imported without the vertex permutation of refine.c:1916 phgRefineInit,
 * which is what PHG does for such files; files without it get that
 * permutation (Liu/Arnold initial marking: longest edge first, then the
 * marked edges of faces 0 and 1) after the numbering, as utils.c:5617 runs
 * it.
 */
#ifndef PHG_GRID_H
#define PHG_GRID_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* boundary type bits (PHG grid.h semantics for the bits used here) */
#define PHG_BT_INTERIOR   0u
#define PHG_BT_DIRICHLET  1u
#define PHG_BT_NEUMANN    2u

typedef struct phg_grid {
    int32_t nvert, nelem, nedge, nface;
    double *verts;          /* [nvert][3] */
    int32_t *elem_verts;    /* [nelem][4] */
    int32_t *elem_edges;    /* [nelem][6] */
    int32_t *elem_faces;    /* [nelem][4] */
    int32_t *elem_neigh;    /* [nelem][4], -1 on the boundary */
    uint8_t *elem_bound;    /* [nelem][4] boundary type of face i (opposite vertex i) */
    uint8_t *types_vert;    /* [nvert] */
    uint8_t *types_edge;    /* [nedge] */
    uint8_t *types_face;    /* [nface] */
    /* bisection type of each element, ALBERT code (albert.c:36): 0 DIAGONAL,
     * 1 FACE, 2 EDGE, 3 MIXED, 4 OPPOSITE */
    uint8_t *elem_type;     /* [nelem] */
    /* element index at each position of PHG's tree traversal (leaves in
     * depth-first order, utils.c:2519 phgTraverseElements); NULL: index order
     * (a flat, imported mesh) */
    int32_t *trav;          /* [nelem] or NULL */
} phg_grid;

/* Local edge -> vertex table (elem-info.c:62 phgTetEdgeVertex). */
extern const int phg_tet_edge_vertex[6][2];

/* Structured Kuhn mesh of the unit cube: n^3 cubes, 6 tetrahedra each, every
 * tetrahedron listing its cube's (0,0,0) and (1,1,1) corners first, in the
 * pattern of test/cube.dat.  bcode[6] gives the ALBERT boundary code of the
 * faces x=0, x=1, y=0, y=1, z=0, z=1 (1 Dirichlet, -1 Neumann, 0 interior).
 * Numbering is applied (phg_grid_number) before returning. */
phg_grid *phg_grid_new_kuhn(int n, const int bcode[6]);

/* ALBERT text format (albert.c:351).  The writer always emits an
 * "element type" section so the reader imports vertex order unchanged.
 * Return 0 on success. */
int phg_grid_write_albert(const phg_grid *g, const char *path);
phg_grid *phg_grid_read_albert(const char *path);

/* PHG's bisection, refine.c:2346 phgRefineMarkedElements in a serial run:
 * every element e is bisected at least marks[e] times (by element index),
 * with PHG's conforming closure (refine.c:1661 refine_path), vertex creation
 * order (refine.c:1230 new_vertex), child vertex order and types
 * (refine.c:749 divide_simplex), and index inheritance (refine.c:2176
 * update_indices): the child holding the refinement edge's smaller vertex
 * keeps its parent's element index and the halves of the parent's edges and
 * faces, the other child takes the next new element index, and new edges and
 * faces are numbered in tree traversal order (utils.c:4282, :4499).  g needs
 * its numbering and element types (any grid made by this layer has them).
 * Returns the refined grid (elements by index, with its traversal order) or
 * NULL on error. */
phg_grid *phg_grid_bisect(const phg_grid *g, const int32_t *marks);

/* mark.c:912 phgMarkElements with an element-wise error indicator err
 * (indexed by element), refinement strategies: ALL marks every element,
 * MAX those above theta * max|err| (mark.c:684), GERS the Dorfler bulk found
 * by mark.c:190's bisection on the factor with theta the bulk fraction
 * (mark.c:406), EQDIST those above theta * tol (mark.c:646).  marks[e] =
 * depth for marked elements, 0 otherwise.  Returns 0, or -1 on bad
 * arguments. */
enum { PHG_MARK_NONE = 0, PHG_MARK_ALL = 1, PHG_MARK_MAX = 2, PHG_MARK_GERS = 3,
       PHG_MARK_EQDIST = 4 };
int phg_grid_mark(const phg_grid *g, int strategy, const double *err, double theta,
                  double tol, int depth, int32_t *marks);

/* Computes neighbours, edge and face numbering and boundary types. */
void phg_grid_number(phg_grid *g);
/* Boundary types of vertices, edges and faces from the elements' face types
 * and neighbours (utils.c:962 types_callback). */
void phg_grid_types(phg_grid *g);
void phg_grid_free(phg_grid *g);

/* Lagrange Pk (k = 1, 2, 3) scalar DOF layout. */
int phg_dof_nbas(int order);
int64_t phg_dof_count(const phg_grid *g, int order);
/* map[nelem][nbas]: element-local basis -> DOF index (global ordering). */
void phg_dof_map(const phg_grid *g, int order, int32_t *map);
/* xyz[ndof][3]: nodal coordinates, the node of each DOF evaluated from the
 * first element (in index order) that holds it, with PHG's barycentric
 * combination in that element's local vertex order (dof.c:5109,
 * dof.c:2210). */
void phg_dof_coords(const phg_grid *g, int order, double *xyz);
/* flags[ndof] = 1 if the DOF's geometric entity carries a boundary bit in
 * `mask` (dof-utils.c:1269 phgDofDirichletBC__, nodal path). */
void phg_dof_bdry(const phg_grid *g, int order, unsigned mask, uint8_t *flags);

#ifdef __cplusplus
}
#endif
#endif
