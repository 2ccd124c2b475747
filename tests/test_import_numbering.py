"""PHG's numbering of an imported mesh, against values derived by hand from
the reference's rules (no code of this repository computed them):

  edges   utils.c:4282 update_edges_callback: elements in index order, local
          edges 0..5 = (0,1) (0,2) (0,3) (1,2) (1,3) (2,3) (elem-info.c:62);
          an edge not yet numbered gets the next number in every element
          sharing it;
  faces   utils.c:4474 update_faces_callback1: elements in index order, face i
          opposite vertex i; a boundary face gets the next number, an interior
          face when e->verts[i] < e1->verts[j] (j its index in the neighbour);
  marking refine.c:1916 phgRefineInit, run by utils.c:5617 after the two steps
          above when the file has no element types: longest edge (refine.c:1814
          comp_edge: squared length, ties by global edge number) becomes edge
          0-1 via refine.c:1829 permute_vertices, then the marked edges of faces
          0 and 1 choose the permutations of refine.c:1987-2021;
  P2 map  dof.c:4821 phgDofMapE2D_: vertices, then nvert + edge number.
"""
from __future__ import annotations

import os

import numpy as np

import phg_amd as pa

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_cube_dat_edges_and_faces():
    """test/cube.dat: 8 vertices, 6 tets, all of the form (0, 5, a, b)."""
    g = pa.Grid.read_albert(os.path.join(GOLDEN, "cube.dat"))
    assert (g.nvert, g.nelem, g.nedge, g.nface) == (8, 6, 19, 18)
    # edges by first encounter (e, local edge):
    #   e0 (0 5 4 1): (0,5)=0 (0,4)=1 (0,1)=2 (5,4)=3 (5,1)=4 (4,1)=5
    #   e1 (0 5 3 1): (0,3)=6 (5,3)=7 (3,1)=8, the rest seen
    #   e2 (0 5 3 2): (0,2)=9 (5,2)=10 (3,2)=11
    #   e3 (0 5 4 6): (0,6)=12 (5,6)=13 (4,6)=14
    #   e4 (0 5 7 6): (0,7)=15 (5,7)=16 (7,6)=17
    #   e5 (0 5 7 2): (7,2)=18
    edges = [[0, 1, 2, 3, 4, 5],
             [0, 6, 2, 7, 4, 8],
             [0, 6, 9, 7, 10, 11],
             [0, 1, 12, 3, 13, 14],
             [0, 15, 12, 16, 13, 17],
             [0, 15, 9, 16, 10, 18]]
    # faces 0 and 1 of every element are on the boundary (numbered on their
    # visit); interior faces: e0.f2/e1.f2 at e1 (3 < 4), e0.f3/e3.f3 at e0
    # (1 < 6), e1.f3/e2.f3 at e1 (1 < 2), e2.f2/e5.f2 at e2 (3 < 7), e3.f2/e4.f2
    # at e3 (4 < 7), e4.f3/e5.f3 at e5 (2 < 6):
    #   e0: b0 b1 skip 2   e1: 3 4 5 6   e2: 7 8 9 .   e3: 10 11 12 .
    #   e4: 13 14 . skip   e5: 15 16 . 17
    faces = [[0, 1, 5, 2],
             [3, 4, 5, 6],
             [7, 8, 9, 6],
             [10, 11, 12, 2],
             [13, 14, 12, 17],
             [15, 16, 9, 17]]
    # no element type section, so the marking runs; in every element the
    # longest edge (0,5), squared length 3, is already edge 0, face 0's marked
    # edge is local 4 (squared length 2 against 1, 1) and face 1's local 1:
    # case 41 (DIAGONAL), no permutation, file order stands
    np.testing.assert_array_equal(g.elem_verts(), [[0, 5, 4, 1], [0, 5, 3, 1], [0, 5, 3, 2],
                                                   [0, 5, 4, 6], [0, 5, 7, 6], [0, 5, 7, 2]])
    np.testing.assert_array_equal(g.elem_edges(), edges)
    np.testing.assert_array_equal(g.elem_faces(), faces)
    m = g.dof_map(2).reshape(6, 10)
    np.testing.assert_array_equal(m[:, 4:], np.array(edges) + 8)


_TET = """DIM: 3
DIM_OF_WORLD: 3
number of vertices: 4
number of elements: 1

vertex coordinates:
 0.0 0.0 0.0
 1.0 0.0 0.0
 0.0 2.0 0.0
 0.0 0.0 3.0

element vertices:
 0 1 2 3

element boundaries:
 1 1 1 1

element neighbours:
 -1 -1 -1 -1
"""


def test_initial_marking_permutes_an_untyped_tet(tmp_path):
    """One tet whose longest edge is local (2,3).  Squared lengths: (0,1) 1,
    (0,2) 4, (0,3) 9, (1,2) 5, (1,3) 10, (2,3) 13; edges 0..5 and faces
    0..3 in local order before the marking.
      longest: local 5 = (2,3).  permute_vertices(0, 2): verts 2 1 0 3,
        edges 3 1 5 0 4 2, faces 2 1 0 3.  permute_vertices(1, 3): verts
        2 3 0 1, edges 5 1 3 2 4 0, faces 2 3 0 1.
      face 0: local 3, 4, 5 = edges 2 (9), 4 (10), 0 (1) -> j = 4.
      face 1: local 1, 2, 5 = edges 1 (4), 3 (5), 0 (1) -> k = 2.
      case 42 (FACE): permute_vertices(2, 3): verts 2 3 1 0, edges
        5 3 1 4 2 0, faces 2 3 1 0."""
    p = tmp_path / "tet.dat"
    p.write_text(_TET)
    g = pa.Grid.read_albert(str(p))
    np.testing.assert_array_equal(g.elem_verts(), [[2, 3, 1, 0]])
    np.testing.assert_array_equal(g.elem_edges(), [[5, 3, 1, 4, 2, 0]])
    np.testing.assert_array_equal(g.elem_faces(), [[2, 3, 1, 0]])
    np.testing.assert_array_equal(g.dof_map(2), [[2, 3, 1, 0, 9, 7, 5, 8, 6, 4]])
    # each local edge still joins the vertices it names
    v = g.elem_verts()[0]
    ev = {0: (0, 1), 1: (0, 2), 2: (0, 3), 3: (1, 2), 4: (1, 3), 5: (2, 3)}
    for k, (a, b) in enumerate([(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]):
        assert set(ev[g.elem_edges()[0][k]]) == {v[a], v[b]}
    # with an element type section (albert.c:195 read_types; the reader
    # returns has_element_type, albert.c:461, so phgRefineInit skips marking)
    # the file order stands
    p2 = tmp_path / "tet_typed.dat"
    p2.write_text(_TET + "\nelement type:\n 0\n")
    g2 = pa.Grid.read_albert(str(p2))
    np.testing.assert_array_equal(g2.elem_verts(), [[0, 1, 2, 3]])
    np.testing.assert_array_equal(g2.elem_edges(), [[0, 1, 2, 3, 4, 5]])
    np.testing.assert_array_equal(g2.elem_faces(), [[0, 1, 2, 3]])
