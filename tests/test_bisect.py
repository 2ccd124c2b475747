"""PHG's bisection refinement (host/bisect.c, refine.c restated).  CPU only.

- tests/golden/cube4.dat is the reference's own test/cube4.dat: cube.dat
  refined six times by one level and exported (albert.c:537: vertices in
  creation order, elements in tree traversal order with their local vertex
  order, boundary codes and bisection types).  Six one-level refinements of
  tests/golden/cube.dat must reproduce it exactly.
- Element, edge and face indices are not in that file; they are checked
  by invariants: bijections onto 0..n-1, one index per geometric edge and
  face, untouched entities keep their indices (refine.c:2176
  update_indices), the element with an old index lies inside the old
  element of that index, new edges and faces numbered in traversal order.
- The DOF layout on refined meshes (P1-P4) is conforming, and the oracle's
  assembly on it reproduces polynomials of the space's degree."""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle as orc
import phg_amd as pa

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EV = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
FV = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]


def _sections(path):
    out, sec = {}, None
    for line in open(path):
        s = line.strip()
        if not s:
            continue
        if ":" in s and not (s[0].isdigit() or s[0] == "-"):
            sec = s.split(":", 1)[0].strip()
            out[sec] = []
            continue
        out[sec].append([float(x) for x in s.split()])
    return {k: np.array(v) for k, v in out.items()}


def _cube():
    return pa.Grid.read_albert(os.path.join(GOLDEN, "cube.dat"))


def test_reproduces_reference_cube4(tmp_path):
    g = _cube()
    for _ in range(6):
        g = g.refine_all(1)
    path = str(tmp_path / "c4.dat")
    g.write_albert(path)
    mine, ref = _sections(path), _sections(os.path.join(GOLDEN, "cube4.dat"))
    for k in ("vertex coordinates", "element vertices", "element boundaries", "element type"):
        assert mine[k].shape == ref[k].shape, k
        assert np.array_equal(mine[k], ref[k]), k


def _local_marks(g, seed):
    rng = np.random.default_rng(seed)
    c = g.verts()[g.elem_verts()].mean(axis=1)
    p = rng.uniform(0.0, 1.0, 3)
    return (np.linalg.norm(c - p, axis=1) < 0.35).astype(np.int32)


def _grids():
    """(parent, child) pairs: uniform one level, uniform three levels in one
    call, local marks on cube.dat refined and on cube4.dat."""
    c = _cube().refine_all(3)
    c4 = pa.Grid.read_albert(os.path.join(GOLDEN, "cube4.dat"))
    return [(c, c.refine_all(1)), (c, c.refine_all(3)),
            (c, c.bisect(_local_marks(c, 1))), (c4, c4.bisect(_local_marks(c4, 2))),
            (c4, c4.bisect(2 * _local_marks(c4, 3)))]


def _edge_keys(g):
    ev = g.elem_verts()
    return np.sort(np.stack([ev[:, [a, b]] for a, b in EV], axis=1), axis=2)


def _face_keys(g):
    ev = g.elem_verts()
    return np.sort(np.stack([ev[:, list(f)] for f in FV], axis=1), axis=2)


def _consistent(idx, keys, n):
    m = {}
    for i, k in zip(idx.ravel(), map(tuple, keys.reshape(-1, keys.shape[-1]))):
        assert m.setdefault(k, i) == i
    assert sorted(set(m.values())) == list(range(n)) and len(m) == n
    return m


@pytest.mark.parametrize("case", range(5))
def test_mesh_conforming_and_volume(case):
    parent, g = _grids()[case]
    x = g.verts()[g.elem_verts()]
    vol = np.abs(np.linalg.det(x[:, 1:] - x[:, :1])) / 6
    assert vol.min() > 0 and abs(vol.sum() - 1.0) < 1e-12
    # every interior face has exactly one neighbour across it, and the
    # neighbour links agree with the vertex triples
    nb = g.elem_neigh()
    fk = _face_keys(g)
    for e in range(g.nelem):
        for f in range(4):
            o = nb[e, f]
            if o < 0:
                continue
            assert tuple(fk[e, f]) in set(map(tuple, fk[o]))
    # Euler characteristic of a ball
    assert g.nvert - g.nedge + g.nface - g.nelem == 1
    # the old vertices stay where they were
    assert np.array_equal(g.verts()[:parent.nvert], parent.verts())


@pytest.mark.parametrize("case", range(5))
def test_indices_inherited_and_consistent(case):
    parent, g = _grids()[case]
    assert np.array_equal(np.sort(g.traversal()), np.arange(g.nelem))
    em = _consistent(g.elem_edges(), _edge_keys(g), g.nedge)
    fm = _consistent(g.elem_faces(), _face_keys(g), g.nface)
    pem = _consistent(parent.elem_edges(), _edge_keys(parent), parent.nedge)
    pfm = _consistent(parent.elem_faces(), _face_keys(parent), parent.nface)
    # edges and faces that were not cut keep their numbers; the halves of
    # cut ones keep the parent's number in exactly one half
    for k, i in pem.items():
        if k in em:
            assert em[k] == i
    for k, i in pfm.items():
        if k in fm:
            assert fm[k] == i
    assert sorted(set(em.values()) & set(pem.values())) == list(range(parent.nedge))
    assert sorted(set(fm.values()) & set(pfm.values())) == list(range(parent.nface))
    # an old element index lives on in an element inside that old element
    px = parent.verts()[parent.elem_verts()]
    cx = g.verts()[g.elem_verts()].mean(axis=1)
    for i in range(parent.nelem):
        X = np.hstack([px[i], np.ones((4, 1))])
        lam = np.linalg.solve(X.T, np.append(cx[i], 1.0))
        assert lam.min() > -1e-12, i
    # new edges in traversal order of first meeting
    ev = g.elem_verts()
    seen = parent.nedge
    for e in g.traversal():
        for k, (a, b) in enumerate(EV):
            i = g.elem_edges()[e, k]
            if i >= parent.nedge and i >= seen:
                assert i == seen
                seen += 1


@pytest.mark.parametrize("case,p", [(0, 1), (2, 2), (3, 3), (4, 4)])
def test_dof_layout_on_refined_mesh(case, p):
    _, g = _grids()[case]
    m = g.dof_map(p)
    xyz = g.dof_coords(p)
    # every element's local nodes sit at the coordinates of their DOFs
    from tests.test_p4_numbering import _local_nodes
    if p == 4:
        nodes = _local_nodes()
    else:
        nodes = np.array([orc_node for orc_node in _nodes_pk(p)])
    X = g.verts()[g.elem_verts()]
    P = np.einsum("bv,evc->ebc", nodes, X)
    assert np.abs(P - xyz[m]).max() < 1e-14
    # the discrete solution reproduces a polynomial of degree p
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    u = {1: x + 2 * y - z, 2: x * x - y * z + z, 3: x ** 3 - x * y * z + y,
         4: x ** 4 - y ** 2 * z ** 2 + x * y}[p]
    f = {1: 0 * x, 2: -2 + 0 * x, 3: -6 * x, 4: -(12 * x ** 2 - 2 * z ** 2 - 2 * y ** 2)}[p]
    nd, bd = g.ndof(p), g.dof_bdry(p)
    A = orc.assemble(orc.problem(p), g.verts(), g.elem_verts(), m, nd, bd, u, f)
    orc.eliminate(A, bd)
    xs, _, _ = orc.pcg(A, A.rhs, rtol=1e-15, btol=1e-15, maxit=100000)
    assert np.abs(xs - u).max() < 1e-10


def _nodes_pk(p):
    nodes = [np.eye(4)[k] for k in range(4)]
    if p == 2:
        for a, b in EV:
            l = np.zeros(4); l[a] = l[b] = .5; nodes.append(l)
    if p == 3:
        for a, b in EV:
            for wa in (2 / 3, 1 / 3):
                l = np.zeros(4); l[a], l[b] = wa, 1 - wa; nodes.append(l)
        for f in FV:
            l = np.zeros(4); l[list(f)] = 1 / 3; nodes.append(l)
    return nodes


def test_marks_and_failures():
    g = _cube().refine_all(3)
    # marks of zero leave the mesh as it is (indices included)
    h = g.bisect(np.zeros(g.nelem, dtype=np.int32))
    assert h.nelem == g.nelem and np.array_equal(h.elem_verts(), g.elem_verts())
    assert np.array_equal(h.elem_edges(), g.elem_edges())
    # one marked element: the closure refines its neighbours around the
    # refinement edge, every element at least as often as marked
    marks = np.zeros(g.nelem, dtype=np.int32)
    marks[5] = 1
    h = g.bisect(marks)
    assert h.nelem > g.nelem + 1


def test_marking_strategies_known_answers():
    """mark.c:912 phgMarkElements on cube.dat's six elements, answers worked
    out by hand from mark.c's steps."""
    g = _cube()
    # MAX (mark.c:684): theta * max = 0.5 * 6 = 3, mark err > 3
    err = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    assert list(g.mark(pa.MARK_MAX, err, theta=0.5, depth=2)) == [0, 0, 0, 2, 2, 2]
    # EQDIST (mark.c:646): theta * tol = 0.5 * 4 = 2, mark err > 2
    assert list(g.mark(pa.MARK_EQDIST, err, theta=0.5, tol=4.0)) == [0, 0, 1, 1, 1, 1]
    assert list(g.mark(pa.MARK_ALL, depth=3)) == [3] * 6
    # GERS (mark.c:406, :190): err = [1 2 3 4 0 0], theta 0.5: eta = 0.5 * 10
    # = 5, max 4.  Factor 0.5: S(0) = 10, S(.5) = 3+4 = 7, S(1) = 0; 7 > 5 so
    # [.5, 1]; S(.75) = 4 <= 5 so [.5, .75]; S(.625) = 7 = S(low): stop at
    # .625, mark err > 2.5
    err = np.array([1.0, 2.0, 3.0, 4.0, 0.0, 0.0])
    assert list(g.mark(pa.MARK_GERS, err, theta=0.5)) == [0, 0, 1, 1, 0, 0]
    # err = [10 9 1 1 1 1], theta 0.4: eta = 9.2, max 10.  S(.5) = 19 > 9.2
    # so [.5, 1]; S(.75) = 19 = S(low): stop, mark err > 7.5 -- both large
    # elements, although 10 alone exceeds eta
    err = np.array([10.0, 9.0, 1.0, 1.0, 1.0, 1.0])
    assert list(g.mark(pa.MARK_GERS, err, theta=0.4)) == [1, 1, 0, 0, 0, 0]
    with pytest.raises(ValueError):
        g.mark(pa.MARK_MAX, None)


def test_mark_then_bisect_refines_marked_elements():
    g = _cube().refine_all(3)
    rng = np.random.default_rng(8)
    err = rng.uniform(0.0, 1.0, g.nelem)
    m = g.mark(pa.MARK_GERS, err, theta=0.3, depth=2)
    h = g.bisect(m)
    # every marked element is cut into at least 4 leaves
    px = g.verts()[g.elem_verts()]
    cx = h.verts()[h.elem_verts()].mean(axis=1)
    inside = np.zeros(g.nelem, dtype=int)
    for i in range(g.nelem):
        X = np.hstack([px[i], np.ones((4, 1))])
        lam = np.linalg.solve(X.T, np.vstack([cx.T, np.ones(h.nelem)]))
        inside[i] = (lam.min(axis=0) > -1e-12).sum()
    assert (inside[m > 0] >= 4).all() and (inside >= 1).all()
    assert inside.sum() == h.nelem
