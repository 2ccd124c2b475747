"""P4 DOF numbering (dof.c:4821 phgDofMapE2D_, lagrange.c:1571 P4_map,
dof.c:978 phgDofMapEdgeData) and the oracle's P4 path end to end.  CPU only.

- Conformity: every element's local node b sits at the physical point of
  the global DOF map[e][b] (a wrong edge reversal or face permutation puts
  two elements' different nodes on one DOF).
- PHG's slot order: edge slot s is the node at lambda = .75 - .25 s from the
  edge's smaller global vertex; face slot s the node near the face vertex of
  rank s among the face's global vertex indices; the element DOF (after the
  vertex, edge and face blocks, in element order) is the centroid.
- P4 reproduces quartics: the assembled, eliminated and solved discrete
  solution of -Lap u = f for a quartic u (f its exact Laplacian, a
  quadratic, interpolated exactly by P4) is the nodal interpolant of u."""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle as orc
import phg_amd as pa
from tests._cases import POISSON_ALL

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EV = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
FV = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]


def _grids():
    base = pa.Grid.kuhn(2, POISSON_ALL)
    return [base, base.refine_toward(np.array([0.0, 1.0, 0.0]), 2, radius=0.6),
            pa.Grid.read_albert(os.path.join(GOLDEN, "cube4.dat"))]


def _local_nodes():
    nodes = [np.eye(4)[k] for k in range(4)]
    for a, b in EV:
        for wa in (.75, .5, .25):
            l = np.zeros(4)
            l[a], l[b] = wa, 1 - wa
            nodes.append(l)
    for f in FV:
        for t in range(3):
            l = np.zeros(4)
            l[list(f)] = .25
            l[f[t]] = .5
            nodes.append(l)
    nodes.append(np.full(4, .25))
    return np.array(nodes)


@pytest.mark.parametrize("gi", range(3))
def test_p4_map_conforming(gi):
    g = _grids()[gi]
    m = g.dof_map(4).reshape(g.nelem, 35)
    nd = g.ndof(4)
    assert nd == g.nvert + 3 * g.nedge + 3 * g.nface + g.nelem
    assert np.array_equal(np.unique(m), np.arange(nd))
    xyz = g.dof_coords(4)
    X = g.verts()[g.elem_verts()]                 # [nelem][4][3]
    P = np.einsum("bv,evc->ebc", _local_nodes(), X)
    assert np.abs(P - xyz[m]).max() < 1e-14


@pytest.mark.parametrize("gi", range(3))
def test_p4_slot_order(gi):
    g = _grids()[gi]
    ev = g.elem_verts()
    m = g.dof_map(4).reshape(g.nelem, 35)
    xyz = g.dof_coords(4)
    V = g.verts()
    v0, f4 = g.nvert, g.nvert + 3 * g.nedge
    e4 = f4 + 3 * g.nface
    for e in range(g.nelem):
        v = ev[e]
        for k, (a, b) in enumerate(EV):
            lo, hi = sorted((v[a], v[b]))
            base = m[e, 4 + 3 * k] - (m[e, 4 + 3 * k] - v0) % 3
            for s in range(3):
                w = .75 - .25 * s
                assert np.allclose(xyz[base + s], w * V[lo] + (1 - w) * V[hi], atol=1e-15)
        for k, f in enumerate(FV):
            gv = sorted(v[list(f)])
            base = m[e, 22 + 3 * k] - (m[e, 22 + 3 * k] - f4) % 3
            for s in range(3):
                pt = .25 * V[gv].sum(axis=0) + .25 * V[gv[s]]
                assert np.allclose(xyz[base + s], pt, atol=1e-15)
        assert m[e, 34] == e4 + e
        assert np.allclose(xyz[e4 + e], V[v].mean(axis=0), atol=1e-15)


def test_p4_boundary_flags():
    g = _grids()[0]
    bd = g.dof_bdry(4).astype(bool)
    xyz = g.dof_coords(4)
    on = (np.isclose(xyz, 0.0) | np.isclose(xyz, 1.0)).any(axis=1)
    assert np.array_equal(bd, on)


@pytest.mark.parametrize("gi", [0, 1])
def test_p4_reproduces_a_quartic(gi):
    g = _grids()[gi]
    xyz = g.dof_coords(4)
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    u = x ** 4 + 2 * x * y ** 3 - y ** 2 * z ** 2 + z
    f = -(12 * x ** 2 + 12 * x * y - 2 * z ** 2 - 2 * y ** 2)
    nd, bd = g.ndof(4), g.dof_bdry(4)
    A = orc.assemble(orc.problem(4), g.verts(), g.elem_verts(), g.dof_map(4), nd, bd, u, f)
    orc.eliminate(A, bd)
    xs, _, _ = orc.pcg(A, A.rhs, rtol=1e-15, btol=1e-15, maxit=100000)
    assert np.abs(xs - u).max() < 1e-10
