"""The reference's own mesh fixtures (tests/golden/*.dat, copied from
test/*.dat): the ALBERT reader, the structured Kuhn generator, neighbour
computation and the element routines are checked against them."""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle as orc
import phg_amd as pa

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _section(path, name, count):
    lines = open(path).read().splitlines()
    i = next(k for k, l in enumerate(lines) if l.strip().startswith(name))
    rows = [l.split() for l in lines[i + 1:] if l.strip()][:count]
    return np.array(rows, dtype=float)


def _neigh(g):
    return np.ctypeslib.as_array(g.s.elem_neigh, (g.nelem, 4)).copy()


def test_cube_dat_is_the_kuhn_pattern():
    """test/cube.dat: 6 tetrahedra of the unit cube.  The structured
    generator's 1^3 mesh reproduces its vertices, elements, boundary faces
    and the file's neighbour table (ALBERT: neighbour opposite vertex i)."""
    path = os.path.join(GOLDEN, "cube.dat")
    g = pa.Grid.read_albert(path)
    k = pa.Grid.kuhn(1)
    assert (g.nvert, g.nelem) == (8, 6)
    xyz_file = _section(path, "vertex coordinates", 8)
    ev_file = _section(path, "element vertices", 6).astype(np.int32)
    assert np.array_equal(g.verts(), xyz_file)
    assert np.array_equal(g.elem_verts(), ev_file)
    # same tetrahedra, same local vertex order, as the generator
    kv = k.verts()
    assert np.array_equal(kv[k.elem_verts()], xyz_file[ev_file])
    nb_file = _section(path, "element neighbours", 6).astype(np.int32)
    assert np.array_equal(_neigh(g), nb_file)
    assert np.array_equal(_neigh(k), nb_file)
    # Dirichlet faces 0 and 1 of every element ("element boundaries: 1 1 0 0")
    bd = np.ctypeslib.as_array(g.s.elem_bound, (6, 4))
    assert np.all(bd[:, :2] != 0) and np.all(bd[:, 2:] == 0)


def test_tetra_dat_reference_element():
    """test/tetra.dat, the reference tetrahedron: volume 1/6, P1 stiffness
    vol * grad(lambda_i) . grad(lambda_j) in closed form, P1 mass
    vol/20 (1 + delta_ij)."""
    g = pa.Grid.read_albert(os.path.join(GOLDEN, "tetra.dat"))
    x = g.verts()[g.elem_verts()[0]]
    assert abs(orc.volume(x) - 1.0 / 6.0) < 1e-16
    K, b = orc.element(orc.problem(1), x, f_loc=np.ones(4))
    want = np.array([[3, -1, -1, -1], [-1, 1, 0, 0], [-1, 0, 1, 0], [-1, 0, 0, 1]]) / 6.0
    assert np.allclose(K, want, rtol=0, atol=1e-16)
    M, _ = orc.element(orc.problem(1, use_kappa=True, mass_coef=1.0), x, 0.0, np.ones(4))
    assert np.allclose(M, (np.ones((4, 4)) + np.eye(4)) / 120.0, rtol=0, atol=1e-17)
    assert np.allclose(b, np.full(4, 1.0 / 24.0), rtol=0, atol=1e-17)


def test_cube4_dat_mesh_and_neighbours():
    """test/cube4.dat (384 tetrahedra from refining cube.dat): volumes sum
    to 1, every face is shared or on the boundary, and the neighbour table
    computed from the vertex lists equals the file's."""
    path = os.path.join(GOLDEN, "cube4.dat")
    g = pa.Grid.read_albert(path)
    assert (g.nvert, g.nelem) == (125, 384)
    xyz, ev = g.verts(), g.elem_verts()
    vol = sum(orc.volume(xyz[e]) for e in ev)
    assert abs(vol - 1.0) < 1e-14
    nb_file = _section(path, "element neighbours", 384).astype(np.int32)
    assert np.array_equal(_neigh(g), nb_file)
    # Lagrange layouts on the imported mesh
    for p in (1, 2, 3):
        m = g.dof_map(p)
        assert m.min() == 0 and m.max() == g.ndof(p) - 1
    assert g.ndof(1) == 125 and g.ndof(2) == 729
