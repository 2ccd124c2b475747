"""Host local refinement (newest-vertex bisection with conforming closure)
used to build the unevenly refined meshes of the batched-solve workload."""
from __future__ import annotations

from collections import Counter

import numpy as np
import pytest

import phg_amd as pa


def _faces(ev):
    f = np.concatenate([ev[:, [1, 2, 3]], ev[:, [0, 2, 3]], ev[:, [0, 1, 3]],
                        ev[:, [0, 1, 2]]])
    return np.sort(f, axis=1)


def _volumes(xyz, ev):
    a = xyz[ev[:, 1]] - xyz[ev[:, 0]]
    b = xyz[ev[:, 2]] - xyz[ev[:, 0]]
    c = xyz[ev[:, 3]] - xyz[ev[:, 0]]
    return np.abs(np.einsum("ij,ij->i", a, np.cross(b, c))) / 6.0


@pytest.mark.parametrize("levels,point", [(2, (0, 0, 0)), (5, (1, 0, 1)), (7, (1, 1, 1))])
def test_refined_mesh_is_conforming(levels, point):
    g = pa.Grid.kuhn(4).refine_toward(point, levels, radius=0.45)
    xyz, ev = g.verts(), g.elem_verts()
    # every face is shared by two elements unless it lies on the cube boundary
    cnt = Counter(map(tuple, _faces(ev)))
    for f, k in cnt.items():
        assert k in (1, 2)
        if k == 1:
            x = xyz[list(f)]
            assert any(np.all(x[:, d] == x[0, d]) and x[0, d] in (0.0, 1.0)
                       for d in range(3)), f
    vol = _volumes(xyz, ev)
    assert np.all(vol > 0) and abs(vol.sum() - 1.0) < 1e-12
    # no vertex sits inside an edge of a coarser element (hanging nodes)
    assert g.nvert == len(np.unique(np.round(xyz * 2 ** 20).astype(np.int64), axis=0))


def test_refinement_is_local_and_deterministic():
    g = pa.Grid.kuhn(4)
    r1 = g.refine_toward((0, 0, 0), 6, radius=0.3)
    r2 = g.refine_toward((0, 0, 0), 6, radius=0.3)
    assert np.array_equal(r1.elem_verts(), r2.elem_verts())
    xyz, ev = r1.verts(), r1.elem_verts()
    vol = _volumes(xyz, ev)
    c = xyz[ev].mean(axis=1)
    near = np.linalg.norm(c, axis=1) < 0.2
    far = np.linalg.norm(c, axis=1) > 0.9
    assert vol[near].max() < vol[far].min()
    # each sweep bisects the marked elements at least once
    assert r1.nelem > g.nelem * 1.5


def test_refined_p3_space_numbering():
    """The refined mesh gets PHG's DOF numbering layer: P3 DOF count =
    vertices + 2 edges + faces, and every element map entry in range."""
    r = pa.Grid.kuhn(3).refine_toward((1, 1, 0), 4, radius=0.4)
    nd = r.ndof(3)
    m = r.dof_map(3)
    assert m.shape == (r.nelem, 20) and m.min() == 0 and m.max() == nd - 1
    ev = r.elem_verts()
    edges = {tuple(sorted(e)) for t in ev for e in
             ((t[0], t[1]), (t[0], t[2]), (t[0], t[3]), (t[1], t[2]), (t[1], t[3]), (t[2], t[3]))}
    faces = set(map(tuple, _faces(ev)))
    assert nd == r.nvert + 2 * len(edges) + len(faces)
