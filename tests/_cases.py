"""Seeded problem setups shared by the tests (host data only)."""
from __future__ import annotations

import numpy as np

import phg_amd as pa

POISSON_ALL = (1, 1, 1, 1, 1, 1)          # Dirichlet on every side
CLAMP_X0 = (1, -1, -1, -1, -1, -1)        # elasticity: x=0 clamped


def u_sin(xyz):
    """C1 manufactured solution u = sin(pi x) sin(pi y) sin(pi z)."""
    return np.sin(np.pi * xyz[:, 0]) * np.sin(np.pi * xyz[:, 1]) * np.sin(np.pi * xyz[:, 2])


def f_sin(xyz):
    return 3.0 * np.pi ** 2 * u_sin(xyz)


def lognormal_kappa(nelem, sigma, seed):
    rng = np.random.default_rng(seed)
    return np.exp(rng.normal(0.0, sigma, size=nelem))


def poisson_case(n, p, kind="sin", sigma=0.5, seed=7, mesh="kuhn"):
    """Returns dict(grid, map, ndof, bdry, coords, f, g, kappa) for an n^3 x 6
    mesh: the structured generator's (mesh="kuhn") or PHG's (mesh="phg":
    test/cube.dat bisected 3 log2(n) times, the workloads' mesh).

    kind: 'sin'   -Lap u = f with u = sin sin sin (C1 family),
          'kappa' -div(kappa grad u) = 1, u = 0 (C2 family),
          'linear' u = x+y+z, f = 0 (examples/functions.h TEST_CASE 1).
    """
    g = pa.Grid.phg_cube(n, POISSON_ALL) if mesh == "phg" else pa.Grid.kuhn(n, POISSON_ALL)
    xyz = g.dof_coords(p)
    nd = g.ndof(p)
    case = dict(grid=g, p=p, map=g.dof_map(p), ndof=nd, bdry=g.dof_bdry(p),
                coords=xyz, verts=g.verts(), elem_verts=g.elem_verts(), kappa=None)
    if kind == "sin":
        case["f"] = f_sin(xyz)
        case["g"] = u_sin(xyz)
    elif kind == "kappa":
        case["f"] = np.ones(nd)
        case["g"] = np.zeros(nd)
        case["kappa"] = lognormal_kappa(g.nelem, sigma, seed)
    elif kind == "linear":
        case["f"] = np.zeros(nd)
        case["g"] = xyz.sum(axis=1)
    else:
        raise ValueError(kind)
    return case


def elasticity_case(n, p=2, seed=11):
    g = pa.Grid.kuhn(n, CLAMP_X0)
    nd = g.ndof(p)
    rng = np.random.default_rng(seed)
    return dict(grid=g, p=p, map=g.dof_map(p), ndof=nd, bdry=g.dof_bdry(p),
                coords=g.dof_coords(p), verts=g.verts(), elem_verts=g.elem_verts(),
                f=rng.uniform(-1.0, 1.0, size=3 * nd), g=np.zeros(3 * nd), kappa=None)
