"""Objective / adjoint helpers shared by the adjoint tests (host side).

J(kappa) = 1/2 int (u_h - ud_h)^2 for -div(kappa grad u) = 1, u = 0 on the
boundary; ud is the solution for a second seeded kappa field.
"""
from __future__ import annotations

import numpy as np

import oracle as orc

from tests._cases import lognormal_kappa, poisson_case


def c3_case(n, p=1, sigma=0.3, seed=31):
    """C3 family: log kappa ~ N(0, sigma^2), target from a second field."""
    c = poisson_case(n, p, "kappa", sigma=sigma, seed=seed)
    c["kappa_d"] = lognormal_kappa(c["grid"].nelem, sigma, seed + 1)
    return c


def oracle_solve(c, kappa, rtol=1e-15):
    pb = orc.problem(c["p"], use_kappa=True)
    A = orc.assemble(pb, c["verts"], c["elem_verts"], c["map"], c["ndof"], c["bdry"],
                     c["g"], c["f"], kappa)
    orc.eliminate(A, c["bdry"])
    x, _, _ = orc.pcg(A, A.rhs, rtol=rtol, btol=rtol, maxit=100000)
    return A, x


def oracle_objective(c, u, ud):
    J, _ = orc.mass_apply(c["p"], c["verts"], c["elem_verts"], c["map"], u - ud)
    return J


def oracle_adjoint(c, kappa, rtol=1e-15):
    """CPU adjoint: (J, grad, u, ud, lambda)."""
    A, u = oracle_solve(c, kappa, rtol)
    _, ud = oracle_solve(c, c["kappa_d"], rtol)
    J, Me = orc.mass_apply(c["p"], c["verts"], c["elem_verts"], c["map"], u - ud)
    rhs = np.where(c["bdry"] != 0, 0.0, Me)
    lam, _, _ = orc.pcg(A, rhs, rtol=rtol, btol=rtol, maxit=100000)
    g = orc.adjoint_grad(c["p"], c["verts"], c["elem_verts"], c["map"], kappa, u, lam)
    return J, g, u, ud, lam
