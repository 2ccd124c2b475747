"""FLGMRES_r, PHG's default GMRES (solver-phg.c:2063, Jacobi PC), on the GPU
against the oracle's restatement, known answers and direct solves."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as orc
from tests.test_bicgstab import _fe_system
from tests.test_oracle import _nonsym, _solver_test0_system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fmt", list(range(10)))
def test_gmres_matches_oracle_every_layout(hip, fmt):
    pa = hip
    A, Ao = _fe_system(pa, 4, 2)
    A.set_format(fmt, fmt)
    xo, nits_o, res_o = orc.gmres(Ao, Ao.rhs, rtol=1e-10, btol=1e-10)
    x = pa.Vec(A.n)
    nits, res = pa.gmres(A, A.rhs(), x, rtol=1e-10, btol=1e-10)
    # same recurrence; FMA contraction on the GPU moves the last bits
    assert abs(nits - nits_o) <= 2
    assert np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo) < 1e-8
    assert res <= 1e-10 * np.linalg.norm(Ao.rhs) * (1 + 1e-9)


@pytest.mark.parametrize("restart,augk", [(10, 2), (3, 2), (5, 1)])
def test_gmres_restarts_nonsymmetric_and_guess(hip, restart, augk):
    pa = hip
    M = _nonsym(200, 7)
    xs = np.random.default_rng(8).standard_normal(200)
    b = M @ xs
    A = pa.mat_from_csr(M.indptr, M.indices, M.data, b, np.zeros(200, np.uint8), True)
    Ao = orc.csr_from_arrays(M.indptr, M.indices, M.data)
    x0 = np.random.default_rng(9).standard_normal(200)
    xo, nits_o, _ = orc.gmres(Ao, b, x0=x0, restart=restart, augk=augk, rtol=1e-12,
                              btol=1e-12)
    x = pa.Vec(200, x0)
    nits, res = pa.gmres(A, pa.Vec(200, b), x, restart=restart, augk=augk, rtol=1e-12,
                         btol=1e-12)
    assert abs(nits - nits_o) <= 2
    assert np.linalg.norm(x.numpy() - xs) / np.linalg.norm(xs) < 1e-9
    assert np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo) < 1e-9


def test_gmres_solver_test0_maxit_and_zero_rhs(hip):
    pa = hip
    rp, ci, va, b = _solver_test0_system()
    nb = len(b)
    A = pa.mat_from_csr(rp, ci, va, b, np.zeros(nb, np.uint8), True)
    x = pa.Vec(nb)
    nits, res = pa.gmres(A, pa.Vec(nb, b), x, rtol=1e-14, btol=1e-14)
    _, nits_o, _ = orc.gmres(orc.csr_from_arrays(rp, ci, va), b, rtol=1e-14, btol=1e-14)
    assert abs(nits - nits_o) <= 1
    np.testing.assert_allclose(x.numpy(), 1.0, rtol=0, atol=1e-12)
    # not converged within maxit: the reference reports |r0|
    A2, Ao = _fe_system(pa, 3, 2)
    x = pa.Vec(A2.n)
    nits, res = pa.gmres(A2, A2.rhs(), x, maxit=4, rtol=1e-30, btol=1e-30)
    xo, nits_o, res_o = orc.gmres(Ao, Ao.rhs, maxit=4, rtol=1e-30, btol=1e-30)
    assert nits == nits_o == 4
    # |r0| (x0 = 0: |b|), summed in different orders by numpy, the oracle and
    # the work-group partials
    nb0 = np.linalg.norm(Ao.rhs)
    assert abs(res_o - nb0) <= 1e-14 * nb0 and abs(res - res_o) <= 1e-14 * res_o
    assert np.linalg.norm(x.numpy() - xo) <= 1e-10 * np.linalg.norm(xo)
    x = pa.Vec(nb, np.ones(nb))
    nits, res = pa.gmres(A, pa.Vec(nb), x)
    assert nits == 0 and res == 1.0 and not x.numpy().any()
