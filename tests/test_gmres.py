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


# ------------------------------------------------------ the other -gmres_type
# gmres_l, gmres_r, fgmres_r, lgmres_l, lgmres_r (solver-phg.c:783-2062) and
# flgmres_r through phg_hip_gmres_type, against the oracle's restatement of
# each variant.
TYPES = orc.GMRES_TYPES
LEFT = ("gmres_l", "lgmres_l")


@pytest.mark.parametrize("fmt", [0, 3, 6, 9])
@pytest.mark.parametrize("t", TYPES)
def test_gmres_type_matches_oracle(hip, t, fmt):
    pa = hip
    A, Ao = _fe_system(pa, 4, 2)
    A.set_format(fmt, fmt)
    xo, nits_o, res_o = orc.gmres(Ao, Ao.rhs, rtol=1e-10, btol=1e-10, gmres_type=t)
    x = pa.Vec(A.n)
    nits, res = pa.gmres(A, A.rhs(), x, rtol=1e-10, btol=1e-10, gmres_type=t)
    print(f"{t} fmt {fmt}: {nits} its (oracle {nits_o}) res {res:.3e} ({res_o:.3e})",
          flush=True)
    assert abs(nits - nits_o) <= 2
    assert np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo) < 1e-8
    tol = 1e-10 * np.linalg.norm(Ao.rhs)
    assert res <= tol * (1 + 1e-9)
    if t in LEFT:
        # the true residual of the returned x
        true = np.linalg.norm(Ao.rhs - Ao.to_scipy() @ x.numpy())
        assert abs(res - true) <= 1e-6 * true


@pytest.mark.parametrize("t", TYPES)
@pytest.mark.parametrize("restart,augk", [(3, 2), (5, 1)])
def test_gmres_type_restarts_nonsymmetric_and_guess(hip, t, restart, augk):
    pa = hip
    M = _nonsym(200, 7)
    xs = np.random.default_rng(8).standard_normal(200)
    b = M @ xs
    A = pa.mat_from_csr(M.indptr, M.indices, M.data, b, np.zeros(200, np.uint8), True)
    Ao = orc.csr_from_arrays(M.indptr, M.indices, M.data)
    x0 = np.random.default_rng(9).standard_normal(200)
    kw = dict(restart=restart, augk=augk, rtol=1e-12, btol=1e-12, gmres_type=t)
    xo, nits_o, _ = orc.gmres(Ao, b, x0=x0, **kw)
    x = pa.Vec(200, x0)
    nits, res = pa.gmres(A, pa.Vec(200, b), x, **kw)
    assert abs(nits - nits_o) <= 2
    assert np.linalg.norm(x.numpy() - xs) / np.linalg.norm(xs) < 1e-9
    assert np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo) < 1e-9
    # capped runs follow the oracle step for step; 4 restart + 4 steps end
    # the fourth cycle of an augmented variant at step restart + 1, where
    # the update reads the second augmentation coefficient of cycle three
    for maxit in (restart, 2 * restart + 1, 4 * restart + 4):
        xo, nits_o, res_o = orc.gmres(Ao, b, x0=x0, maxit=maxit, **kw)
        x = pa.Vec(200, x0)
        nits, res = pa.gmres(A, pa.Vec(200, b), x, maxit=maxit, **kw)
        assert nits == nits_o
        assert np.linalg.norm(x.numpy() - xo) <= 1e-10 * np.linalg.norm(xo)
        assert abs(res - res_o) <= 1e-10 * res_o


@pytest.mark.parametrize("t", LEFT)
def test_gmres_left_false_convergence(hip, t):
    """Rows scaled by ~1e3: the preconditioned estimate meets the tolerance
    of the unpreconditioned |r0| early; the left variants check |b - A x| and
    run on (solver-phg.c:979-1003) to the true tolerance, as the oracle."""
    import scipy.sparse as sp
    pa = hip
    n = 80
    Md = _nonsym(n, 31).toarray() * (1e3 * (1.0 + np.arange(n) % 3))[:, None]
    S = sp.csr_matrix(Md)
    S.sort_indices()
    b = Md @ np.random.default_rng(32).standard_normal(n)
    A = pa.mat_from_csr(S.indptr, S.indices, S.data, b, np.zeros(n, np.uint8), True)
    Ao = orc.csr_from_arrays(S.indptr, S.indices, S.data)
    kw = dict(restart=200, augk=2, rtol=1e-6, btol=1e-30, gmres_type=t)
    xo, nits_o, res_o = orc.gmres(Ao, b, **kw)
    x = pa.Vec(n)
    nits, res = pa.gmres(A, pa.Vec(n, b), x, **kw)
    print(f"{t}: {nits} its (oracle {nits_o}) res {res:.3e} ({res_o:.3e})", flush=True)
    assert nits == nits_o
    assert np.linalg.norm(x.numpy() - xo) <= 1e-9 * np.linalg.norm(xo)
    assert res <= 1e-6 * np.linalg.norm(b) and abs(res - res_o) <= 1e-6 * res_o


@pytest.mark.parametrize("t", TYPES)
def test_gmres_type_zero_diagonal_and_zero_rhs(hip, t):
    """PHG_HIP_E_ZERO_DIAG wherever the oracle's jacobi_pc stops (-1), no
    error where the flexible variants never reach the PC; b = 0 gives x = 0
    and residual 1."""
    pa = hip
    rp, ci, va, b = _solver_test0_system()
    nb = len(b)
    A = pa.mat_from_csr(rp, ci, va, b, np.zeros(nb, np.uint8), True)
    x = pa.Vec(nb, np.ones(nb))
    nits, res = pa.gmres(A, pa.Vec(nb), x, gmres_type=t)
    assert nits == 0 and res == 1.0 and not x.numpy().any()
    va2 = np.array(va, dtype=np.float64)
    for jj in range(rp[5], rp[6]):
        if ci[jj] == 5:
            va2[jj] = 0.0
    Ao2 = orc.csr_from_arrays(rp, ci, va2)
    xs = np.ones(nb)
    b3 = Ao2.to_scipy() @ xs
    for rhs, x0 in ((b, None), (b3, xs)):
        A2 = pa.mat_from_csr(rp, ci, va2, rhs, np.zeros(nb, np.uint8), True)
        _, nits_o, _ = orc.gmres(Ao2, rhs, x0=x0, gmres_type=t)
        x = pa.Vec(nb, x0) if x0 is not None else pa.Vec(nb)
        if nits_o == -1:
            with pytest.raises(pa.PhgHipError) as ei:
                pa.gmres(A2, pa.Vec(nb, rhs), x, gmres_type=t)
            assert ei.value.code == -4
        else:
            nits, _ = pa.gmres(A2, pa.Vec(nb, rhs), x, gmres_type=t)
            assert nits == nits_o == 0
    with pytest.raises(ValueError):
        pa.gmres(A, pa.Vec(nb, b), pa.Vec(nb), gmres_type="gmres")
