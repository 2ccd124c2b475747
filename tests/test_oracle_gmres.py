"""The oracle's GMRES family (solver-phg.c:59-77: gmres_l, gmres_r, fgmres_r,
lgmres_l, lgmres_r, flgmres_r, all with the Jacobi PC) pinned against known
answers (test/solver_test0.c), direct solves, the minimal-residual property
of one cycle and of an augmented second cycle (closed forms by dense least
squares), and the equivalences a fixed preconditioner implies."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as orc
from tests._cases import poisson_case
from tests.test_oracle import _nonsym, _solver_test0_system

TYPES = orc.GMRES_TYPES
LEFT = ("gmres_l", "lgmres_l")
AUG = ("lgmres_l", "lgmres_r", "flgmres_r")


def _fe():
    c = poisson_case(3, 2, "kappa")
    Ao = orc.assemble(orc.problem(2, use_kappa=True), c["verts"], c["elem_verts"], c["map"],
                      c["ndof"], c["bdry"], c["g"], c["f"], c["kappa"])
    orc.eliminate(Ao, c["bdry"])
    return Ao


@pytest.mark.parametrize("t", TYPES)
def test_solver_test0_known_answer(t):
    rp, ci, va, b = _solver_test0_system()
    A = orc.csr_from_arrays(rp, ci, va)
    x, nits, res = orc.gmres(A, b, rtol=1e-14, btol=1e-14, gmres_type=t)
    assert nits > 0 and res <= 1e-14 * np.linalg.norm(b) * (1 + 1e-12)
    np.testing.assert_allclose(x, 1.0, rtol=0, atol=1e-11)


@pytest.mark.parametrize("t", TYPES)
@pytest.mark.parametrize("restart,augk", [(10, 2), (3, 2), (5, 1)])
def test_direct_solves(t, restart, augk):
    import scipy.sparse.linalg as spla
    Ao = _fe()
    xd = spla.spsolve(Ao.to_scipy().tocsc(), Ao.rhs)
    x, nits, res = orc.gmres(Ao, Ao.rhs, restart=restart, augk=augk, rtol=1e-12, btol=1e-12,
                             gmres_type=t)
    assert nits > 0
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9
    M = _nonsym(200, 7)
    xs = np.random.default_rng(8).standard_normal(200)
    b = M @ xs
    A = orc.csr_from_arrays(M.indptr, M.indices, M.data)
    x0 = np.random.default_rng(9).standard_normal(200)
    x, nits, res = orc.gmres(A, b, x0=x0, restart=restart, augk=augk, rtol=1e-12,
                             btol=1e-12, gmres_type=t)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-9
    true = np.linalg.norm(b - M @ x)
    tol = 1e-12 * max(np.linalg.norm(b), np.linalg.norm(b - M @ x0))
    if t in LEFT:
        # the reported residual is the true one, checked against the tolerance
        assert abs(res - true) <= 1e-9 * true and res <= tol
    else:
        # the Givens estimate; the true residual agrees
        assert res <= tol * (1 + 1e-9) and true < 100 * res


def _krylov_basis(Op, r, m):
    """Orthonormal basis of K_m(Op, r) by Arnoldi with two passes of
    classical Gram-Schmidt (numerically orthogonal for the m used here)."""
    Q = np.zeros((len(r), m))
    v = r / np.linalg.norm(r)
    for j in range(m):
        Q[:, j] = v
        w = Op @ v
        for _ in range(2):
            w -= Q[:, :j + 1] @ (Q[:, :j + 1].T @ w)
        v = w / np.linalg.norm(w)
    return Q


def _min_res(Op, rhs, Q):
    u, *_ = np.linalg.lstsq(Op @ Q, rhs, rcond=None)
    return np.linalg.norm(rhs - Op @ Q @ u), Q @ u


@pytest.mark.parametrize("t", TYPES)
def test_first_cycle_is_minimal_residual(t):
    """One cycle of m steps: right variants give x = M^-1 V y minimising
    |b - A x| over M^-1 K_m(A M^-1, b); left ones minimise |M^-1 (b - A x)|
    over x in K_m(M^-1 A, M^-1 b)."""
    n, m = 60, 6
    M = _nonsym(n, 11)
    A = M.toarray()
    d = np.diag(A)
    b = np.random.default_rng(12).standard_normal(n)
    Ao = orc.csr_from_arrays(M.indptr, M.indices, M.data)
    x, nits, res = orc.gmres(Ao, b, restart=m, augk=2, maxit=m, rtol=1e-30, btol=1e-30,
                             gmres_type=t)
    assert nits == m
    assert res == np.linalg.norm(b)          # no tolerance check: |r0| is reported
    if t in LEFT:
        MA = A / d[:, None]
        Q = _krylov_basis(MA, b / d, m)
        u, *_ = np.linalg.lstsq(MA @ Q, b / d, rcond=None)
        best = np.linalg.norm(b / d - MA @ Q @ u)
        got = np.linalg.norm((b - A @ x) / d)
    else:
        AM = A / d[None, :]
        Q = _krylov_basis(AM, b, m)
        u, *_ = np.linalg.lstsq(AM @ Q, b, rcond=None)
        best = np.linalg.norm(b - AM @ Q @ u)
        got = np.linalg.norm(b - A @ x)
    assert abs(got - best) <= 1e-9 * np.linalg.norm(b)


@pytest.mark.parametrize("t", AUG)
def test_second_cycle_is_augmented_minimal_residual(t):
    """Cycle 1 of an augmented variant has restart + 1 steps: restart
    Krylov directions from r1 and the correction of cycle 0 (the correction
    itself for the left variant, its preimage M (x1 - x0) for the right
    ones).  Its x minimises the (left: preconditioned) residual over that
    space; dropping the augmentation direction gives a strictly larger
    residual."""
    n, mk = 80, 4
    M = _nonsym(n, 21)
    A = M.toarray()
    d = np.diag(A)
    b = np.random.default_rng(22).standard_normal(n)
    Ao = orc.csr_from_arrays(M.indptr, M.indices, M.data)
    kw = dict(restart=mk, augk=2, rtol=1e-30, btol=1e-30, gmres_type=t)
    x1, n1, _ = orc.gmres(Ao, b, maxit=mk, **kw)
    x2, n2, _ = orc.gmres(Ao, b, maxit=2 * mk + 1, **kw)
    assert (n1, n2) == (mk, 2 * mk + 1)
    r1 = b - A @ x1
    if t in LEFT:
        Op, rhs, aug = A / d[:, None], r1 / d, x1
    else:
        Op, rhs, aug = A / d[None, :], r1, d * x1
    Q = _krylov_basis(Op, rhs, mk)
    S = np.column_stack([Q, aug])
    u, *_ = np.linalg.lstsq(Op @ S, rhs, rcond=None)
    best = np.linalg.norm(rhs - Op @ S @ u)
    uq, *_ = np.linalg.lstsq(Op @ Q, rhs, rcond=None)
    plain = np.linalg.norm(rhs - Op @ Q @ uq)
    got = np.linalg.norm((b - A @ x2) / d) if t in LEFT else np.linalg.norm(b - A @ x2)
    assert abs(got - best) <= 1e-9 * np.linalg.norm(b)
    assert plain > best * (1 + 1e-6)


@pytest.mark.parametrize("a,b_", [("gmres_r", "fgmres_r"), ("lgmres_r", "flgmres_r")])
def test_flexible_equals_plain_for_a_fixed_pc(a, b_):
    """With the same (Jacobi) PC in every step the flexible and plain right
    variants build the same iterates; only the rounding of M^-1 (V y) versus
    sum y_j M^-1 v_j differs."""
    M = _nonsym(200, 7)
    b = M @ np.random.default_rng(8).standard_normal(200)
    A = orc.csr_from_arrays(M.indptr, M.indices, M.data)
    for maxit in (7, 25, 10000):
        xa, na, ra = orc.gmres(A, b, restart=5, augk=2, maxit=maxit, rtol=1e-12, btol=1e-12,
                               gmres_type=a)
        xb, nb, rb = orc.gmres(A, b, restart=5, augk=2, maxit=maxit, rtol=1e-12, btol=1e-12,
                               gmres_type=b_)
        assert na == nb
        assert np.linalg.norm(xa - xb) <= 1e-10 * np.linalg.norm(xb)


@pytest.mark.parametrize("t", LEFT)
def test_left_false_convergence_closed_form(t):
    """Rows scaled by ~1e3: Jacobi divides the scaling out of M^-1 A, but the
    tolerance comes from the unpreconditioned |r0|, so the Givens estimate
    |M^-1 r| meets it at a step k where |b - A x| is still far above it.
    The left variants then check the true residual, restart from z_c =
    M^-1 r_c with s0 = |r_c| / estimate (solver-phg.c:995-1003) and continue
    that cycle until estimate * s0 meets the tolerance (solver-phg.c:904-908).
    k, the continued cycle's length j and its x follow by dense least
    squares over Krylov bases; restarting after each step, or stopping on
    the plain estimate, gives another x."""
    n = 80
    M = _nonsym(n, 31).toarray()
    sc = 1e3 * (1.0 + np.arange(n) % 3)
    A = M * sc[:, None]
    d = np.diag(A)
    b = A @ np.random.default_rng(32).standard_normal(n)
    rows, cols = np.nonzero(A)
    import scipy.sparse as sp
    S = sp.csr_matrix(A)
    S.sort_indices()
    Ao = orc.csr_from_arrays(S.indptr, S.indices, S.data)
    rtol = 1e-6
    tol = rtol * np.linalg.norm(b)
    kw = dict(restart=200, augk=2, rtol=rtol, btol=1e-30, gmres_type=t)
    MA = A / d[:, None]
    z0 = b / d
    for k in range(1, n):
        est_k, _ = _min_res(MA, z0, _krylov_basis(MA, z0, k))
        if est_k <= tol:
            break
    xc, nc, _ = orc.gmres(Ao, b, maxit=k, **kw)
    assert nc == k
    rc = b - A @ xc
    assert np.linalg.norm(rc) > 10 * tol
    s0 = np.linalg.norm(rc) / est_k
    zc = rc / d
    for j in range(1, n):
        est_j, dx = _min_res(MA, zc, _krylov_basis(MA, zc, j))
        if est_j * s0 <= tol:
            break
    assert j >= 3
    xj, nj, _ = orc.gmres(Ao, b, maxit=k + j, **kw)
    assert nj == k + j
    xp = xc + dx
    assert np.linalg.norm(xj - xp) <= 1e-8 * np.linalg.norm(xp)
    x, nits, res = orc.gmres(Ao, b, **kw)
    true = np.linalg.norm(b - A @ x)
    assert nits >= k + j and res <= tol and abs(res - true) <= 1e-9 * true


@pytest.mark.parametrize("t", TYPES)
def test_zero_rhs_and_zero_diagonal(t):
    rp, ci, va, b = _solver_test0_system()
    A = orc.csr_from_arrays(rp, ci, va)
    x, nits, res = orc.gmres(A, np.zeros(len(b)), x0=np.ones(len(b)), gmres_type=t)
    assert nits == 0 and res == 1.0 and not x.any()
    # a zero diagonal: jacobi_pc (solver-phg.c:230) stops the run
    va2 = np.array(va)
    for i in range(len(b)):
        for jj in range(rp[i], rp[i + 1]):
            if ci[jj] == i and i == 5:
                va2[jj] = 0.0
    A2 = orc.csr_from_arrays(rp, ci, va2)
    _, nits, _ = orc.gmres(A2, b, gmres_type=t)
    assert nits == -1
    # converged start: the left variants precondition r0 before any check,
    # the plain right ones precondition the (empty) correction, the flexible
    # ones never reach the PC
    xs = np.ones(len(b))
    A3 = orc.csr_from_arrays(rp, ci, va2)
    b3 = A3.to_scipy() @ xs
    _, nits, _ = orc.gmres(A3, b3, x0=xs, gmres_type=t)
    assert nits == (0 if t in ("fgmres_r", "flgmres_r") else -1)
