"""BiCGstab (solver-phg.c:607 solver_BiCGstab, Jacobi PC) on the GPU against
the oracle's restatement, known answers and direct solves."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as orc
from tests._cases import poisson_case
from tests.test_oracle import _solver_test0_system

pytestmark = pytest.mark.gpu


def _fe_system(pa, n, p):
    c = poisson_case(n, p, "kappa")
    pb = pa.Problem.poisson(use_kappa=True)
    sp = pa.Space(c["grid"], p, bdry=c["bdry"])
    A = pa.assemble(sp, pb, kappa=c["kappa"], f=c["f"], g=c["g"])
    A.eliminate()
    Ao = orc.assemble(orc.problem(p, use_kappa=True), c["verts"], c["elem_verts"], c["map"],
                      c["ndof"], c["bdry"], c["g"], c["f"], c["kappa"])
    orc.eliminate(Ao, c["bdry"])
    return A, Ao


@pytest.mark.parametrize("fmt", list(range(10)))
def test_bicgstab_matches_oracle_every_layout(hip, fmt):
    pa = hip
    A, Ao = _fe_system(pa, 4, 2)
    A.set_format(fmt, fmt)
    xo, nits_o, res_o = orc.bicgstab(Ao, Ao.rhs, rtol=1e-10, btol=1e-10)
    x = pa.Vec(A.n)
    nits, res = pa.bicgstab(A, A.rhs(), x, rtol=1e-10, btol=1e-10)
    # same recurrence; FMA contraction on the GPU moves the last bits
    assert abs(nits - nits_o) <= 2
    assert np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo) < 1e-8
    assert res <= 1e-10 * np.linalg.norm(Ao.rhs) * (1 + 1e-9)


def test_bicgstab_nonzero_guess_matches_oracle(hip):
    pa = hip
    A, Ao = _fe_system(pa, 3, 3)
    x0 = np.random.default_rng(2).standard_normal(A.n)
    xo, nits_o, _ = orc.bicgstab(Ao, Ao.rhs, x0=x0, rtol=1e-11, btol=1e-11)
    x = pa.Vec(A.n, x0)
    nits, res = pa.bicgstab(A, A.rhs(), x, rtol=1e-11, btol=1e-11)
    assert abs(nits - nits_o) <= 2
    assert np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo) < 1e-8


def test_bicgstab_solver_test0_and_nonsymmetric(hip):
    import scipy.sparse as sp
    pa = hip
    rp, ci, va, b = _solver_test0_system()
    A = pa.mat_from_csr(rp, ci, va, b, np.zeros(len(b), np.uint8), True)
    x = pa.Vec(len(b))
    nits, res = pa.bicgstab(A, pa.Vec(len(b), b), x, rtol=1e-14, btol=1e-14)
    _, nits_o, _ = orc.bicgstab(orc.csr_from_arrays(rp, ci, va), b, rtol=1e-14, btol=1e-14)
    # at rtol 1e-14 the last step's residual is at rounding level, so the
    # step count may differ by one
    assert abs(nits - nits_o) <= 1
    np.testing.assert_allclose(x.numpy(), 1.0, rtol=0, atol=1e-12)
    rng = np.random.default_rng(3)
    n = 300
    M = sp.random(n, n, density=0.03, random_state=4, format="csr") * 2.0 - \
        sp.random(n, n, density=0.03, random_state=5, format="csr")
    M = (M + sp.diags(np.abs(M).sum(axis=1).A.ravel() + 1.0)).tocsr()
    M.sort_indices()
    xs = rng.standard_normal(n)
    bv = M @ xs
    A = pa.mat_from_csr(M.indptr, M.indices, M.data, bv, np.zeros(n, np.uint8), True)
    x = pa.Vec(n)
    nits, res = pa.bicgstab(A, pa.Vec(n, bv), x, rtol=1e-13, btol=1e-13)
    assert np.linalg.norm(x.numpy() - xs) / np.linalg.norm(xs) < 1e-10
    assert np.linalg.norm(bv - M @ x.numpy()) < 10 * max(res, 1e-13 * np.linalg.norm(bv))


def test_bicgstab_small_s_exit_and_zero_rhs(hip):
    """A diagonal system takes the |s| <= 1e-60 exit in the first step with
    x = b / diag and the true residual (0); b = 0 returns x = 0."""
    pa = hip
    n = 9
    d = np.arange(1.0, n + 1.0)
    b = np.linspace(-2.0, 3.0, n)
    A = pa.mat_from_csr(np.arange(n + 1), np.arange(n), d, b, np.zeros(n, np.uint8), True)
    x = pa.Vec(n)
    nits, res = pa.bicgstab(A, pa.Vec(n, b), x)
    assert nits == 1 and res == 0.0
    assert np.array_equal(x.numpy(), b / d)
    x = pa.Vec(n, np.ones(n))
    nits, res = pa.bicgstab(A, pa.Vec(n), x)
    assert nits == 0 and res == 1.0 and not x.numpy().any()


def test_bicgstab_maxit(hip):
    pa = hip
    A, Ao = _fe_system(pa, 3, 2)
    x = pa.Vec(A.n)
    nits, res = pa.bicgstab(A, A.rhs(), x, maxit=3, rtol=1e-30, btol=1e-30)
    _, nits_o, res_o = orc.bicgstab(Ao, Ao.rhs, maxit=3, rtol=1e-30, btol=1e-30)
    assert nits == nits_o == 3
    assert abs(res - res_o) <= 1e-8 * res_o
