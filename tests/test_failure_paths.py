"""Failure and edge paths of the Krylov solvers, forced from the host with
inputs that end every kernel normally, against the oracle's restatement of
solver-phg.c:485-604:
  zero diagonal   jacobi_pc stops (solver-phg.c:233): PHG_HIP_E_ZERO_DIAG
                  from PCG, BiCGstab and GMRES, oracle PCG -1;
  (p, Ap) = 0     the loop ends after the iteration with a warning, not an
                  error (solver-phg.c:563-570): same iterations and
                  residual as the oracle;
  NaN in b        |r| > tol is false for NaN, so no iteration runs and the
                  residual is NaN (solver-phg.c:535);
  maxit = 1       one iteration, the residual after it;
  maxit = 0       no iteration, the initial residual."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as orc
from tests._cases import poisson_case

pytestmark = pytest.mark.gpu


def _csr(pa, rp, ci, va, b):
    n = len(rp) - 1
    A = pa.mat_from_csr(np.asarray(rp, dtype=np.int64), np.asarray(ci, dtype=np.int32),
                        np.asarray(va, dtype=np.float64), np.asarray(b, dtype=np.float64),
                        np.zeros(n, dtype=np.uint8), True)
    return A, orc.csr_from_arrays(rp, ci, va)


def test_zero_diagonal(hip):
    pa = hip
    rp, ci, va = [0, 2, 4, 6], [0, 1, 1, 0, 2, 1], [4.0, 1.0, 0.0, 1.0, 4.0, 1.0]
    b = np.ones(3)
    A, Ao = _csr(pa, rp, ci, va, b)
    for solve in (lambda x: pa.pcg(A, A.rhs(), x),
                  lambda x: pa.bicgstab(A, A.rhs(), x),
                  lambda x: pa.gmres(A, A.rhs(), x)):
        with pytest.raises(pa.PhgHipError) as ei:
            solve(pa.Vec(3))
        assert ei.value.code == -4
    _, nits, _ = orc.pcg(Ao, b)
    assert nits == -1
    x = pa.Vec(3)
    with pytest.raises(pa.PhgHipError) as ei:
        pa.pc_jacobi(A, A.rhs(), x)
    assert ei.value.code == -4


def test_pap_zero_breakdown(hip):
    """diag(1, -1), b = (1, 1): p = z = (1, -1), A p = (1, 1), (p, Ap) = 0."""
    pa = hip
    rp, ci, va = [0, 1, 2], [0, 1], [1.0, -1.0]
    b = np.array([1.0, 1.0])
    A, Ao = _csr(pa, rp, ci, va, b)
    x = pa.Vec(2)
    nits, res = pa.pcg(A, A.rhs(), x)
    xo, nits_o, res_o = orc.pcg(Ao, b)
    assert (nits, nits_o) == (1, 1)
    assert res == res_o == np.sqrt(2.0)
    assert not x.numpy().any() and not xo.any()


def test_nan_rhs_runs_no_iteration(hip):
    pa = hip
    c = poisson_case(3, 2, "kappa")
    sp = pa.Space(c["grid"], 2, bdry=c["bdry"])
    A = pa.assemble(sp, pa.Problem.poisson(use_kappa=True), kappa=c["kappa"], f=c["f"],
                    g=c["g"])
    A.eliminate()
    bv = A.rhs().numpy()
    bv[5] = np.nan
    x = pa.Vec(A.n)
    nits, res = pa.pcg(A, pa.Vec(A.n, bv), x)
    rp, ci, va, _ = A.download()
    _, nits_o, res_o = orc.pcg(orc.csr_from_arrays(rp, ci, va), bv)
    assert nits == nits_o == 0
    assert np.isnan(res) and np.isnan(res_o)


@pytest.mark.parametrize("maxit", [0, 1, 2])
def test_small_maxit(hip, maxit):
    pa = hip
    c = poisson_case(4, 2, "kappa")
    sp = pa.Space(c["grid"], 2, bdry=c["bdry"])
    A = pa.assemble(sp, pa.Problem.poisson(use_kappa=True), kappa=c["kappa"], f=c["f"],
                    g=c["g"])
    A.eliminate()
    rp, ci, va, rhs = A.download()
    x = pa.Vec(A.n)
    nits, res = pa.pcg(A, A.rhs(), x, maxit=maxit)
    xo, nits_o, res_o = orc.pcg(orc.csr_from_arrays(rp, ci, va), rhs, maxit=maxit)
    assert nits == nits_o == maxit
    assert abs(res - res_o) <= 1e-12 * res_o
    assert np.linalg.norm(x.numpy() - xo) <= 1e-12 * max(np.linalg.norm(xo), 1.0)
