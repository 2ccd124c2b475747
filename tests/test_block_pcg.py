"""Block of right-hand sides (phg_hip_pcg_block): each column against the
single-vector GPU PCG and the oracle's PCG on the same system."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as orc
from tests._cases import elasticity_case, poisson_case

pytestmark = pytest.mark.gpu


def _loads(pa, c, pb, k, seed):
    """k right-hand sides: body forces with i.i.d. U(-1, 1) nodal values,
    assembled and eliminated like the matrix (C4 generator)."""
    sp = pa.Space(c["grid"], c["p"], bdry=c["bdry"])
    rng = np.random.default_rng(seed)
    dim = 3 if pb.dim == 3 else 1
    A = None
    bs, fs = [], []
    for j in range(k):
        f = rng.uniform(-1.0, 1.0, size=dim * c["ndof"])
        Aj = pa.assemble(sp, pb, kappa=c["kappa"], f=f, g=c["g"])
        Aj.eliminate()
        bs.append(Aj.rhs())
        fs.append(f)
        if A is None:
            A = Aj
    return sp, A, bs, fs


def _oracle_solution(c, pb_o, f, tol=1e-14):
    dim = pb_o.dim
    bd = np.repeat(c["bdry"], dim) if dim == 3 else c["bdry"]
    Ao = orc.assemble(pb_o, c["verts"], c["elem_verts"], c["map"], c["ndof"],
                      c["bdry"], c["g"], f, c["kappa"])
    orc.eliminate(Ao, bd)
    x, _, _ = orc.pcg(Ao, Ao.rhs, rtol=tol, btol=tol, maxit=100000)
    return Ao, x


@pytest.mark.parametrize("shape", [0, 2, 3, 5])
@pytest.mark.parametrize("mixed", [False, True])
def test_block_elasticity_16_rhs(hip, mixed, shape):
    pa = hip
    c = elasticity_case(3, 2)
    c["kappa"] = None
    pb = pa.Problem.elasticity()
    sp, A, bs, fs = _loads(pa, c, pb, 16, seed=4)
    pa.spmm_set_shape(A, shape, shape)
    xs = [pa.Vec(A.n) for _ in bs]
    nits, res = pa.pcg_block(A, bs, xs, rtol=1e-10, btol=1e-10, mixed=mixed)
    pb_o = orc.problem(2, dim=3)
    for j in (0, 7, 15):
        x1 = pa.Vec(A.n)
        n1, _ = pa.pcg(A, bs[j], x1, rtol=1e-10, btol=1e-10)
        _, xo = _oracle_solution(c, pb_o, fs[j])
        xb = xs[j].numpy()
        if not mixed:
            # the same recurrence with reductions in another order
            assert abs(int(nits[j]) - n1) <= 2, (nits[j], n1)
        else:
            assert nits[j] <= 1.3 * n1 + 5, (nits[j], n1)
        assert np.linalg.norm(xb - xo) / np.linalg.norm(xo) < 1e-8


@pytest.mark.parametrize("shape", [1, 4])
@pytest.mark.parametrize("mixed", [False, True])
def test_block_scalar_5_rhs_residuals(hip, mixed, shape):
    """Scalar P2 kappa-Poisson, 5 columns (1x1 blocks): the returned
    residual is the FP64 residual of the returned x, below each column's
    tolerance."""
    pa = hip
    c = poisson_case(5, 2, "kappa")
    pb = pa.Problem.poisson(use_kappa=True)
    sp, A, bs, fs = _loads(pa, c, pb, 5, seed=8)
    pa.spmm_set_shape(A, shape, shape)
    xs = [pa.Vec(A.n) for _ in bs]
    nits, res = pa.pcg_block(A, bs, xs, rtol=1e-11, btol=1e-11, mixed=mixed)
    pb_o = orc.problem(2, use_kappa=True)
    for j in range(5):
        Ao, xo = _oracle_solution(c, pb_o, fs[j])
        xb = xs[j].numpy()
        r = Ao.rhs - orc.matvec(Ao, xb)
        tol = 1e-11 * np.linalg.norm(Ao.rhs)
        if mixed:
            assert np.linalg.norm(r) <= tol * (1 + 1e-6)
        assert np.linalg.norm(xb - xo) / np.linalg.norm(xo) < 1e-8
        assert nits[j] > 0


def test_block_zero_column_and_initial_guess(hip):
    pa = hip
    c = poisson_case(4, 2, "kappa")
    pb = pa.Problem.poisson(use_kappa=True)
    sp, A, bs, fs = _loads(pa, c, pb, 3, seed=9)
    bs[1] = pa.Vec(A.n)      # zero right-hand side: x = 0 (solver-phg.c:511-518)
    x0 = np.random.default_rng(1).standard_normal(A.n)
    xs = [pa.Vec(A.n, x0), pa.Vec(A.n, x0), pa.Vec(A.n)]
    nits, res = pa.pcg_block(A, bs, xs, rtol=1e-14, btol=0.0)
    assert nits[1] == 0 and np.all(xs[1].numpy() == 0.0)
    pb_o = orc.problem(2, use_kappa=True)
    for j in (0, 2):
        _, xo = _oracle_solution(c, pb_o, fs[j])
        assert np.linalg.norm(xs[j].numpy() - xo) / np.linalg.norm(xo) < 1e-8


def test_spmm_tuner_picks_fastest(hip):
    pa = hip
    c = elasticity_case(3, 2)
    c["kappa"] = None
    pb = pa.Problem.elasticity()
    sp, A, bs, fs = _loads(pa, c, pb, 2, seed=5)
    t = pa.spmm_tune(A, reps=5)
    assert set(t) == set(pa.SPMM_SHAPES)
    assert all(v[0] > 0 and v[1] > 0 for v in t.values())
    assert pa.spmm_bytes(A, True) < pa.spmm_bytes(A, False)
