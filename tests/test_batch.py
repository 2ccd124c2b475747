"""Assembly and solve of several systems in captured HIP graphs
(phg_hip_batch_*).  Every run assembles each matrix again on the device from
the vectors' current contents with the kernels of phg_hip_assemble +
phg_hip_mat_eliminate, device scans included, so the PHG-order CSR after a
run must be bitwise the CSR of a fresh assembly with the same coefficients
-- pattern included: PHG's exact-zero drop (matvec.c:5460) makes the
pattern depend on kappa for P2/P3, and the coefficients drawn here change it
-- and that CSR is the oracle's (the assembly parity tests).  The solutions
must match the oracle's FP64 PCG solution to 1e-8."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as orc
from phg_amd.multimesh import batch_meshes

pytestmark = pytest.mark.gpu


def _oracle_solution(g, p, kappa, f):
    nd = g.ndof(p)
    A = orc.assemble(orc.problem(p, use_kappa=True), g.verts(), g.elem_verts(),
                     g.dof_map(p), nd, g.dof_bdry(p), np.zeros(nd), f, kappa)
    orc.eliminate(A, g.dof_bdry(p))
    x, _, _ = orc.pcg(A, A.rhs, rtol=1e-14, btol=1e-14, maxit=100000)
    return x


def _fresh(pa, sp, pb, kappa, f, eop=True):
    A = pa.assemble(sp, pb, kappa=kappa, f=f, g=np.zeros(sp.ndof))
    A.eliminate()
    if eop:
        A.element_operator(sp, pb, kappa=kappa)
    return A


def _same_csr(A, A1):
    return all(np.array_equal(a.view(np.uint8), b.view(np.uint8))
               for a, b in zip(A.download(), A1.download()))


def _relerr(x, y):
    return np.linalg.norm(x - y) / np.linalg.norm(y)


@pytest.mark.parametrize("p", [2, 3])
def test_batch_assembly_and_solve(hip, p):
    pa = hip
    meshes = batch_meshes(nmesh=4, n=3, seed=7, lmin=1, lmax=3)
    rng = np.random.default_rng(11)
    pb = pa.Problem.poisson(use_kappa=True)
    systems, grids = [], []
    for g, _, _ in meshes:
        sp = pa.Space(g, p)
        nd = g.ndof(p)
        k0 = np.exp(rng.normal(0.0, 0.5, g.nelem))
        f0 = rng.uniform(0.5, 1.5, nd)
        A = _fresh(pa, sp, pb, k0, f0)
        systems.append((sp, pb, A, pa.Vec(g.nelem, k0), pa.Vec(nd, f0),
                        pa.Vec(nd, np.zeros(nd)), pa.Vec(nd, np.zeros(nd))))
        grids.append(g)
    B = pa.Batch(systems, rtol=1e-10, btol=1e-10)
    moved = 0
    for trial in range(2):
        # new coefficients and loads in the device vectors, cold start
        ks, fs, nnz0 = [], [], []
        for (sp, _, A, kv, fv, _, xv), g in zip(systems, grids):
            k = np.exp(rng.normal(0.0, 0.5, g.nelem))
            f = rng.uniform(0.5, 1.5, g.ndof(p))
            kv.upload(k)
            fv.upload(f)
            xv.upload(np.zeros(g.ndof(p)))
            ks.append(k)
            fs.append(f)
            nnz0.append(A.nnz)
        nits, res = B.run()
        for i, ((sp, _, A, _, _, _, xv), g) in enumerate(zip(systems, grids)):
            A1 = _fresh(pa, sp, pb, ks[i], fs[i], eop=False)
            assert A.nnz == A1.nnz, (trial, i)
            assert _same_csr(A, A1), (trial, i)
            moved += A.nnz != nnz0[i]
            xo = _oracle_solution(g, p, ks[i], fs[i])
            assert _relerr(xv.numpy(), xo) < 1e-8, (trial, i)
            assert res[i] <= 1e-10 * np.linalg.norm(A1.rhs().numpy()) * 1.0001
    if p == 3:
        # the drawn coefficients moved the pattern of some system, and the
        # refill followed it
        assert moved > 0
    # the matrices work with the other solvers (the solver copy is rebuilt
    # from the refilled arrays); the element operator is back after the batch
    (sp, _, A, kv, fv, _, xv), g = systems[0], grids[0]
    x64 = pa.Vec(A.n)
    pa.pcg(A, A.rhs(), x64, rtol=1e-12, btol=1e-12)
    xo = _oracle_solution(g, p, ks[0], fs[0])
    assert _relerr(x64.numpy(), xo) < 1e-9
    del B
    y, y1 = pa.Vec(A.n), pa.Vec(A.n)
    A.element_apply(x64, y)
    A.matvec(x64, y1)
    assert _relerr(y.numpy(), y1.numpy()) < 2e-6


def test_batch_warm_start_and_repeat(hip):
    """A second run from the converged x stops at once; a run after x is
    reset repeats the first bit for bit."""
    pa = hip
    meshes = batch_meshes(nmesh=2, n=3, seed=2, lmin=1, lmax=2)
    pb = pa.Problem.poisson(use_kappa=True)
    rng = np.random.default_rng(5)
    systems = []
    for g, _, _ in meshes:
        sp = pa.Space(g, 2)
        nd = g.ndof(2)
        k = np.exp(rng.normal(0.0, 0.5, g.nelem))
        A = _fresh(pa, sp, pb, k, np.ones(nd))
        systems.append((sp, pb, A, pa.Vec(g.nelem, k), pa.Vec(nd, np.ones(nd)),
                        pa.Vec(nd, np.zeros(nd)), pa.Vec(nd, np.zeros(nd))))
    B = pa.Batch(systems, rtol=1e-10, btol=1e-10)
    n1, r1 = B.run()
    x1 = [s[6].numpy() for s in systems]
    n2, _ = B.run()
    assert (n2 <= 2).all() and (n1 > 20).all()
    for s in systems:
        s[6].upload(np.zeros(s[2].n))
    n3, r3 = B.run()
    assert np.array_equal(n1, n3) and np.array_equal(r1, r3)
    for s, x in zip(systems, x1):
        assert np.array_equal(s[6].numpy(), x)


def test_batch_locks_and_requirements(hip):
    """While the batch exists its matrix, space and vectors cannot be freed
    and the matrix gets no new element operator; a matrix without an element
    operator is refused."""
    pa = hip
    (g, _, _), = batch_meshes(nmesh=1, n=3, seed=4, lmin=1, lmax=1)
    pb = pa.Problem.poisson(use_kappa=True)
    sp = pa.Space(g, 2)
    nd = g.ndof(2)
    k = np.ones(g.nelem)
    vecs = (pa.Vec(g.nelem, k), pa.Vec(nd, np.ones(nd)), pa.Vec(nd, np.zeros(nd)),
            pa.Vec(nd, np.zeros(nd)))
    A0 = _fresh(pa, sp, pb, k, np.ones(nd), eop=False)
    with pytest.raises(pa.PhgHipError):
        pa.Batch([(sp, pb, A0) + vecs])
    A = _fresh(pa, sp, pb, k, np.ones(nd))
    B = pa.Batch([(sp, pb, A) + vecs], maxit=500)
    n1, _ = B.run()
    assert 0 < n1[0] < 500
    with pytest.raises(pa.PhgHipError) as ei:
        A.element_operator(sp, pb, kappa=k)
    assert ei.value.code == -5
    lib = pa.hip_lib
    for fn, h in ((lib.phg_hip_mat_destroy, A._h), (lib.phg_hip_space_destroy, sp._h),
                  (lib.phg_hip_vec_destroy, vecs[0]._h), (lib.phg_hip_vec_destroy, vecs[3]._h)):
        assert fn(h) == -5
    del B
    A.element_operator(sp, pb, kappa=k)
