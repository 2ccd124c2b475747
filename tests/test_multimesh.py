"""Batched solve of independent systems (phg_hip_pcg_multi): every system's
solution is bitwise the one phg_hip_pcg_reliable gives on its own (same
kernels, same fixed-order reductions), with and without the HIP graph, and
matches the oracle's PCG solution."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as orc
from phg_amd.multimesh import assemble_batch, batch_meshes

pytestmark = pytest.mark.gpu


def _oracle_solution(g, p, kappa):
    nd = g.ndof(p)
    A = orc.assemble(orc.problem(p, use_kappa=True), g.verts(), g.elem_verts(),
                     g.dof_map(p), nd, g.dof_bdry(p), np.zeros(nd), np.ones(nd), kappa)
    orc.eliminate(A, g.dof_bdry(p))
    x, _, _ = orc.pcg(A, A.rhs, rtol=1e-14, btol=1e-14, maxit=100000)
    return x


@pytest.mark.parametrize("graph", [True, False])
def test_multi_matches_single_and_oracle(hip, graph):
    pa = hip
    meshes = batch_meshes(nmesh=4, n=3, seed=3, lmin=1, lmax=4)
    sys_ = assemble_batch(meshes, p=2, seed=4)
    As = [A for _, A, _ in sys_]
    bs = [b for _, _, b in sys_]
    for A, fmt in zip(As, (0, 4, 6, 7)):
        A.set_format(fmt, fmt)
    xs = [pa.Vec(A.n) for A in As]
    nits, res = pa.pcg_multi(As, bs, xs, rtol=1e-10, btol=1e-10, use_graph=graph)
    rng = np.random.default_rng(4)
    for i, (A, b, x) in enumerate(zip(As, bs, xs)):
        x1 = pa.Vec(A.n)
        n1, _, r1 = pa.pcg_reliable(A, b, x1, rtol=1e-10, btol=1e-10)
        assert nits[i] == n1 and res[i] == r1
        assert np.array_equal(x.numpy(), x1.numpy())
        g = meshes[i][0]
        kappa = np.exp(rng.normal(0.0, 0.5, size=g.nelem))
        xo = _oracle_solution(g, 2, kappa)
        assert np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo) < 1e-8


def test_multi_rejects_duplicates(hip):
    pa = hip
    meshes = batch_meshes(nmesh=1, n=2, seed=1, lmin=1, lmax=1)
    (_, A, b), = assemble_batch(meshes, p=1)
    x, y = pa.Vec(A.n), pa.Vec(A.n)
    with pytest.raises(pa.PhgHipError):
        pa.pcg_multi([A, A], [b, b], [x, y])
