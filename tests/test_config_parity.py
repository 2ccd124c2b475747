"""Parity at the sizes of the benchmark workloads, through the C ABI.

C1  P1 Poisson sin, 32^3 x 6 Kuhn mesh (35,937 DOFs): CSR bit for bit before
    and after elimination, Jacobi-PCG to rtol 1e-10 against the oracle's
    (iteration count within 2, solution within 1e-9).
C2  P2 kappa-Poisson, 64^3 x 6 (2,146,689 DOFs, 57.3M entries): CSR columns
    in storage order, values and right-hand side bit for bit before and after
    elimination, so the exact-zero drop set of matvec.c:5460 matches at full
    size; the mixed and FP64 solutions against the oracle's FP64 solution at
    the committed sample (tests/golden/c2_64_oracle.npz).
C3  P1 adjoint on the 48^3 x 6 mesh read from an ALBERT file: the GPU adjoint
    gradient against the oracle's CPU adjoint to 1e-8 (max-norm relative).
C5  the batch of 8 unevenly refined P3 meshes: every system of the graph-
    replayed batch on the element operator against the oracle's PCG (1e-8),
    and the P3 CSR of a refined mesh bit for bit against the oracle.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle as orc
from phg_amd.multimesh import assemble_batch, batch_meshes
from tests._adjoint import c3_case, oracle_adjoint
from tests._cases import poisson_case

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def _assert_csr_equal(A, Ao):
    rp, ci, va, rhs = A.download()
    assert np.array_equal(rp, Ao.rowptr), "row pointers differ"
    assert np.array_equal(ci, Ao.cols), "column storage order differs"
    assert _bits_equal(va, Ao.vals), "values differ"
    assert _bits_equal(rhs, Ao.rhs), "right-hand side differs"


def _both(pa, c, use_kappa):
    sp = pa.Space(c["grid"], c["p"], bdry=c["bdry"])
    pb = pa.Problem.poisson(use_kappa=use_kappa)
    A = pa.assemble(sp, pb, kappa=c["kappa"], f=c["f"], g=c["g"])
    Ao = orc.assemble(orc.problem(c["p"], use_kappa=use_kappa), c["verts"], c["elem_verts"],
                      c["map"], c["ndof"], c["bdry"], c["g"], c["f"], c["kappa"])
    return sp, pb, A, Ao


def test_c1_32_csr_and_pcg(hip):
    pa = hip
    c = poisson_case(32, 1, "sin", mesh="phg")
    assert c["ndof"] == 35937 and c["grid"].nelem == 196608
    sp, pb, A, Ao = _both(pa, c, False)
    _assert_csr_equal(A, Ao)
    A.eliminate()
    orc.eliminate(Ao, c["bdry"])
    _assert_csr_equal(A, Ao)
    xo, nits_o, res_o = orc.pcg(Ao, Ao.rhs, rtol=1e-10, btol=1e-10)
    x = pa.Vec(A.n)
    nits, res = pa.pcg(A, A.rhs(), x, rtol=1e-10, btol=1e-10)
    rel = np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo)
    print(f"C1 32^3: its {nits} (oracle {nits_o}), residual {res:.3e} (oracle "
          f"{res_o:.3e}), rel. diff {rel:.2e}", flush=True)
    assert abs(nits - nits_o) <= 2
    assert rel < 1e-9


def test_c2_64_csr_bitexact_and_solutions(hip):
    pa = hip
    c = poisson_case(64, 2, "kappa", seed=2024, mesh="phg")
    assert c["ndof"] == 2146689 and c["grid"].nelem == 1572864
    sp, pb, A, Ao = _both(pa, c, True)
    _assert_csr_equal(A, Ao)
    A.eliminate()
    orc.eliminate(Ao, c["bdry"])
    _assert_csr_equal(A, Ao)
    n, nnz = A.info()
    print(f"C2 64^3: CSR bit-exact, n={n} nnz={nnz}", flush=True)
    del Ao
    ref = np.load(os.path.join(GOLDEN, "c2_64_oracle.npz"))
    idx, xs = ref["index"], ref["x"]
    A.element_operator(sp, pb, kappa=c["kappa"])
    for name, solve in (("mixed", lambda x: pa.pcg_reliable(A, A.rhs(), x, rtol=1e-8,
                                                            btol=1e-8)),
                        ("fp64", lambda x: pa.pcg(A, A.rhs(), x, maxit=100000, rtol=1e-8,
                                                  btol=1e-8))):
        x = pa.Vec(A.n)
        solve(x)
        xv = x.numpy()
        rel = np.linalg.norm(xv[idx] - xs) / np.linalg.norm(xs)
        rel_norm = abs(np.linalg.norm(xv) - float(ref["norm2"])) / float(ref["norm2"])
        print(f"  {name}: sampled rel. error vs oracle {rel:.2e}, |x| rel. diff "
              f"{rel_norm:.2e}", flush=True)
        assert rel < 1e-8 and rel_norm < 1e-8


def test_c3_48_adjoint_matches_cpu_adjoint(hip, tmp_path):
    pa = hip
    c = c3_case(48, 1)
    path = str(tmp_path / "cube48.dat")
    c["grid"].write_albert(path)
    grid = pa.Grid.read_albert(path)
    assert np.array_equal(grid.dof_map(1), c["map"])
    c["grid"] = grid
    sp = pa.Space(grid, 1, bdry=c["bdry"])
    pb = pa.Problem.poisson(use_kappa=True)

    def solve(kappa):
        A = pa.assemble(sp, pb, kappa=kappa, f=c["f"], g=c["g"])
        A.eliminate()
        u = pa.Vec(A.n)
        pa.pcg(A, A.rhs(), u, maxit=100000, rtol=1e-14, btol=1e-14)
        return A, u
    _, ud = solve(c["kappa_d"])
    A, u = solve(c["kappa"])
    J, g, nits = pa.adjoint_logkappa(sp, pb, c["kappa"], A, u, ud, rtol=1e-14)
    Jo, go, _, _, _ = oracle_adjoint(c, c["kappa"])
    err = np.abs(g - go).max() / np.abs(go).max()
    print(f"C3 48^3: J {J:.15e} (oracle {Jo:.15e}), adjoint its {nits}, gradient "
          f"max-norm rel. diff {err:.2e}", flush=True)
    assert abs(J - Jo) <= 1e-10 * abs(Jo)
    assert err <= 1e-8


def _oracle_pcg(g, p, kappa):
    nd = g.ndof(p)
    Ao = orc.assemble(orc.problem(p, use_kappa=True), g.verts(), g.elem_verts(),
                      g.dof_map(p), nd, g.dof_bdry(p), np.zeros(nd), np.ones(nd), kappa)
    orc.eliminate(Ao, g.dof_bdry(p))
    x, nits, _ = orc.pcg(Ao, Ao.rhs, rtol=1e-13, btol=1e-13, maxit=100000)
    return x, nits


def test_c5_batch_p3_graph_matches_oracle(hip):
    """The C5 workload family at a size the oracle solves in seconds: 8 P3
    kappa-Poisson problems on 5^3 x 6 meshes refined 2-8 levels toward seeded
    corners (4,582-44,836 DOFs), FP32 iterations on the split-bf16 MFMA
    element operator, one graph replayed for the whole batch."""
    pa = hip
    meshes = batch_meshes(nmesh=8, n=5, seed=5, lmin=2, lmax=8)
    kappas = []
    sys_ = assemble_batch(meshes, p=3, seed=6, element_operator=True, kappas=kappas)
    As = [A for _, A, _ in sys_]
    bs = [b for _, _, b in sys_]
    xs = [pa.Vec(A.n) for A in As]
    nits, res = pa.pcg_multi(As, bs, xs, rtol=1e-10, btol=1e-10, use_graph=True)
    for i, ((g, _, lev), x) in enumerate(zip(meshes, xs)):
        xo, nits_o = _oracle_pcg(g, 3, kappas[i])
        rel = np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo)
        print(f"  mesh {i}: {levels_str(g, lev)} its {nits[i]} (oracle FP64 {nits_o}), "
              f"rel. error {rel:.2e}", flush=True)
        assert rel < 1e-8


def levels_str(g, lev):
    return f"{lev} levels, {g.nelem} tets, {g.ndof(3)} DOFs,"


def test_c5_refined_mesh_p3_csr_bitexact(hip):
    """P3 CSR of the most refined C5-type mesh (8 levels toward a corner)
    against the oracle, before and after elimination."""
    pa = hip
    meshes = batch_meshes(nmesh=8, n=5, seed=5, lmin=2, lmax=8)
    g, _, lev = max(meshes, key=lambda m: m[2])
    assert lev == 8
    p = 3
    nd = g.ndof(p)
    bd = g.dof_bdry(p)
    xyz = g.dof_coords(p)
    kappa = np.exp(np.random.default_rng(8).normal(0.0, 0.5, size=g.nelem))
    f = 1.0 + xyz[:, 0] * xyz[:, 1]
    gv = np.sin(xyz[:, 2])
    sp = pa.Space(g, p, bdry=bd)
    A = pa.assemble(sp, pa.Problem.poisson(use_kappa=True), kappa=kappa, f=f, g=gv)
    Ao = orc.assemble(orc.problem(p, use_kappa=True), g.verts(), g.elem_verts(),
                      g.dof_map(p), nd, bd, gv, f, kappa)
    _assert_csr_equal(A, Ao)
    A.eliminate()
    orc.eliminate(Ao, bd)
    _assert_csr_equal(A, Ao)
