"""Parity of the HIP path with the oracle, through the C ABI.

Bar: element matrices, CSR storage order, CSR values and right-hand sides are
bit-identical to the oracle (the oracle keeps PHG's floating-point operation
order); solutions agree with the oracle's FP64 PCG to 1e-9 relative.
"""
import numpy as np
import pytest

import oracle as orc
from tests._cases import elasticity_case, poisson_case

pytestmark = pytest.mark.gpu


def _space(pa, c, dim=1):
    return pa.Space(c["grid"], c["p"], bdry=c["bdry"])


def _oracle_problem(c, pb):
    return orc.problem(c["p"], dim=pb.dim, use_kappa=pb.use_kappa,
                       mass_coef=pb.mass_coef, rhs_order=pb.rhs_order, E=pb.E, nu=pb.nu,
                       qcache=bool(pb.qcache))


def _bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


CASES = [
    ("sin", 1, dict()), ("sin", 2, dict()), ("sin", 3, dict()),
    ("kappa", 1, dict(use_kappa=True)), ("kappa", 2, dict(use_kappa=True)),
    ("kappa", 3, dict(use_kappa=True)),
    ("sin", 2, dict(mass_coef=1.0)), ("sin", 3, dict(mass_coef=2.5)),
    ("sin", 1, dict(qcache=True)), ("sin", 2, dict(qcache=True)),
    ("sin", 3, dict(qcache=True)),
    ("sin", 4, dict()), ("kappa", 4, dict(use_kappa=True)), ("sin", 4, dict(mass_coef=1.5)),
]


@pytest.mark.parametrize("kind,p,opts", CASES)
def test_element_matrices_bitexact(hip, kind, p, opts):
    pa = hip
    c = poisson_case(3, p, kind)
    sp = _space(pa, c)
    pb = pa.Problem.poisson(**opts)
    K, B = pa.element_matrices(sp, pb, kappa=c["kappa"], f=c["f"])
    opb = _oracle_problem(c, pb)
    verts, ev, m = c["verts"], c["elem_verts"], c["map"]
    for e in range(c["grid"].nelem):
        Ko, Bo = orc.element(opb, verts[ev[e]], 0.0 if c["kappa"] is None else c["kappa"][e],
                             c["f"][m[e]])
        assert _bits_equal(K[e], Ko), f"element {e} stiffness differs"
        assert _bits_equal(B[e], Bo), f"element {e} load differs"


def test_qcache_rejects_other_integrands(hip):
    pa = hip
    c = poisson_case(2, 2, "kappa")
    sp = _space(pa, c)
    for pb in (pa.Problem.poisson(use_kappa=True, qcache=True),
               pa.Problem.poisson(mass_coef=1.0, qcache=True),
               pa.Problem.poisson(rhs_order=2, qcache=True)):
        with pytest.raises(pa.PhgHipError):
            pa.assemble(sp, pb, kappa=c["kappa"], f=c["f"], g=c["g"])


def test_elasticity_element_bitexact(hip):
    pa = hip
    c = elasticity_case(2)
    sp = _space(pa, c)
    pb = pa.Problem.elasticity()
    K, B = pa.element_matrices(sp, pb, f=c["f"])
    opb = _oracle_problem(c, pb)
    for e in range(c["grid"].nelem):
        floc = c["f"].reshape(-1, 3)[c["map"][e]].ravel()
        Ko, Bo = orc.element(opb, c["verts"][c["elem_verts"][e]], 0.0, floc)
        assert _bits_equal(K[e], Ko)
        assert _bits_equal(B[e], Bo)


ASM_CASES = [("sin", 1, 4), ("sin", 2, 3), ("sin", 3, 2), ("kappa", 2, 3),
             ("kappa", 3, 2), ("linear", 1, 5), ("sin", 4, 2), ("kappa", 4, 2)]


def _assemble_both(pa, c, pb):
    sp = _space(pa, c)
    A = pa.assemble(sp, pb, kappa=c["kappa"], f=c["f"], g=c["g"])
    Ao = orc.assemble(_oracle_problem(c, pb), c["verts"], c["elem_verts"], c["map"],
                      c["ndof"], c["bdry"], c["g"], c["f"], c["kappa"])
    return sp, A, Ao


@pytest.mark.parametrize("kind,p,n,qc", [a + (False,) for a in ASM_CASES] +
                         [("sin", 1, 4, True), ("sin", 2, 3, True), ("sin", 3, 2, True)])
def test_csr_bitexact(hip, kind, p, n, qc):
    pa = hip
    c = poisson_case(n, p, kind)
    pb = pa.Problem.poisson(use_kappa=(kind == "kappa"), qcache=qc)
    sp, A, Ao = _assemble_both(pa, c, pb)
    rp, ci, va, rhs = A.download()
    assert np.array_equal(rp, Ao.rowptr)
    assert np.array_equal(ci, Ao.cols), "column storage order differs"
    assert _bits_equal(va, Ao.vals)
    assert _bits_equal(rhs, Ao.rhs)
    # after Dirichlet elimination and RHS lifting
    A.eliminate()
    orc.eliminate(Ao, c["bdry"])
    rp, ci, va, rhs = A.download()
    assert np.array_equal(rp, Ao.rowptr)
    assert np.array_equal(ci, Ao.cols)
    assert _bits_equal(va, Ao.vals)
    assert _bits_equal(rhs, Ao.rhs)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_elasticity_csr_bitexact(hip, p):
    # element rows of 12, 30 and 60 entries: four, two and one rows per
    # wavefront in the insertion kernel
    pa = hip
    c = elasticity_case(2, p)
    pb = pa.Problem.elasticity()
    bd3 = np.repeat(c["bdry"], 3)
    sp, A, Ao = _assemble_both(pa, c, pb)
    rp, ci, va, rhs = A.download()
    assert np.array_equal(rp, Ao.rowptr)
    assert np.array_equal(ci, Ao.cols)
    assert _bits_equal(va, Ao.vals)
    assert _bits_equal(rhs, Ao.rhs)
    A.eliminate()
    orc.eliminate(Ao, bd3)
    rp, ci, va, rhs = A.download()
    assert np.array_equal(ci, Ao.cols)
    assert _bits_equal(va, Ao.vals)
    assert _bits_equal(rhs, Ao.rhs)


@pytest.mark.parametrize("kind,p,n", [("sin", 1, 8), ("sin", 2, 4), ("kappa", 2, 4),
                                      ("kappa", 3, 3), ("linear", 2, 3), ("kappa", 4, 3)])
def test_pcg_matches_oracle(hip, kind, p, n):
    pa = hip
    c = poisson_case(n, p, kind)
    pb = pa.Problem.poisson(use_kappa=(kind == "kappa"))
    sp, A, Ao = _assemble_both(pa, c, pb)
    A.eliminate()
    orc.eliminate(Ao, c["bdry"])
    xo, nits_o, res_o = orc.pcg(Ao, Ao.rhs, rtol=1e-12, btol=1e-12)
    b = A.rhs()
    x = pa.Vec(A.n)
    nits, res = pa.pcg(A, b, x, rtol=1e-12, btol=1e-12)
    xg = x.numpy()
    rel = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    # reductions run in a different order, so trajectories agree to rounding
    assert abs(nits - nits_o) <= 2, (nits, nits_o)
    assert rel < 1e-9, rel


@pytest.mark.parametrize("fmt", list(range(10)))
def test_pcg_every_spmv_layout(hip, fmt):
    """FP64 PCG and the mixed solver agree with the oracle on each layout."""
    pa = hip
    c = poisson_case(5, 2, "kappa")
    pb = pa.Problem.poisson(use_kappa=True)
    sp, A, Ao = _assemble_both(pa, c, pb)
    A.eliminate()
    A.set_format(fmt, fmt)
    orc.eliminate(Ao, c["bdry"])
    xo, nits_o, _ = orc.pcg(Ao, Ao.rhs, rtol=1e-12, btol=1e-12)
    x = pa.Vec(A.n)
    nits, _ = pa.pcg(A, A.rhs(), x, rtol=1e-12, btol=1e-12)
    assert abs(nits - nits_o) <= 2
    assert np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo) < 1e-9
    x2 = pa.Vec(A.n)
    pa.pcg_mixed(A, A.rhs(), x2, rtol=1e-11, btol=1e-11, inner_rtol=1e-4)
    assert np.linalg.norm(x2.numpy() - xo) / np.linalg.norm(xo) < 1e-8
    x3 = pa.Vec(A.n)
    pa.pcg_reliable(A, A.rhs(), x3, rtol=1e-11, btol=1e-11)
    assert np.linalg.norm(x3.numpy() - xo) / np.linalg.norm(xo) < 1e-8


def test_tuner_picks_measured_layout(hip):
    pa = hip
    c = poisson_case(6, 2, "kappa")
    pb = pa.Problem.poisson(use_kappa=True)
    sp = _space(pa, c)
    A = pa.assemble(sp, pb, kappa=c["kappa"], f=c["f"], g=c["g"])
    A.eliminate()
    t = A.tune(reps=5)
    names = tuple(pa.FMT_NAMES.values())
    assert all(t[k][0] > 0 and t[k][1] > 0 for k in names), t
    f64, f32 = t["chosen"]
    assert t[f64][0] == min(t[k][0] for k in names)
    assert t[f32][1] == min(t[k][1] for k in names)
    x = pa.Vec(A.n)
    nits, res = pa.pcg(A, A.rhs(), x, rtol=1e-10, btol=1e-10)
    assert nits > 0


def test_matvec_bitexact_with_phg_row_sums(hip):
    """phgMatVec sums each row in storage order (matvec.c:920-928); the
    streaming SpMV stages products in LDS and sums them in the same order."""
    pa = hip
    c = poisson_case(5, 3, "kappa")
    pb = pa.Problem.poisson(use_kappa=True)
    sp, A, Ao = _assemble_both(pa, c, pb)
    A.eliminate()
    orc.eliminate(Ao, c["bdry"])
    xv = np.random.default_rng(9).standard_normal(A.n)
    x, y = pa.Vec(A.n, xv), pa.Vec(A.n)
    A.matvec(x, y)
    assert _bits_equal(y.numpy(), orc.matvec(Ao, xv))


def test_pcg_nonzero_initial_guess(hip):
    pa = hip
    c = poisson_case(4, 2, "kappa")
    pb = pa.Problem.poisson(use_kappa=True)
    sp, A, Ao = _assemble_both(pa, c, pb)
    A.eliminate()
    orc.eliminate(Ao, c["bdry"])
    x0 = np.random.default_rng(3).standard_normal(Ao.n)
    xo, nits_o, _ = orc.pcg(Ao, Ao.rhs, x0=x0, rtol=1e-12, btol=1e-14)
    x = pa.Vec(A.n, x0)
    nits, _ = pa.pcg(A, A.rhs(), x, rtol=1e-12, btol=1e-14)
    assert abs(nits - nits_o) <= 2
    assert np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo) < 1e-9


def test_pcg_zero_rhs_returns_zero(hip):
    pa = hip
    c = poisson_case(3, 1, "linear")
    pb = pa.Problem.poisson()
    sp = _space(pa, c)
    A = pa.assemble(sp, pb, f=np.zeros(c["ndof"]), g=np.zeros(c["ndof"]))
    A.eliminate()
    x = pa.Vec(A.n, np.ones(A.n))
    nits, res = pa.pcg(A, pa.Vec(A.n), x)
    # solver-phg.c:497-501: x = 0 and *res untouched, so PCG_Solve reports
    # its initial 1.0
    assert nits == 0 and np.all(x.numpy() == 0.0)
    assert res == 1.0


def test_matvec_and_blas1(hip):
    pa = hip
    c = poisson_case(4, 2, "kappa")
    pb = pa.Problem.poisson(use_kappa=True)
    sp, A, Ao = _assemble_both(pa, c, pb)
    A.eliminate()
    orc.eliminate(Ao, c["bdry"])
    rng = np.random.default_rng(5)
    xv, yv = rng.standard_normal(A.n), rng.standard_normal(A.n)
    x, y = pa.Vec(A.n, xv), pa.Vec(A.n, yv)
    A.matvec(x, y, alpha=2.0, beta=-0.5)
    ref = 2.0 * orc.matvec(Ao, xv) - 0.5 * yv
    assert np.allclose(y.numpy(), ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    assert abs(x.dot(pa.Vec(A.n, yv)) - xv @ yv) < 1e-12 * np.abs(xv) @ np.abs(yv)
    assert abs(x.norm2() - np.linalg.norm(xv)) < 1e-13 * np.linalg.norm(xv)
    y2 = pa.Vec(A.n, yv)
    y2.axpby(3.0, x, 0.25)
    assert np.allclose(y2.numpy(), 3.0 * xv + 0.25 * yv, rtol=1e-15, atol=1e-15)


def test_mixed_precision_refinement(hip):
    pa = hip
    c = poisson_case(6, 2, "kappa", sigma=0.5)
    pb = pa.Problem.poisson(use_kappa=True)
    sp, A, Ao = _assemble_both(pa, c, pb)
    A.eliminate()
    orc.eliminate(Ao, c["bdry"])
    xo, _, _ = orc.pcg(Ao, Ao.rhs, rtol=1e-14, btol=1e-14)
    x = pa.Vec(A.n)
    outer, inner, res = pa.pcg_mixed(A, A.rhs(), x, rtol=1e-11, btol=1e-11,
                                     inner_rtol=1e-4)
    rel = np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo)
    assert outer >= 2 and inner > 0
    assert rel < 1e-8, rel   # 1e-8 relative to the FP64 oracle


@pytest.mark.parametrize("delta", [0.1, 0.01])
def test_reliable_update_solver(hip, delta):
    """FP32 PCG with FP64 residual replacement: the stopping test runs on the
    FP64 residual, so |b - A x| <= tol holds for the returned x, and the
    recurrence is not restarted, so the iteration count stays close to FP64
    PCG's."""
    pa = hip
    c = poisson_case(8, 2, "kappa", sigma=0.5)
    pb = pa.Problem.poisson(use_kappa=True)
    sp, A, Ao = _assemble_both(pa, c, pb)
    A.eliminate()
    orc.eliminate(Ao, c["bdry"])
    b = Ao.rhs
    xo, _, _ = orc.pcg(Ao, b, rtol=1e-14, btol=1e-14)
    _, nits64, _ = orc.pcg(Ao, b, rtol=1e-10, btol=1e-10)
    x = pa.Vec(A.n)
    nits, upd, res = pa.pcg_reliable(A, A.rhs(), x, rtol=1e-10, btol=1e-10,
                                     delta=delta)
    xv = x.numpy()
    r = b - orc.matvec(Ao, xv)
    tol = 1e-10 * np.linalg.norm(b)        # r0 = b for x0 = 0
    assert np.linalg.norm(r) <= tol * (1 + 1e-6), (np.linalg.norm(r), tol)
    assert abs(res - np.linalg.norm(r)) <= 1e-3 * tol
    assert upd >= 2
    assert nits <= 1.3 * nits64 + 5, (nits, nits64)
    assert np.linalg.norm(xv - xo) / np.linalg.norm(xo) < 1e-8


def test_reliable_nonzero_guess_and_zero_rhs(hip):
    pa = hip
    c = poisson_case(4, 2, "kappa")
    pb = pa.Problem.poisson(use_kappa=True)
    sp, A, Ao = _assemble_both(pa, c, pb)
    A.eliminate()
    orc.eliminate(Ao, c["bdry"])
    xo, _, _ = orc.pcg(Ao, Ao.rhs, rtol=1e-14, btol=1e-14)
    x0 = np.random.default_rng(3).standard_normal(A.n)
    x = pa.Vec(A.n, x0)
    nits, upd, res = pa.pcg_reliable(A, A.rhs(), x, rtol=1e-11, btol=0.0)
    assert nits > 0
    assert np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo) < 1e-8
    z = pa.Vec(A.n)
    x2 = pa.Vec(A.n, x0)
    nits, upd, res = pa.pcg_reliable(A, z, x2)
    assert nits == 0 and res == 0.0
    assert np.all(x2.numpy() == 0.0)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_dof_norms(hip, p):
    """phgDofNormL2 of a DOF and of its gradient: closed forms for v = x
    (1/3 and 1 on the unit cube) and the oracle's quadratic forms
    v^T M v, v^T K1 v for a random DOF."""
    pa = hip
    c = poisson_case(3, p, "kappa")
    sp = _space(pa, c)
    l2, h1 = pa.dof_norms(sp, pa.Vec(c["ndof"], c["coords"][:, 0].copy()))
    assert abs(l2 * l2 - 1.0 / 3.0) < 1e-13 and abs(h1 * h1 - 1.0) < 1e-13
    v = np.random.default_rng(p).standard_normal(c["ndof"])
    l2, h1 = pa.dof_norms(sp, pa.Vec(c["ndof"], v))
    half_vMv, _ = orc.mass_apply(p, c["verts"], c["elem_verts"], c["map"], v)
    vKv = -orc.adjoint_grad(p, c["verts"], c["elem_verts"], c["map"],
                            np.ones(c["grid"].nelem), v, v).sum()
    assert abs(l2 * l2 - 2 * half_vMv) <= 1e-12 * 2 * half_vMv
    assert abs(h1 * h1 - vKv) <= 1e-12 * vKv


@pytest.mark.parametrize("p", [2, 3])
def test_cube4_fixture_assembly_and_solve(hip, p):
    """The reference's test/cube4.dat mesh (a refined, unstructured
    tetrahedralisation): CSR bit for bit against the oracle before and after
    elimination, and the PCG solution."""
    import os
    pa = hip
    g = pa.Grid.read_albert(os.path.join(os.path.dirname(__file__), "golden", "cube4.dat"))
    nd = g.ndof(p)
    bd = g.dof_bdry(p)
    xyz = g.dof_coords(p)
    kappa = np.exp(np.random.default_rng(12).normal(0.0, 0.5, size=g.nelem))
    f = 1.0 + xyz[:, 0]
    gv = xyz[:, 1] * xyz[:, 2]
    sp = pa.Space(g, p, bdry=bd)
    pb = pa.Problem.poisson(use_kappa=True)
    A = pa.assemble(sp, pb, kappa=kappa, f=f, g=gv)
    Ao = orc.assemble(orc.problem(p, use_kappa=True), g.verts(), g.elem_verts(),
                      g.dof_map(p), nd, bd, gv, f, kappa)
    rp, ci, va, rhs = A.download()
    assert np.array_equal(ci, Ao.cols) and _bits_equal(va, Ao.vals)
    assert _bits_equal(rhs, Ao.rhs)
    A.eliminate()
    orc.eliminate(Ao, bd)
    rp, ci, va, rhs = A.download()
    assert np.array_equal(ci, Ao.cols) and _bits_equal(va, Ao.vals)
    assert _bits_equal(rhs, Ao.rhs)
    xo, _, _ = orc.pcg(Ao, Ao.rhs, rtol=1e-13, btol=1e-13)
    x = pa.Vec(A.n)
    pa.pcg(A, A.rhs(), x, rtol=1e-12, btol=1e-12)
    assert np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo) < 1e-9


@pytest.mark.parametrize("eliminated", [False, True])
def test_mat_from_csr_phg_packed_matrix(hip, eliminated):
    """A system handed over as a host CSR (the form of PHG's packed matrix):
    before boundary treatment, eliminated on the GPU bit for bit as the
    oracle does; or already eliminated; the PCG solutions agree."""
    pa = hip
    c = poisson_case(5, 2, "sin")
    Ao = orc.assemble(orc.problem(2), c["verts"], c["elem_verts"], c["map"], c["ndof"],
                      c["bdry"], c["g"], c["f"])
    if eliminated:
        orc.eliminate(Ao, c["bdry"])
    A = pa.mat_from_csr(Ao.rowptr, Ao.cols, Ao.vals, Ao.rhs, c["bdry"], eliminated)
    if not eliminated:
        A.eliminate()
        orc.eliminate(Ao, c["bdry"])
        rp, ci, va, rhs = A.download()
        assert np.array_equal(ci, Ao.cols) and _bits_equal(va, Ao.vals)
        assert _bits_equal(rhs, Ao.rhs)
    xo, nits_o, _ = orc.pcg(Ao, Ao.rhs, rtol=1e-12, btol=1e-12)
    x = pa.Vec(A.n)
    nits, _ = pa.pcg(A, A.rhs(), x, rtol=1e-12, btol=1e-12)
    assert abs(nits - nits_o) <= 2
    assert np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo) < 1e-10


def _fan_mesh(path, k):
    """k tetrahedra above and k below a ring of k vertices around the origin:
    the centre vertex touches 2k elements, so its row holds every node of
    the mesh (P1: k + 3 columns) and, for k large enough, more distinct
    columns than the insertion kernel's LDS table."""
    th = 2 * np.pi * np.arange(k) / k
    v = [(0.0, 0.0, 0.0)] + [(np.cos(t), np.sin(t), 0.0) for t in th] + \
        [(0.0, 0.0, 1.0), (0.0, 0.0, -1.0)]
    el = []
    for i in range(1, k + 1):
        j = i % k + 1
        el.append((0, i, j, k + 1))
        el.append((0, i, j, k + 2))
    with open(path, "w") as fh:
        fh.write(f"DIM: 3\nDIM_OF_WORLD: 3\nnumber of vertices: {len(v)}\n"
                 f"number of elements: {len(el)}\n\nvertex coordinates:\n")
        fh.writelines("%.17g %.17g %.17g\n" % (x, y, z) for x, y, z in v)
        fh.write("\nelement vertices:\n")
        fh.writelines(" ".join(map(str, e)) + "\n" for e in el)
        # face j lies opposite vertex j: only the face away from the centre
        # is on the boundary
        fh.write("\nelement boundaries:\n" + "1 0 0 0\n" * len(el))
        fh.write("\nelement type:\n" + "0\n" * len(el))


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_high_valence_rows_bitexact(hip, tmp_path, p):
    """Rows with more distinct columns than the insertion kernel's hash table
    holds take its serial replay; the CSR stays bit for bit the oracle's."""
    pa = hip
    path = str(tmp_path / "fan.dat")
    _fan_mesh(path, 144)
    g = pa.Grid.read_albert(path)
    nd = g.ndof(p)
    bd = g.dof_bdry(p)
    xyz = g.dof_coords(p)
    kappa = np.exp(np.random.default_rng(5).normal(0.0, 0.5, size=g.nelem))
    f = 1.0 + xyz[:, 0]
    gv = xyz[:, 1] * xyz[:, 2]
    sp = pa.Space(g, p, bdry=bd)
    pb = pa.Problem.poisson(use_kappa=True)
    A = pa.assemble(sp, pb, kappa=kappa, f=f, g=gv)
    Ao = orc.assemble(orc.problem(p, use_kappa=True), g.verts(), g.elem_verts(),
                      g.dof_map(p), nd, bd, gv, f, kappa)
    rp, ci, va, rhs = A.download()
    # the centre vertex's row exceeds the table (P1, P2: 16 lanes and 128
    # entries per row; P3: 32 lanes, 128 entries; P4: 64 lanes, 512 entries)
    assert np.diff(Ao.rowptr).max() > {1: 128, 2: 128, 3: 128, 4: 512}[p]
    assert np.array_equal(rp, Ao.rowptr)
    assert np.array_equal(ci, Ao.cols) and _bits_equal(va, Ao.vals)
    assert _bits_equal(rhs, Ao.rhs)
    A.eliminate()
    orc.eliminate(Ao, bd)
    rp, ci, va, rhs = A.download()
    assert np.array_equal(ci, Ao.cols) and _bits_equal(va, Ao.vals)
    assert _bits_equal(rhs, Ao.rhs)
