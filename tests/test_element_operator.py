"""Matrix-free FP32 operator (phg_hip_mat_element_operator): MFMA element
products with three-part bf16 operands against the assembled FP64 matrix,
and the reliable-update / batched solvers running their FP32 iterations on
it.  The FP64 side (assembled CSR, oracle) is the reference."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as orc
from phg_amd.multimesh import assemble_batch, batch_meshes
from tests._cases import elasticity_case, poisson_case

pytestmark = pytest.mark.gpu


def _system(pa, kind, p, n, mass=0.0):
    c = poisson_case(n, p, kind)
    pb = pa.Problem.poisson(use_kappa=(kind == "kappa"), mass_coef=mass)
    sp = pa.Space(c["grid"], p, bdry=c["bdry"])
    A = pa.assemble(sp, pb, kappa=c["kappa"], f=c["f"], g=c["g"])
    A.eliminate()
    return c, pb, sp, A


# (kind, order, mesh n): kappa and plain Laplacian, P1-P3
APPLY_CASES = [("kappa", 1, 6), ("sin", 1, 5), ("kappa", 2, 4), ("sin", 2, 3),
               ("kappa", 3, 3), ("sin", 3, 2)]


@pytest.mark.parametrize("kind,p,n", APPLY_CASES)
def test_apply_matches_assembled_matrix(hip, kind, p, n):
    """FP32 accuracy against the FP64 SpMV of the eliminated matrix.  The
    bound 2e-6 is what FP32 products and sums reach here; keeping only two
    bf16 parts (16 bits) of the operands gives errors near 1e-5, a wrong
    fragment layout or coefficient order errors of order 1."""
    pa = hip
    c, pb, sp, A = _system(pa, kind, p, n)
    A.element_operator(sp, pb, kappa=c["kappa"])
    rng = np.random.default_rng(10 + p)
    for trial in range(2):
        xv = rng.standard_normal(A.n)
        x, y64, y32 = pa.Vec(A.n, xv), pa.Vec(A.n), pa.Vec(A.n)
        A.matvec(x, y64)
        A.element_apply(x, y32)
        a, b = y64.numpy(), y32.numpy()
        rel = np.linalg.norm(b - a) / np.linalg.norm(a)
        assert rel < 2e-6, (trial, rel)
        assert np.max(np.abs(b - a)) < 2e-5 * np.max(np.abs(a))
    te, tg = A.element_time(reps=3)
    eb, gb = A.element_bytes()
    assert te > 0 and tg > 0 and eb > 0 and gb > 0


# meshes of several element blocks (blocks of 2048 / 1024 / 512 elements
# for P1 / P2 / P3), so rows shared between blocks go through the partials
# 32^3 x 6 = 196,608 elements: the block split runs on three element ranges
# of 65,536 elements each, whole blocks; 33^3 x 6 = 215,622 elements: three
# ranges whose ends fall inside a block, so each range ends with a short block
MULTI_BLOCK_CASES = [("kappa", 1, 16), ("kappa", 2, 9), ("sin", 3, 6), ("kappa", 1, 32),
                     ("kappa", 1, 33)]


# -div(kappa grad u) + a u = f (simplest.c:70 with a = mass_coef): the mass
# tensor in slot 6 of the element products, P1-P3, and one mesh of many
# blocks; a = 50 makes the mass term dominate the P2 kappa matrix
MASS_CASES = [("kappa", 1, 6, 1.0), ("sin", 1, 16, 3.0), ("sin", 2, 3, 1.0),
              ("kappa", 2, 4, 50.0), ("kappa", 2, 9, 0.5), ("sin", 3, 2, 1.0),
              ("kappa", 3, 3, 10.0)]


@pytest.mark.parametrize("kind,p,n,a", MASS_CASES)
def test_apply_with_mass_term(hip, kind, p, n, a):
    """Against the FP64 SpMV of the assembled matrix (stiffness plus a times
    the mass matrix of quad.c:5582, whose entries the CSR parity tests pin
    bit for bit to the oracle).  The products act on x - x_0 per element
    and the mass term takes x_0 (T^6 1)_i back: the third trial adds a large
    constant to x, which only the mass term sees -- a missing or misplaced
    correction gives errors of order 1 there."""
    pa = hip
    c, pb, sp, A = _system(pa, kind, p, n, mass=a)
    A.element_operator(sp, pb, kappa=c["kappa"])
    rng = np.random.default_rng(30 + p)
    bd = c["bdry"].astype(bool)
    for trial in range(3):
        xv = rng.standard_normal(A.n)
        if trial == 1:
            xv = np.where(bd, 0.0, 1.0)
        elif trial == 2:
            xv = xv + 1000.0
        x, y64, y32 = pa.Vec(A.n, xv), pa.Vec(A.n), pa.Vec(A.n)
        A.matvec(x, y64)
        A.element_apply(x, y32)
        ya, yb = y64.numpy(), y32.numpy()
        rel = np.linalg.norm(yb - ya) / np.linalg.norm(ya)
        assert rel < 2e-6, (trial, rel)
        assert np.max(np.abs(yb - ya)) < 2e-5 * np.max(np.abs(ya)), trial
    eb, _ = A.element_bytes()
    lay = A.element_layout()
    assert eb > 0 and lay["interior"] + lay["shared"] + lay["dirichlet"] == A.n


@pytest.mark.parametrize("p,n", [(1, 8), (2, 4), (3, 3)])
def test_reliable_solver_with_mass_term(hip, p, n):
    """Mixed-precision PCG with FP32 iterations on the operator, against the
    oracle's FP64 PCG on its own assembly of the same problem."""
    pa = hip
    a = 2.5
    c, pb, sp, A = _system(pa, "kappa", p, n, mass=a)
    Ao = orc.assemble(orc.problem(p, use_kappa=True, mass_coef=a), c["verts"],
                      c["elem_verts"], c["map"], c["ndof"], c["bdry"], c["g"], c["f"],
                      c["kappa"])
    orc.eliminate(Ao, c["bdry"])
    xo, _, _ = orc.pcg(Ao, Ao.rhs, rtol=1e-13, btol=1e-13, maxit=100000)
    x_csr = pa.Vec(A.n)
    n_csr, _, _ = pa.pcg_reliable(A, A.rhs(), x_csr, rtol=1e-10, btol=1e-10)
    A.element_operator(sp, pb, kappa=c["kappa"])
    x = pa.Vec(A.n)
    nits, upd, _ = pa.pcg_reliable(A, A.rhs(), x, rtol=1e-10, btol=1e-10)
    print("mass", p, n, "iterations csr", n_csr, "element operator", nits, flush=True)
    assert upd >= 1
    assert abs(nits - n_csr) <= max(3, n_csr // 10)
    assert np.linalg.norm(x.numpy() - xo) / np.linalg.norm(xo) < 1e-8


def test_layout_shared_with_and_without_mass(hip):
    """A mass matrix on a space whose layout a stiffness matrix built takes
    that layout and its own tiles and coefficients; both operators match
    their matrices, and the stiffness one is unchanged by the other's
    set-up."""
    pa = hip
    c = poisson_case(6, 2, "kappa")
    sp = pa.Space(c["grid"], 2, bdry=c["bdry"])
    outs = []
    mats = []
    for a in (0.0, 4.0):
        pb = pa.Problem.poisson(use_kappa=True, mass_coef=a)
        A = pa.assemble(sp, pb, kappa=c["kappa"], f=c["f"], g=c["g"])
        A.eliminate()
        A.element_operator(sp, pb, kappa=c["kappa"])
        mats.append(A)
    xv = np.random.default_rng(8).standard_normal(mats[0].n)
    x = pa.Vec(mats[0].n, xv)
    for A in mats:
        y64, y32 = pa.Vec(A.n), pa.Vec(A.n)
        A.matvec(x, y64)
        A.element_apply(x, y32)
        ya, yb = y64.numpy(), y32.numpy()
        assert np.linalg.norm(yb - ya) / np.linalg.norm(ya) < 2e-6
        outs.append(yb)
    assert not np.allclose(outs[0], outs[1])
    y = pa.Vec(mats[0].n)
    mats[0].element_apply(x, y)
    assert np.array_equal(y.numpy(), outs[0])


@pytest.mark.parametrize("kind,p,n", MULTI_BLOCK_CASES)
def test_apply_over_many_blocks(hip, kind, p, n):
    pa = hip
    c, pb, sp, A = _system(pa, kind, p, n)
    A.element_operator(sp, pb, kappa=c["kappa"])
    lay = A.element_layout()
    print("layout", p, n, lay, flush=True)
    assert lay["blocks"] > 2 and lay["shared"] > 0
    # every row is set once: by its only block, from partials, or Dirichlet
    assert lay["interior"] + lay["shared"] + lay["dirichlet"] == A.n
    assert lay["dirichlet"] == int(c["bdry"].astype(bool).sum())
    assert lay["entries"] == lay["interior"] + lay["partials"]
    rng = np.random.default_rng(20 + p)
    xv = rng.standard_normal(A.n)
    x, y64 = pa.Vec(A.n, xv), pa.Vec(A.n)
    A.matvec(x, y64)
    ys = []
    for _ in range(2):
        y32 = pa.Vec(A.n)
        A.element_apply(x, y32)
        ys.append(y32.numpy())
    a = y64.numpy()
    rel = np.linalg.norm(ys[0] - a) / np.linalg.norm(a)
    assert rel < 2e-6, rel
    assert np.max(np.abs(ys[0] - a)) < 2e-5 * np.max(np.abs(a))
    assert np.array_equal(ys[0], ys[1])            # fixed summation order


def test_apply_dirichlet_rows_are_diagonal(hip):
    """Rows of Dirichlet DOFs are diag * x; a vector supported on Dirichlet
    DOFs only touches no interior row (elimination removed those columns)."""
    pa = hip
    c, pb, sp, A = _system(pa, "kappa", 2, 3)
    A.element_operator(sp, pb, kappa=c["kappa"])
    bd = c["bdry"].astype(bool)
    xv = np.where(bd, 1.0 + np.arange(A.n) % 7, 0.0)
    x, y64, y32 = pa.Vec(A.n, xv), pa.Vec(A.n), pa.Vec(A.n)
    A.matvec(x, y64)
    A.element_apply(x, y32)
    b = y32.numpy()
    assert np.all(b[~bd] == 0.0)
    np.testing.assert_allclose(b[bd], y64.numpy()[bd], rtol=1e-7)


@pytest.mark.parametrize("p,n", [(2, 4), (3, 3)])
def test_reliable_solver_on_element_operator(hip, p, n):
    pa = hip
    c, pb, sp, A = _system(pa, "kappa", p, n)
    Ao = orc.assemble(orc.problem(p, use_kappa=True), c["verts"], c["elem_verts"], c["map"],
                      c["ndof"], c["bdry"], c["g"], c["f"], c["kappa"])
    orc.eliminate(Ao, c["bdry"])
    xo, nits_o, _ = orc.pcg(Ao, Ao.rhs, rtol=1e-13, btol=1e-13, maxit=100000)
    x_csr = pa.Vec(A.n)
    n_csr, _, _ = pa.pcg_reliable(A, A.rhs(), x_csr, rtol=1e-10, btol=1e-10)
    A.element_operator(sp, pb, kappa=c["kappa"])
    xs = []
    for _ in range(2):
        x = pa.Vec(A.n)
        nits, upd, res = pa.pcg_reliable(A, A.rhs(), x, rtol=1e-10, btol=1e-10)
        xs.append(x.numpy())
        # FP64 residual tolerance reached; FP32 iteration count as on the CSR
        assert upd >= 1
        assert abs(nits - n_csr) <= max(3, n_csr // 10)
        assert np.linalg.norm(xs[-1] - xo) / np.linalg.norm(xo) < 1e-8
    assert np.array_equal(xs[0], xs[1])           # deterministic
    A.use_element_operator(False)
    x2 = pa.Vec(A.n)
    n2, _, _ = pa.pcg_reliable(A, A.rhs(), x2, rtol=1e-10, btol=1e-10)
    assert n2 == n_csr and np.array_equal(x2.numpy(), x_csr.numpy())


def test_reliable_iterations_as_on_the_fp32_csr(hip):
    """The FP32 products act on x - x_0 per element (every T^m annihilates
    constants).  Products on x itself carry FP32 errors of the size of x
    rather than of its variation over the element, which cost the reliable
    solver about 10% more iterations on this P2 kappa problem (159 against
    145 in a CPU emulation); on x - x_0 it converges as on the FP32 CSR."""
    pa = hip
    c, pb, sp, A = _system(pa, "kappa", 2, 16)
    x_csr = pa.Vec(A.n)
    n_csr, _, _ = pa.pcg_reliable(A, A.rhs(), x_csr, rtol=1e-8, btol=1e-8)
    A.element_operator(sp, pb, kappa=c["kappa"])
    x = pa.Vec(A.n)
    nits, _, _ = pa.pcg_reliable(A, A.rhs(), x, rtol=1e-8, btol=1e-8)
    print("iterations csr", n_csr, "element operator", nits, flush=True)
    assert nits <= n_csr + max(2, n_csr // 50)
    xr = x_csr.numpy()
    assert np.linalg.norm(x.numpy() - xr) / np.linalg.norm(xr) < 1e-7


@pytest.mark.parametrize("graph", [True, False])
def test_multi_on_element_operator_matches_single(hip, graph):
    pa = hip
    meshes = batch_meshes(nmesh=3, n=3, seed=3, lmin=1, lmax=3)
    sys_ = assemble_batch(meshes, p=3, seed=4, element_operator=True)
    As = [A for _, A, _ in sys_]
    bs = [b for _, _, b in sys_]
    xs = [pa.Vec(A.n) for A in As]
    nits, res = pa.pcg_multi(As, bs, xs, rtol=1e-10, btol=1e-10, use_graph=graph)
    for i, (A, b, x) in enumerate(zip(As, bs, xs)):
        x1 = pa.Vec(A.n)
        n1, _, r1 = pa.pcg_reliable(A, b, x1, rtol=1e-10, btol=1e-10)
        assert nits[i] == n1 and res[i] == r1
        assert np.array_equal(x.numpy(), x1.numpy())


def test_element_operator_rejects_other_problems(hip):
    pa = hip
    c = elasticity_case(2)
    sp = pa.Space(c["grid"], c["p"], bdry=c["bdry"])
    pb = pa.Problem.elasticity()
    A = pa.assemble(sp, pb, f=c["f"], g=c["g"])
    A.eliminate()
    with pytest.raises(pa.PhgHipError):
        A.element_operator(sp, pb)
    c, pb, sp, A = _system(pa, "sin", 2, 2)
    with pytest.raises(pa.PhgHipError):
        A.use_element_operator(True)    # not built on this matrix


def test_layout_shared_by_matrices_of_one_space(hip):
    """The block layout depends on the mesh, its Dirichlet rows and the
    solver numbering, not on kappa: a second matrix assembled on the same
    space (another kappa) takes the space's layout and only computes its
    coefficients.  Its operator equals, bit for bit, the one built from
    scratch on a fresh space, and matches its own FP64 matrix."""
    pa = hip
    c = poisson_case(8, 2, "kappa")
    pb = pa.Problem.poisson(use_kappa=True)
    sp = pa.Space(c["grid"], 2, bdry=c["bdry"])
    kappa2 = np.exp(np.random.default_rng(99).normal(0.0, 0.5, c["grid"].nelem))
    A1 = pa.assemble(sp, pb, kappa=c["kappa"], f=c["f"], g=c["g"])
    A1.eliminate()
    A1.element_operator(sp, pb, kappa=c["kappa"])
    A2 = pa.assemble(sp, pb, kappa=kappa2, f=c["f"], g=c["g"])
    A2.eliminate()
    A2.element_operator(sp, pb, kappa=kappa2)          # the space's layout
    sp3 = pa.Space(c["grid"], 2, bdry=c["bdry"])
    A3 = pa.assemble(sp3, pb, kappa=kappa2, f=c["f"], g=c["g"])
    A3.eliminate()
    A3.element_operator(sp3, pb, kappa=kappa2)         # built from scratch
    assert A2.element_layout() == A1.element_layout() == A3.element_layout()
    xv = np.random.default_rng(5).standard_normal(A2.n)
    x, y64, y2, y3 = pa.Vec(A2.n, xv), pa.Vec(A2.n), pa.Vec(A2.n), pa.Vec(A2.n)
    A2.matvec(x, y64)
    A2.element_apply(x, y2)
    A3.element_apply(x, y3)
    assert np.array_equal(y2.numpy(), y3.numpy())
    a = y64.numpy()
    assert np.linalg.norm(y2.numpy() - a) / np.linalg.norm(a) < 2e-6
    del A1                                              # the layout stays with A2
    z = pa.Vec(A2.n)
    A2.element_apply(x, z)
    assert np.array_equal(z.numpy(), y2.numpy())


@pytest.mark.parametrize("graph", [False, True])
def test_without_shared_row_codes(hip, monkeypatch, graph):
    """With the per-row codes switched off (PHG_HIP_EOP_NO_ROW_CODES) the
    shared rows are summed by the fix-up kernel in every iteration slot
    instead of by the solver's update kernel, in the same block order, so
    the reliable and the batched (graph-replayed) solves give the same
    solutions bit for bit; the fix-up kernel returns at once in the slots
    that are not iteration slots and after convergence."""
    pa = hip
    meshes = batch_meshes(nmesh=2, n=8, seed=3, lmin=1, lmax=2)   # several blocks each
    runs = []
    for off in (False, True):
        if off:
            monkeypatch.setenv("PHG_HIP_EOP_NO_ROW_CODES", "1")
        sys_ = assemble_batch(meshes, p=2, seed=4, element_operator=True)
        As = [A for _, A, _ in sys_]
        bs = [b for _, _, b in sys_]
        xs = [pa.Vec(A.n) for A in As]
        nits, res = pa.pcg_multi(As, bs, xs, rtol=1e-10, btol=1e-10, use_graph=graph)
        x1 = pa.Vec(As[0].n)
        n1, _, r1 = pa.pcg_reliable(As[0], bs[0], x1, rtol=1e-10, btol=1e-10)
        runs.append((list(nits), [x.numpy() for x in xs], n1, x1.numpy()))
        monkeypatch.delenv("PHG_HIP_EOP_NO_ROW_CODES", raising=False)
    (na, xa, n1a, x1a), (nb, xb, n1b, x1b) = runs
    assert na == nb and n1a == n1b
    assert all(np.array_equal(u, v) for u, v in zip(xa, xb))
    assert np.array_equal(x1a, x1b)


# (order, cubes per side, refined): meshes from test/cube.dat bisected, and
# one locally refined P3 mesh of the batch workload
LAYOUT_CASES = [(1, 16, False), (2, 16, False), (2, 32, False), (3, 8, False), (3, 8, True)]


@pytest.mark.parametrize("p,n,refined", LAYOUT_CASES)
def test_device_layout_matches_host_layout(hip, monkeypatch, p, n, refined):
    """The layout built on the device (blocks of a fixed number of elements
    in Morton order) against the host's greedy split.  Where the host cuts
    the same blocks -- no block reaches the row limit -- both layouts hold
    the same tables, so the operator is the same bit for bit; elsewhere the
    device falls back to the host split, or its own blocks give an operator
    within FP32 accuracy of the FP64 matrix."""
    pa = hip
    g = batch_meshes(nmesh=1, n=n, seed=4, lmin=3, lmax=3)[0][0] if refined \
        else pa.Grid.phg_cube(n)
    bd = g.dof_bdry(p)
    kappa = np.exp(np.random.default_rng(3).normal(0.0, 0.5, g.nelem))
    pb = pa.Problem.poisson(use_kappa=True)
    ys, lays = [], []
    for host in (False, True):
        if host:
            monkeypatch.setenv("PHG_HIP_EOP_HOST_LAYOUT", "1")
        sp = pa.Space(g, p, bdry=bd)
        A = pa.assemble(sp, pb, kappa=kappa, f=np.ones(g.ndof(p)), g=np.zeros(g.ndof(p)))
        A.eliminate()
        A.element_operator(sp, pb, kappa=kappa)
        lay = A.element_layout()
        if host:
            assert lay["device"] == 0
        xv = np.random.default_rng(5).standard_normal(A.n)
        x, y64, y = pa.Vec(A.n, xv), pa.Vec(A.n), pa.Vec(A.n)
        A.matvec(x, y64)
        A.element_apply(x, y)
        a, b = y64.numpy(), y.numpy()
        assert np.linalg.norm(b - a) / np.linalg.norm(a) < 2e-6, (host, lay)
        ys.append(b)
        lays.append(lay)
        del A, sp
    monkeypatch.delenv("PHG_HIP_EOP_HOST_LAYOUT")
    dev, hst = lays
    same = {k: v for k, v in dev.items() if k != "device"} == \
        {k: v for k, v in hst.items() if k != "device"}
    if not refined:
        assert dev["device"] == 1, dev
    if dev["device"] == 1 and same:
        assert np.array_equal(ys[0], ys[1])
