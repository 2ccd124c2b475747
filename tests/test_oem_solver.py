"""PHG's solver plugin interface (include/phg_oem_hip.h) on the GPU.

The tests act as PHG: they hand the OEM_SOLVER vtables PHG-shaped MAT / VEC
/ SOLVER objects and call them in the order solver.c does -- phgSolverAssemble
(Create, AddMatrixEntries per row, Assemble), phgSolverUpdateRHS, then
Solve -- directly through the vtable and through the serial driver
stand-ins, and compare with the oracle's PCG / BiCGstab / GMRES
(solver-phg.c:485, :607, :2063).  The matrices are the oracle's assembled
and eliminated CSR in PHG's packed form (what phgMatAssemble + phgMatPack
leave), or a matrix-free MAT around the device assembly (phgMatCreateHip)."""
from __future__ import annotations

import ctypes as C
import subprocess
import sys
import textwrap

import numpy as np
import pytest

import oracle as orc
from tests import _phg_types as T
from tests._cases import poisson_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oem(hip):
    return T.load()


class PackedMat:
    """A PHG MAT in packed form (matvec.h:102-115) over host arrays kept
    alive here, with its row map and boundary rows (map.h:115)."""

    def __init__(self, rowptr, cols, vals, bdry_rows):
        n = len(rowptr) - 1
        self.ind = np.ascontiguousarray(rowptr, dtype=np.uintp)
        self.cols = np.ascontiguousarray(cols, dtype=np.int32)
        self.vals = np.ascontiguousarray(vals, dtype=np.float64)
        self.bdry = np.ascontiguousarray(np.flatnonzero(bdry_rows), dtype=np.int32)
        self.part = np.array([0, n], dtype=np.int32)
        m = T.MAP()
        m.nlocal = m.nglobal = m.localsize = n
        m.partition = self.part.ctypes.data_as(C.POINTER(T.INT))
        m.bdry = self.bdry.ctypes.data_as(C.POINTER(T.INT))
        m.bdry_nlocal = m.bdry_localsize = m.bdry_nglobal = len(self.bdry)
        m.nprocs = 1
        self.map = m
        a = T.MAT()
        a.rmap = a.cmap = C.pointer(self.map)
        a.type = T.PHG_PACKED
        a.packed_ind = self.ind.ctypes.data_as(C.POINTER(C.c_size_t))
        a.packed_cols = self.cols.ctypes.data_as(C.POINTER(T.INT))
        a.packed_data = self.vals.ctypes.data_as(C.POINTER(T.FLOAT))
        a.nnz_d = len(self.cols)
        a.localsize = n
        a.handle_bdry_eqns = 1
        a.assembled = 1
        a.mode = 1
        self.mat = a
        self.n = n

    @property
    def ptr(self):
        return C.pointer(self.mat)


def _vec_np(v, n):
    return np.ctypeslib.as_array(v.contents.data, (n,))


def _system(n=5, p=2, kind="kappa"):
    c = poisson_case(n, p, kind)
    Ao = orc.assemble(orc.problem(p, use_kappa=(kind == "kappa")), c["verts"],
                      c["elem_verts"], c["map"], c["ndof"], c["bdry"], c["g"], c["f"],
                      c["kappa"])
    orc.eliminate(Ao, c["bdry"])
    return c, Ao


def test_hip_pcg_through_the_driver(oem):
    """phgSolverMat2Solver -> phgSolverVecSolve (Assemble, UpdateRHS, Solve)
    -> phgSolverDestroy with SOLVER_HIP_PCG: nits, residual and solution as
    the oracle's PCG with PHG's default tolerances (rtol = btol = 1e-10)."""
    c, Ao = _system()
    M = PackedMat(Ao.rowptr, Ao.cols, Ao.vals, c["bdry"])
    s = oem.phgSolverMat2Solver(C.addressof(T.vtable(oem, "phgSolverHipPCG_")), M.ptr)
    assert s.contents.rtol == 1e-10 and s.contents.maxit == 10000
    assert M.mat.refcount == 1                      # solver.c:879
    _vec_np(s.contents.rhs, M.n)[:] = Ao.rhs        # lifted by the oracle's elimination
    x = oem.phgMapCreateVec(C.pointer(M.map), 1)
    nits = oem.phgSolverVecSolve(s, 0, x)
    xo, nits_o, res_o = orc.pcg(Ao, Ao.rhs, rtol=1e-10, btol=1e-10)
    xv = _vec_np(x, M.n).copy()
    print(f"hip_pcg: {nits} its (oracle {nits_o}), residual {s.contents.residual:.3e} "
          f"(oracle {res_o:.3e})", flush=True)
    assert nits == s.contents.nits and abs(nits - nits_o) <= 2
    assert np.linalg.norm(xv - xo) / np.linalg.norm(xo) < 1e-9
    assert abs(s.contents.residual - res_o) <= 1e-3 * res_o
    assert s.contents.assembled and s.contents.sent_to_oem and s.contents.rhs_updated
    # a second solve reuses the device matrix; b = 0 gives x = 0, 0 its and
    # residual 1.0 (PCG_Solve's initial value); destroy frees the matrix
    _vec_np(s.contents.rhs, M.n)[:] = 0.0
    _vec_np(x, M.n)[:] = 1.0
    assert oem.phgSolverVecSolve(s, 1, x) == 0
    assert s.contents.residual == 1.0 and not _vec_np(x, M.n).any()
    assert not s.contents.mat                        # phgMatDestroy(&solver->mat)
    assert M.mat.refcount == 0
    oem.phgVecDestroy(C.byref(x))
    oem.phgSolverDestroy(C.byref(s))
    assert not s


def test_vtable_in_solver_c_order(oem):
    """The vtable called directly, as solver.c's phgSolverAssemble
    (send_matrix_to_oem_solver: Create, AddMatrixEntries row by row) and
    phgSolverVecSolve (Solve) do."""
    c, Ao = _system(4, 2)
    M = PackedMat(Ao.rowptr, Ao.cols, Ao.vals, c["bdry"])
    vt = T.vtable(oem, "phgSolverHipPCG_")
    assert vt.name == b"hip_pcg" and vt.req_symmetry == 3 and vt.iterative
    assert vt.allow_matrix_free and not vt.parallel
    s = T.SOLVER()
    s.mat = M.ptr
    s.rtol, s.btol, s.atol, s.maxit = 1e-12, 1e-12, 0.0, 10000
    rhs = Ao.rhs.copy()
    v = T.VEC()
    v.data = rhs.ctypes.data_as(C.POINTER(T.FLOAT))
    v.map = C.pointer(M.map)
    s.rhs = C.pointer(v)
    sp = C.pointer(s)
    assert vt.Init(sp) == 0 and vt.Create(sp) == 0
    for i in range(M.n):
        a, b = Ao.rowptr[i], Ao.rowptr[i + 1]
        row = (T.INT * 1)(i)
        cols = np.ascontiguousarray(Ao.cols[a:b], dtype=np.int32)
        vals = np.ascontiguousarray(Ao.vals[a:b])
        assert vt.AddMatrixEntries(sp, 1, row, b - a, cols.ctypes.data_as(C.POINTER(T.INT)),
                                   vals.ctypes.data_as(C.POINTER(T.FLOAT))) == 0
    assert vt.Assemble(sp) == 0
    x0 = np.zeros(M.n)
    xv = T.VEC()
    xv.data = x0.ctypes.data_as(C.POINTER(T.FLOAT))
    xv.map = C.pointer(M.map)
    nits = vt.Solve(sp, C.pointer(xv), 0)
    xo, nits_o, _ = orc.pcg(Ao, Ao.rhs, rtol=1e-12, btol=1e-12)
    assert nits == s.nits and abs(nits - nits_o) <= 2
    assert np.linalg.norm(x0 - xo) / np.linalg.norm(xo) < 1e-9
    assert vt.Destroy(sp) == 0 and not s.oem_data


@pytest.mark.parametrize("name", ["phgSolverHipBiCGstab_", "phgSolverHipGMRES_"])
def test_bicgstab_and_gmres_vtables(oem, name):
    """hip_bicgstab and hip_gmres (PHG's flgmres_r, restart 10, augk 2) on
    a nonsymmetric system (the kappa-Poisson matrix plus a skew part), with
    the oracle's restatements of solver-phg.c:607 / :2063."""
    c, Ao = _system(3, 2)
    rng = np.random.default_rng(3)
    vals = Ao.vals.copy()
    rows = np.repeat(np.arange(Ao.n), np.diff(Ao.rowptr))
    off = (Ao.cols != rows) & ~c["bdry"].astype(bool)[rows]
    vals[off] *= 1.0 + 0.1 * rng.uniform(-1, 1, off.sum())
    A2 = orc.csr_from_arrays(Ao.rowptr, Ao.cols, vals)
    M = PackedMat(Ao.rowptr, Ao.cols, vals, c["bdry"])
    s = oem.phgSolverMat2Solver(C.addressof(T.vtable(oem, name)), M.ptr)
    _vec_np(s.contents.rhs, M.n)[:] = Ao.rhs
    x = oem.phgMapCreateVec(C.pointer(M.map), 1)
    nits = oem.phgSolverVecSolve(s, 0, x)
    if name == "phgSolverHipBiCGstab_":
        xo, nits_o, _ = orc.bicgstab(A2, Ao.rhs, rtol=1e-10, btol=1e-10)
    else:
        xo, nits_o, _ = orc.gmres(A2, Ao.rhs, restart=10, augk=2, rtol=1e-10, btol=1e-10)
    xv = _vec_np(x, M.n).copy()
    print(f"{name}: {nits} its (oracle {nits_o})", flush=True)
    assert abs(nits - nits_o) <= max(2, nits_o // 20)
    assert np.linalg.norm(xv - xo) / np.linalg.norm(xo) < 1e-8
    oem.phgVecDestroy(C.byref(x))
    oem.phgSolverDestroy(C.byref(s))


@pytest.mark.parametrize("gtype", orc.GMRES_TYPES)
def test_gmres_vtable_every_gmres_type(oem, gtype):
    """hip_gmres with each -gmres_type keyword (solver-phg.c:71) against the
    oracle's restatement of that variant; unknown keywords and other
    solvers are refused."""
    c, Ao = _system(3, 2)
    M = PackedMat(Ao.rowptr, Ao.cols, Ao.vals, c["bdry"])
    s = oem.phgSolverMat2Solver(C.addressof(T.vtable(oem, "phgSolverHipGMRES_")), M.ptr)
    assert oem.phgSolverHipSetGMRESType(s, b"gmres") == -1
    assert oem.phgSolverHipSetGMRESType(s, gtype.encode()) == 0
    _vec_np(s.contents.rhs, M.n)[:] = Ao.rhs
    x = oem.phgMapCreateVec(C.pointer(M.map), 1)
    nits = oem.phgSolverVecSolve(s, 0, x)
    xo, nits_o, res_o = orc.gmres(Ao, Ao.rhs, rtol=1e-10, btol=1e-10, gmres_type=gtype)
    xv = _vec_np(x, M.n).copy()
    print(f"{gtype}: {nits} its (oracle {nits_o}), residual {s.contents.residual:.3e} "
          f"(oracle {res_o:.3e})", flush=True)
    assert abs(nits - nits_o) <= max(2, nits_o // 20)
    assert np.linalg.norm(xv - xo) / np.linalg.norm(xo) < 1e-8
    oem.phgVecDestroy(C.byref(x))
    oem.phgSolverDestroy(C.byref(s))
    s = oem.phgSolverMat2Solver(C.addressof(T.vtable(oem, "phgSolverHipPCG_")), M.ptr)
    assert oem.phgSolverHipSetGMRESType(s, gtype.encode()) == -1
    oem.phgSolverDestroy(C.byref(s))


def test_matrix_free_device_matrix(oem, hip):
    """The element loop lowered to the GPU (phg_hip_assemble + eliminate)
    wrapped as a matrix-free PHG MAT: SOLVER_HIP_PCG solves it without host
    rows, FP64 and mixed precision, against the oracle; phgHipMatVec is
    PHG's row-order product bit for bit and phgHipJacobiPC is b ./ diag."""
    pa = hip
    c, Ao = _system(5, 3)
    sp = pa.Space(c["grid"], c["p"], bdry=c["bdry"])
    A = pa.assemble(sp, pa.Problem.poisson(use_kappa=True), kappa=c["kappa"], f=c["f"],
                    g=c["g"])
    A.eliminate()
    M = oem.phgMatCreateHip(A._h, 1)
    assert M.contents.type == T.PHG_MATRIX_FREE and M.contents.assembled
    n = A.n
    xo, nits_o, _ = orc.pcg(Ao, Ao.rhs, rtol=1e-12, btol=1e-12)
    for mixed in (0, 1):
        s = oem.phgSolverMat2Solver(C.addressof(T.vtable(oem, "phgSolverHipPCG_")), M)
        oem.phgSolverHipSetOptions(s, mixed, 0, 0, 1)
        oem.phgSolverSetTol(s, 1e-12)
        s.contents.btol = 1e-12
        _vec_np(s.contents.rhs, n)[:] = A.rhs().numpy()
        x = oem.phgMapCreateVec(M.contents.rmap, 1)
        nits = oem.phgSolverVecSolve(s, 0, x)
        rel = np.linalg.norm(_vec_np(x, n) - xo) / np.linalg.norm(xo)
        print(f"matrix-free hip_pcg mixed={mixed}: {nits} its (oracle {nits_o}), "
              f"rel. error {rel:.2e}", flush=True)
        assert rel < (1e-9 if not mixed else 1e-8)
        if not mixed:
            assert abs(nits - nits_o) <= 2
        oem.phgVecDestroy(C.byref(x))
        oem.phgSolverDestroy(C.byref(s))
    # MV_FUNC and PC_PROC
    rng = np.random.default_rng(8)
    xin = oem.phgMapCreateVec(M.contents.rmap, 1)
    yout = oem.phgMapCreateVec(M.contents.rmap, 1)
    xv = rng.standard_normal(n)
    _vec_np(xin, n)[:] = xv
    for op in (T.MAT_OP_N, T.MAT_OP_T):
        assert oem.phgHipMatVec(op, M, xin, yout) == 0
        y = _vec_np(yout, n).copy()
        assert np.array_equal(y.view(np.uint64), orc.matvec(Ao, xv).view(np.uint64))
    assert oem.phgHipMatVec(T.MAT_OP_D, M, xin, yout) == -1
    oem.phgHipJacobiPC(C.cast(M, C.c_void_p), xin, C.byref(yout))
    diag = np.array([Ao.vals[Ao.rowptr[i] + np.flatnonzero(
        Ao.cols[Ao.rowptr[i]:Ao.rowptr[i + 1]] == i)[0]] for i in range(n)])
    assert np.array_equal(_vec_np(yout, n), xv / diag)
    oem.phgVecDestroy(C.byref(xin))
    oem.phgVecDestroy(C.byref(yout))
    oem.phgMatDestroyHip(C.byref(M))


def test_zero_diagonal_is_fatal():
    """jacobi_pc (solver-phg.c:233): a zero diagonal stops the run through
    phgError, i.e. exit(1) with the message, in a child process."""
    code = textwrap.dedent("""
        import ctypes as C, numpy as np, sys
        sys.path.insert(0, %r)
        import phg_amd as pa
        from tests import _phg_types as T
        from tests.test_oem_solver import PackedMat, _vec_np
        pa.init(0)
        oem = T.load()
        # 3 rows, the middle one with a zero diagonal entry
        rp = [0, 2, 4, 6]; ci = [0, 1, 1, 0, 2, 1]; va = [4.0, 1.0, 0.0, 1.0, 4.0, 1.0]
        M = PackedMat(rp, ci, va, np.zeros(3, dtype=np.uint8))
        s = oem.phgSolverMat2Solver(C.addressof(T.vtable(oem, "phgSolverHipPCG_")), M.ptr)
        _vec_np(s.contents.rhs, 3)[:] = 1.0
        x = oem.phgMapCreateVec(C.pointer(M.map), 1)
        print("solving", flush=True)
        oem.phgSolverVecSolve(s, 0, x)
        print("not reached", flush=True)
    """) % T.ROOT
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       timeout=120)
    print(p.stdout, p.stderr[-400:], flush=True)
    assert p.returncode == 1
    assert "solving" in p.stdout and "not reached" not in p.stdout
    assert "zero diagonal" in p.stderr
