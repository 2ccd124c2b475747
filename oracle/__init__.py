"""Oracle: CPU restatement of PHG's finite-element hot path.

TEST INFRASTRUCTURE ONLY -- the checker for the HIP path.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
this package.  The product (``phg_amd``) never imports, links or calls it.

The C library (``phg_oracle.c``) follows the reference files cited in
``phg_oracle.h`` with the same floating-point operation order.  Build it with
``make -C oracle`` (also done by ``__graft_entry__.build()``).

Pinning (how each function is checked, see tests/test_oracle.py):
  quad3d        closed-form simplex moments (test/quad_test.c:129-153) and
                sum of weights = 1
  bas/grad      nodal delta property, partition of unity, finite-difference
                gradients
  jacobian      J * [x;1] = I and volume = |det|/6
  element       closed-form Kuhn P1 stiffness, P1/P2 mass matrices,
                f == 1 loads (= integral of the basis), brute-force
                high-order quadrature of the bilinear forms on one
                tetrahedron; elasticity vs a direct strain-energy evaluation
  assemble      dense brute-force assembly on tiny meshes; PHG insertion
                order checked against a Python restatement of
                add_columns_to_row
  eliminate     u = x+y+z reproduced to round-off by P1 (examples/functions.h
                TEST_CASE 1)
  pcg           test/solver_test0.c arrowhead system (x = 1),
                test/mat_test1.c A(I,J) = 1+I+J product, scipy cg
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"oracle library missing: {LIB_PATH}; run `make -C oracle`")

lib = C.CDLL(LIB_PATH)
_i32, _i64, _dbl = C.c_int32, C.c_int64, C.c_double
_P = C.POINTER


class Problem(C.Structure):
    _fields_ = [("p", C.c_int), ("dim", C.c_int), ("use_kappa", C.c_int),
                ("mass_coef", _dbl), ("rhs_order", C.c_int), ("E", _dbl), ("nu", _dbl),
                ("qcache", C.c_int)]


class _Csr(C.Structure):
    _fields_ = [("n", _i64), ("nnz", _i64), ("rowptr", _P(_i64)), ("cols", _P(_i32)),
                ("vals", _P(_dbl)), ("rhs", _P(_dbl))]


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("orc_quad3d", C.c_int, [C.c_int, _P(_P(_dbl)), _P(_P(_dbl))])
_sig("orc_bas", None, [C.c_int, _P(_dbl), _P(_dbl)])
_sig("orc_grad_lambda", None, [C.c_int, _P(_dbl), _P(_dbl)])
_sig("orc_jacobian", None, [_P(_dbl), _P(_dbl)])
_sig("orc_volume", _dbl, [_P(_dbl)])
_sig("orc_element", None, [_P(Problem), _P(_dbl), _dbl, _P(_dbl), _P(_dbl), _P(_dbl)])
_sig("orc_assemble", _P(_Csr), [_P(Problem), _i64, _P(_dbl), _P(_i32), _P(_i32), _i64,
                                _P(C.c_uint8), _P(_dbl), _P(_dbl), _P(_dbl)])
_sig("orc_eliminate", None, [_P(_Csr), _P(C.c_uint8)])
_sig("orc_pcg", C.c_int, [_P(_Csr), _P(_dbl), _P(_dbl), C.c_int, _dbl, _dbl, _dbl,
                          _P(_dbl)])
_sig("orc_bicgstab", C.c_int, [_P(_Csr), _P(_dbl), _P(_dbl), C.c_int, _dbl, _dbl, _dbl,
                               _P(_dbl)])
_sig("orc_gmres", C.c_int, [C.c_int, _P(_Csr), _P(_dbl), _P(_dbl), C.c_int, C.c_int, C.c_int,
                            _dbl, _dbl, _dbl, _P(_dbl)])
_sig("orc_matvec", None, [_P(_Csr), _P(_dbl), _P(_dbl)])
_sig("orc_csr_free", None, [_P(_Csr)])
_sig("orc_mass_apply", _dbl, [C.c_int, _i64, _P(_dbl), _P(_i32), _P(_i32), _i64,
                              _P(_dbl), _P(_dbl)])
_sig("orc_adjoint_grad", None, [C.c_int, _i64, _P(_dbl), _P(_i32), _P(_i32), _P(_dbl),
                                _P(_dbl), _P(_dbl), _P(_dbl)])


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(_P(t))


def nbas(p):
    return {1: 4, 2: 10, 3: 20, 4: 35}[p]


def problem(p, dim=1, use_kappa=False, mass_coef=0.0, rhs_order=-1, E=2.5, nu=0.25,
            qcache=False):
    return Problem(p, dim, int(use_kappa), float(mass_coef), int(rhs_order), E, nu,
                   int(qcache))


def quad3d(order):
    pp, ww = _P(_dbl)(), _P(_dbl)()
    n = lib.orc_quad3d(order, C.byref(pp), C.byref(ww))
    if n < 0:
        raise ValueError(order)
    return (np.ctypeslib.as_array(pp, (n, 4)).copy(),
            np.ctypeslib.as_array(ww, (n,)).copy())


def bas(p, lam):
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    out = np.empty(nbas(p))
    lib.orc_bas(p, _ptr(lam, _dbl), _ptr(out, _dbl))
    return out


def grad_lambda(p, lam):
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    out = np.empty((nbas(p), 4))
    lib.orc_grad_lambda(p, _ptr(lam, _dbl), _ptr(out, _dbl))
    return out


def jacobian(x):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(4, 3)
    J = np.empty((4, 4))
    lib.orc_jacobian(_ptr(x, _dbl), _ptr(J, _dbl))
    return J


def volume(x):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(4, 3)
    return lib.orc_volume(_ptr(x, _dbl))


def element(pb, x, kappa=0.0, f_loc=None):
    M = nbas(pb.p) * pb.dim
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(4, 3)
    f_loc = np.zeros(M) if f_loc is None else np.ascontiguousarray(f_loc, dtype=np.float64)
    K = np.empty((M, M))
    b = np.empty(M)
    lib.orc_element(C.byref(pb), _ptr(x, _dbl), float(kappa), _ptr(f_loc, _dbl),
                    _ptr(K, _dbl), _ptr(b, _dbl))
    return K, b


class Csr:
    """Owned copy of an oracle CSR (rowptr, cols, vals, rhs)."""

    def __init__(self, ptr):
        s = ptr.contents
        n, nnz = s.n, s.nnz
        self.n = n
        self.rowptr = np.ctypeslib.as_array(s.rowptr, (n + 1,)).copy()
        self.cols = np.ctypeslib.as_array(s.cols, (max(nnz, 1),))[:nnz].copy()
        self.vals = np.ctypeslib.as_array(s.vals, (max(nnz, 1),))[:nnz].copy()
        self.rhs = np.ctypeslib.as_array(s.rhs, (n,)).copy()
        self._p = ptr

    def refresh(self):
        self.__init__(self._p)

    def __del__(self):
        if getattr(self, "_p", None):
            lib.orc_csr_free(self._p)
            self._p = None

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.vals, self.cols, self.rowptr), shape=(self.n, self.n))


def assemble(pb, verts, elem_verts, dof_map, ndof, bdry, g, f, kappa=None):
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    ev = np.ascontiguousarray(elem_verts, dtype=np.int32)
    m = np.ascontiguousarray(dof_map, dtype=np.int32)
    bd = np.ascontiguousarray(bdry, dtype=np.uint8)
    g = np.ascontiguousarray(g, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    kap = None if kappa is None else np.ascontiguousarray(kappa, dtype=np.float64)
    p = lib.orc_assemble(C.byref(pb), ev.shape[0], _ptr(verts, _dbl), _ptr(ev, _i32),
                         _ptr(m, _i32), ndof, _ptr(bd, C.c_uint8), _ptr(g, _dbl),
                         _ptr(f, _dbl), _ptr(kap, _dbl))
    return Csr(p)


def eliminate(A: Csr, bdry_rows):
    b = np.ascontiguousarray(bdry_rows, dtype=np.uint8)
    lib.orc_eliminate(A._p, _ptr(b, C.c_uint8))
    A.refresh()
    return A


def _arg(A: Csr):
    """The C struct of an oracle CSR or of one built by csr_from_arrays."""
    return A._p if A._p is not None else C.byref(A._s)


def pcg(A: Csr, b, x0=None, maxit=10000, rtol=1e-10, btol=1e-10, atol=0.0):
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64)
    res = _dbl(1.0)     # PCG_Solve's initial value (solver-phg.c:2433)
    nits = lib.orc_pcg(_arg(A), _ptr(b, _dbl), _ptr(x, _dbl), maxit, rtol, btol, atol,
                       C.byref(res))
    return x, nits, res.value


def matvec(A: Csr, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(A.n)
    lib.orc_matvec(_arg(A), _ptr(x, _dbl), _ptr(y, _dbl))
    return y


def bicgstab(A: Csr, b, x0=None, maxit=10000, rtol=1e-10, btol=1e-10, atol=0.0):
    """solver-phg.c:607 solver_BiCGstab (Jacobi PC) on an oracle CSR or one
    built by csr_from_arrays."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64)
    res = _dbl()
    p = A._p if A._p is not None else C.byref(A._s)
    nits = lib.orc_bicgstab(p, _ptr(b, _dbl), _ptr(x, _dbl), maxit, rtol, btol, atol,
                            C.byref(res))
    return x, nits, res.value


GMRES_TYPES = ("gmres_l", "gmres_r", "fgmres_r", "lgmres_l", "lgmres_r", "flgmres_r")


def gmres(A: Csr, b, x0=None, restart=10, augk=2, maxit=10000, rtol=1e-10, btol=1e-10,
          atol=0.0, gmres_type="flgmres_r"):
    """solver-phg.c:59-77 gmres_funcs with the Jacobi PC; gmres_type is one of
    GMRES_TYPES (PHG's -gmres_type names, default flgmres_r)."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(A.n) if x0 is None else np.array(x0, dtype=np.float64)
    res = _dbl()
    nits = lib.orc_gmres(GMRES_TYPES.index(gmres_type), _arg(A), _ptr(b, _dbl),
                         _ptr(x, _dbl), restart, augk, maxit, rtol, btol, atol,
                         C.byref(res))
    return x, nits, res.value


def csr_from_arrays(rowptr, cols, vals, rhs=None):
    """Builds an oracle CSR from plain arrays (for known-answer systems)."""
    n = len(rowptr) - 1
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    ci = np.ascontiguousarray(cols, dtype=np.int32)
    va = np.ascontiguousarray(vals, dtype=np.float64)
    rh = np.zeros(n) if rhs is None else np.ascontiguousarray(rhs, dtype=np.float64)
    obj = Csr.__new__(Csr)
    obj.n, obj.rowptr, obj.cols, obj.vals, obj.rhs = n, rp, ci, va, rh
    s = _Csr(n, len(va), _ptr(rp, _i64), _ptr(ci, _i32), _ptr(va, _dbl), _ptr(rh, _dbl))
    obj._keep = (rp, ci, va, rh, s)
    obj._p = None
    obj._s = s
    return obj


def pcg_arrays(rowptr, cols, vals, b, maxit=10000, rtol=1e-10, btol=1e-10, atol=0.0):
    A = csr_from_arrays(rowptr, cols, vals)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(A.n)
    res = _dbl(1.0)     # PCG_Solve's initial value (solver-phg.c:2433)
    nits = lib.orc_pcg(C.byref(A._s), _ptr(b, _dbl), _ptr(x, _dbl), maxit, rtol, btol,
                       atol, C.byref(res))
    return x, nits, res.value


def mass_apply(p, verts, elem_verts, dof_map, e):
    """(1/2 e^T M e, M e) for the Pk mass matrix, element order."""
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    ev = np.ascontiguousarray(elem_verts, dtype=np.int32)
    m = np.ascontiguousarray(dof_map, dtype=np.int32)
    e = np.ascontiguousarray(e, dtype=np.float64)
    Me = np.empty_like(e)
    J = lib.orc_mass_apply(p, ev.shape[0], _ptr(verts, _dbl), _ptr(ev, _i32),
                           _ptr(m, _i32), len(e), _ptr(e, _dbl), _ptr(Me, _dbl))
    return J, Me


def adjoint_grad(p, verts, elem_verts, dof_map, kappa, u, lam):
    """dJ/dlog(kappa_e) = -kappa_e lambda_e^T K1_e u_e per element."""
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    ev = np.ascontiguousarray(elem_verts, dtype=np.int32)
    m = np.ascontiguousarray(dof_map, dtype=np.int32)
    kap = np.ascontiguousarray(kappa, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    g = np.empty(ev.shape[0])
    lib.orc_adjoint_grad(p, ev.shape[0], _ptr(verts, _dbl), _ptr(ev, _i32), _ptr(m, _i32),
                         _ptr(kap, _dbl), _ptr(u, _dbl), _ptr(lam, _dbl), _ptr(g, _dbl))
    return g
