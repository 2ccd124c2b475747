"""phg_amd -- MI355X (gfx950) finite-element hot path of PHG.

Python bindings over the two in-tree native libraries:

* ``lib/libphg_grid.so`` -- host grid layer with PHG's import numbering
  (``phg_amd/host/phg_grid.h``);
* ``lib/libphg_hip.so``  -- HIP kernels behind the C ABI of
  ``include/phg_hip.h``.

The bindings are plumbing for drivers, tests and the benchmark: all numerical
work runs in the HIP library.  Importing the package without the built HIP
library raises ``ImportError`` -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "lib")

__all__ = [
    "Grid", "Space", "Mat", "Vec", "Problem", "init", "finalize",
    "assemble", "element_matrices", "pcg", "bicgstab", "gmres", "pcg_mixed", "pcg_reliable", "pcg_block", "pcg_multi", "mat_from_csr", "dof_norms", "spmm_tune", "adjoint_logkappa", "PhgHipError",
    "hip_lib", "grid_lib", "HIP_LIB_PATH", "GRID_LIB_PATH",
]

HIP_LIB_PATH = os.path.join(_LIB, "libphg_hip.so")
GRID_LIB_PATH = os.path.join(_LIB, "libphg_grid.so")


def _load(path: str) -> C.CDLL:
    if not os.path.exists(path):
        raise ImportError(
            f"phg_amd native library missing: {path}. Build it with "
            "`python phg_amd/build.py` (or __graft_entry__.build()).")
    return C.CDLL(path)


grid_lib = _load(GRID_LIB_PATH)
hip_lib = _load(HIP_LIB_PATH)

_i32, _i64, _dbl, _vp, _cp = C.c_int32, C.c_int64, C.c_double, C.c_void_p, C.c_char_p
_P = C.POINTER


class _GridStruct(C.Structure):
    _fields_ = [
        ("nvert", _i32), ("nelem", _i32), ("nedge", _i32), ("nface", _i32),
        ("verts", _P(_dbl)), ("elem_verts", _P(_i32)), ("elem_edges", _P(_i32)),
        ("elem_faces", _P(_i32)), ("elem_neigh", _P(_i32)),
        ("elem_bound", _P(C.c_uint8)), ("types_vert", _P(C.c_uint8)),
        ("types_edge", _P(C.c_uint8)), ("types_face", _P(C.c_uint8)),
        ("elem_type", _P(C.c_uint8)), ("trav", _P(_i32)),
    ]


class Problem(C.Structure):
    """Integrand selector (include/phg_hip.h phg_hip_problem)."""
    _fields_ = [("dim", _i32), ("use_kappa", _i32), ("mass_coef", _dbl),
                ("rhs_order", _i32), ("E", _dbl), ("nu", _dbl), ("qcache", _i32)]

    @classmethod
    def poisson(cls, use_kappa=False, mass_coef=0.0, rhs_order=-1, qcache=False):
        """qcache: the integrals of examples/poisson.c -qcache (quad-cache.c)."""
        return cls(1, int(use_kappa), float(mass_coef), int(rhs_order), 0.0, 0.0,
                   int(qcache))

    @classmethod
    def elasticity(cls, E=2.5, nu=0.25, rhs_order=-1):
        return cls(3, 0, 0.0, int(rhs_order), float(E), float(nu), 0)


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


# grid library
_sig(grid_lib, "phg_grid_new_kuhn", _P(_GridStruct), [_i32, _P(C.c_int)])
_sig(grid_lib, "phg_grid_read_albert", _P(_GridStruct), [_cp])
_sig(grid_lib, "phg_grid_write_albert", C.c_int, [_P(_GridStruct), _cp])
_sig(grid_lib, "phg_grid_bisect", _P(_GridStruct), [_P(_GridStruct), _P(_i32)])
_sig(grid_lib, "phg_grid_mark", C.c_int,
     [_P(_GridStruct), C.c_int, _P(_dbl), _dbl, _dbl, C.c_int, _P(_i32)])
MARK_NONE, MARK_ALL, MARK_MAX, MARK_GERS, MARK_EQDIST = range(5)
_sig(grid_lib, "phg_grid_free", None, [_P(_GridStruct)])
_sig(grid_lib, "phg_dof_nbas", C.c_int, [C.c_int])
_sig(grid_lib, "phg_dof_count", _i64, [_P(_GridStruct), C.c_int])
_sig(grid_lib, "phg_dof_map", None, [_P(_GridStruct), C.c_int, _P(_i32)])
_sig(grid_lib, "phg_dof_coords", None, [_P(_GridStruct), C.c_int, _P(_dbl)])
_sig(grid_lib, "phg_dof_bdry", None, [_P(_GridStruct), C.c_int, C.c_uint,
                                      _P(C.c_uint8)])

# HIP library
_sig(hip_lib, "phg_hip_init", C.c_int, [_i32])
_sig(hip_lib, "phg_hip_finalize", C.c_int, [])
_sig(hip_lib, "phg_hip_last_error", _cp, [])
_sig(hip_lib, "phg_hip_synchronize", C.c_int, [])
_sig(hip_lib, "phg_hip_space_create", C.c_int,
     [_i32, _i64, _P(_dbl), _i64, _P(_i32), _P(_i32), _i64, _P(C.c_uint8), _P(_vp)])
_sig(hip_lib, "phg_hip_space_destroy", C.c_int, [_vp])
_sig(hip_lib, "phg_hip_element_matrices", C.c_int,
     [_vp, _P(Problem), _P(_dbl), _P(_dbl), _P(_dbl), _P(_dbl)])
_sig(hip_lib, "phg_hip_assemble", C.c_int,
     [_vp, _P(Problem), _P(_dbl), _P(_dbl), _P(_dbl), _P(_vp)])
_sig(hip_lib, "phg_hip_mat_eliminate", C.c_int, [_vp])
_sig(hip_lib, "phg_hip_mat_from_csr", C.c_int,
     [_i64, _P(_i64), _P(_i32), _P(_dbl), _P(_dbl), _P(C.c_uint8), _i32, _P(_vp)])
_sig(hip_lib, "phg_hip_dof_norms", C.c_int, [_vp, _vp, _P(_dbl), _P(_dbl)])
_sig(hip_lib, "phg_hip_adjoint_logkappa", C.c_int,
     [_vp, _P(Problem), _P(_dbl), _vp, _vp, _vp, _i32, _dbl, _P(_dbl), _P(_dbl), _vp,
      _P(_i32)])
_sig(hip_lib, "phg_hip_mat_info", C.c_int, [_vp, _P(_i64), _P(_i64)])
_sig(hip_lib, "phg_hip_mat_download", C.c_int,
     [_vp, _P(_i64), _P(_i32), _P(_dbl), _P(_dbl)])
_sig(hip_lib, "phg_hip_mat_destroy", C.c_int, [_vp])
_sig(hip_lib, "phg_hip_mat_dump_csr", C.c_int, [_vp, _cp])
_sig(hip_lib, "phg_hip_vec_dump", C.c_int, [_vp, _cp])
_sig(hip_lib, "phg_hip_vec_create", C.c_int, [_i64, _P(_vp)])
_sig(hip_lib, "phg_hip_vec_destroy", C.c_int, [_vp])
_sig(hip_lib, "phg_hip_vec_upload", C.c_int, [_vp, _P(_dbl)])
_sig(hip_lib, "phg_hip_vec_download", C.c_int, [_vp, _P(_dbl)])
_sig(hip_lib, "phg_hip_mat_rhs", C.c_int, [_vp, _vp])
_sig(hip_lib, "phg_hip_matvec", C.c_int, [_dbl, _vp, _vp, _dbl, _vp])
_sig(hip_lib, "phg_hip_vec_dot", C.c_int, [_vp, _vp, _P(_dbl)])
_sig(hip_lib, "phg_hip_vec_norm2", C.c_int, [_vp, _P(_dbl)])
_sig(hip_lib, "phg_hip_vec_axpby", C.c_int, [_dbl, _vp, _dbl, _vp])
_sig(hip_lib, "phg_hip_vec_copy", C.c_int, [_vp, _vp])
_sig(hip_lib, "phg_hip_pc_jacobi", C.c_int, [_vp, _vp, _vp])
_sig(hip_lib, "phg_hip_pcg", C.c_int,
     [_vp, _vp, _vp, _i32, _dbl, _dbl, _dbl, _P(_dbl)])
_sig(hip_lib, "phg_hip_bicgstab", C.c_int,
     [_vp, _vp, _vp, _i32, _dbl, _dbl, _dbl, _P(_dbl)])
_sig(hip_lib, "phg_hip_gmres", C.c_int,
     [_vp, _vp, _vp, _i32, _i32, _i32, _dbl, _dbl, _dbl, _P(_dbl)])
_sig(hip_lib, "phg_hip_gmres_type", C.c_int,
     [_vp, _vp, _vp, _i32, _i32, _i32, _i32, _dbl, _dbl, _dbl, _P(_dbl)])
_sig(hip_lib, "phg_hip_pcg_mixed", C.c_int,
     [_vp, _vp, _vp, _i32, _dbl, _dbl, _dbl, _dbl, _P(_i32), _P(_dbl)])
_sig(hip_lib, "phg_hip_pcg_reliable", C.c_int,
     [_vp, _vp, _vp, _i32, _dbl, _dbl, _dbl, _dbl, _P(_i32), _P(_dbl)])
_sig(hip_lib, "phg_hip_pcg_block", C.c_int,
     [_vp, _i32, _P(_vp), _P(_vp), _i32, _dbl, _dbl, _dbl, _i32, _dbl, _P(_i32), _P(_dbl)])
_sig(hip_lib, "phg_hip_pcg_multi", C.c_int,
     [_i32, _P(_vp), _P(_vp), _P(_vp), _i32, _dbl, _dbl, _dbl, _dbl, _i32, _P(_i32),
      _P(_dbl)])
_sig(hip_lib, "phg_hip_batch_create", C.c_int,
     [_i32, _P(_vp), _P(_P(Problem)), _P(_vp), _P(_vp), _P(_vp), _P(_vp), _P(_vp), _i32,
      _dbl, _dbl, _dbl, _dbl, _P(_vp)])
_sig(hip_lib, "phg_hip_batch_run", C.c_int, [_vp, _P(_i32), _P(_dbl)])
_sig(hip_lib, "phg_hip_batch_destroy", C.c_int, [_vp])
_sig(hip_lib, "phg_hip_mat_element_operator", C.c_int, [_vp, _vp, _P(Problem), _P(_dbl)])
_sig(hip_lib, "phg_hip_mat_use_element_operator", C.c_int, [_vp, _i32])
_sig(hip_lib, "phg_hip_element_apply", C.c_int, [_vp, _vp, _vp])
_sig(hip_lib, "phg_hip_element_bytes", C.c_int, [_vp, _P(_i64), _P(_i64)])
_sig(hip_lib, "phg_hip_element_time", C.c_int, [_vp, _i32, _P(_dbl), _P(_dbl)])
_sig(hip_lib, "phg_hip_element_layout", C.c_int, [_vp, _P(_i64)])
_sig(hip_lib, "phg_hip_spmm_tune", C.c_int, [_vp, _i32])
_sig(hip_lib, "phg_hip_spmm_tune_time", C.c_int, [_vp, _i32, _i32, _P(_dbl)])
_sig(hip_lib, "phg_hip_spmm_bytes", C.c_int, [_vp, _i32, _P(_i64)])
_sig(hip_lib, "phg_hip_spmm_set_shape", C.c_int, [_vp, _i32, _i32])
_sig(hip_lib, "phg_hip_mat_tune", C.c_int, [_vp, _i32])
_sig(hip_lib, "phg_hip_mat_tune_time", C.c_int, [_vp, _i32, _i32, _P(_dbl)])
_sig(hip_lib, "phg_hip_mat_set_format", C.c_int, [_vp, _i32, _i32])
_sig(hip_lib, "phg_hip_mat_get_format", C.c_int, [_vp, _P(_i32), _P(_i32)])
_sig(hip_lib, "phg_hip_mat_format_bytes", C.c_int, [_vp, _i32, _i32, _P(_i64)])
_sig(hip_lib, "phg_hip_profile", C.c_int, [_i32])
_sig(hip_lib, "phg_hip_stream_copy", C.c_int, [_i64, _i32, _P(_dbl)])
_sig(hip_lib, "phg_hip_profile_read", C.c_int, [_i32, _P(_i64), _P(_dbl)])

# SpMV layouts of the solvers (include/phg_hip.h)
FMT_NAMES = {0: "csr_b2k", 1: "csr_b4k", 2: "csr_b2k_par", 3: "csr_b4k_par",
             4: "csr_wave", 5: "sell64", 6: "csr_dict",
             7: "csr_dict_np", 8: "csr_xstage", 9: "csr_xstage_pf"}

# kernel classes of phg_hip_profile_read (include/phg_hip.h)
K_PCG_P64, K_SPMV64, K_UPDATE64, K_PCG_P32, K_SPMV32, K_UPDATE32 = range(6)


class PhgHipError(RuntimeError):
    def __init__(self, code, where):
        msg = hip_lib.phg_hip_last_error().decode(errors="replace")
        super().__init__(f"{where}: error {code}: {msg}")
        self.code = code


def _check(rc, where):
    if rc < 0:
        raise PhgHipError(rc, where)
    return rc


def _ptr(a, ctype):
    return a.ctypes.data_as(_P(ctype)) if a is not None else None


def init(device: int = 0) -> None:
    _check(hip_lib.phg_hip_init(device), "phg_hip_init")


def finalize() -> None:
    _check(hip_lib.phg_hip_finalize(), "phg_hip_finalize")


def synchronize() -> None:
    _check(hip_lib.phg_hip_synchronize(), "phg_hip_synchronize")


class Grid:
    """Host grid with PHG numbering (phg_amd/host/phg_grid.h)."""

    def __init__(self, ptr):
        if not ptr:
            raise RuntimeError("grid construction failed")
        self._p = ptr

    @classmethod
    def kuhn(cls, n: int, bcode=(1, 1, 1, 1, 1, 1)):
        arr = (C.c_int * 6)(*bcode)
        return cls(grid_lib.phg_grid_new_kuhn(n, arr))

    @classmethod
    def read_albert(cls, path: str):
        return cls(grid_lib.phg_grid_read_albert(path.encode()))

    @classmethod
    def phg_cube(cls, n: int, bcode=(1, 1, 1, 1, 1, 1)):
        """The n^3 x 6 mesh as PHG builds it: test/cube.dat (6 Kuhn
        tetrahedra, which kuhn(1) reproduces with its boundary codes and
        types) bisected 3 log2(n) times by one phgRefineAllElements call, so
        vertices, elements, edges and faces carry PHG's numbering.  n a power
        of 2."""
        lev = int(n).bit_length() - 1
        if n < 1 or (1 << lev) != n:
            raise ValueError("phg_cube: n must be a power of 2")
        g = cls.kuhn(1, bcode)
        return g.refine_all(3 * lev) if lev > 0 else g

    def refine_toward(self, point, levels: int, radius: float = 0.45,
                      shrink: float = 1.0) -> "Grid":
        """Local refinement toward `point` with PHG's bisection: sweep l marks
        (depth 1) every element whose centroid lies within radius * shrink^l
        of the point or which contains it, and refines them with PHG's
        conforming closure (phg_grid_bisect)."""
        pt = np.asarray(point, dtype=np.float64)
        g = self
        for lev in range(int(levels)):
            x = g.verts()[g.elem_verts()]
            near = np.linalg.norm(x.mean(axis=1) - pt, axis=1) < radius * shrink ** lev
            X = np.concatenate([x, np.ones((g.nelem, 4, 1))], axis=2)
            rhs = np.broadcast_to(np.append(pt, 1.0)[:, None], (g.nelem, 4, 1))
            lam = np.linalg.solve(np.transpose(X, (0, 2, 1)), rhs)[..., 0]
            inside = lam.min(axis=1) >= -1e-12
            g = g.bisect((near | inside).astype(np.int32))
        return g

    def bisect(self, marks) -> "Grid":
        """PHG's bisection (refine.c:2346 phgRefineMarkedElements): every
        element e bisected at least marks[e] times, with PHG's closure,
        vertex creation order and index inheritance (phg_grid_bisect)."""
        m = np.ascontiguousarray(np.broadcast_to(marks, (self.nelem,)), dtype=np.int32)
        p = grid_lib.phg_grid_bisect(self._p, _ptr(m, _i32))
        if not p:
            raise RuntimeError("phg_grid_bisect failed")
        return Grid(p)

    def mark(self, strategy: int, err=None, theta: float = 0.5, tol: float = 1.0,
             depth: int = 1) -> np.ndarray:
        """mark.c:912 phgMarkElements with an element-wise indicator; returns
        the marks by element index (phg_grid_mark)."""
        m = np.zeros(self.nelem, dtype=np.int32)
        e = None if err is None else np.ascontiguousarray(err, dtype=np.float64)
        if e is not None and e.size != self.nelem:
            raise ValueError("one indicator per element")
        if grid_lib.phg_grid_mark(self._p, int(strategy), _ptr(e, _dbl), float(theta),
                                  float(tol), int(depth), _ptr(m, _i32)) != 0:
            raise ValueError("phg_grid_mark: bad arguments")
        return m

    def refine_all(self, level: int) -> "Grid":
        """refine.c:2541 phgRefineAllElements(g, level)."""
        return self.bisect(np.full(self.nelem, int(level), dtype=np.int32))

    def elem_type(self):
        return np.ctypeslib.as_array(self.s.elem_type, (self.nelem,)).copy()

    def traversal(self):
        """Element indices in PHG's tree traversal order (identity for a flat
        mesh)."""
        if not self.s.trav:
            return np.arange(self.nelem, dtype=np.int32)
        return np.ctypeslib.as_array(self.s.trav, (self.nelem,)).copy()

    def elem_bound(self):
        return np.ctypeslib.as_array(self.s.elem_bound, (self.nelem, 4)).copy()

    def write_albert(self, path: str) -> None:
        if grid_lib.phg_grid_write_albert(self._p, path.encode()) != 0:
            raise OSError(f"cannot write {path}")

    def __del__(self):
        if getattr(self, "_p", None):
            grid_lib.phg_grid_free(self._p)
            self._p = None

    @property
    def s(self):
        return self._p.contents

    nvert = property(lambda self: self.s.nvert)
    nelem = property(lambda self: self.s.nelem)
    nedge = property(lambda self: self.s.nedge)
    nface = property(lambda self: self.s.nface)

    def verts(self):
        return np.ctypeslib.as_array(self.s.verts, (self.nvert, 3)).copy()

    def elem_verts(self):
        return np.ctypeslib.as_array(self.s.elem_verts, (self.nelem, 4)).copy()

    def elem_edges(self):
        return np.ctypeslib.as_array(self.s.elem_edges, (self.nelem, 6)).copy()

    def elem_faces(self):
        return np.ctypeslib.as_array(self.s.elem_faces, (self.nelem, 4)).copy()

    def elem_neigh(self):
        return np.ctypeslib.as_array(self.s.elem_neigh, (self.nelem, 4)).copy()

    def types_vert(self):
        return np.ctypeslib.as_array(self.s.types_vert, (self.nvert,)).copy()

    def ndof(self, order: int) -> int:
        return int(grid_lib.phg_dof_count(self._p, order))

    def dof_map(self, order: int) -> np.ndarray:
        nb = grid_lib.phg_dof_nbas(order)
        m = np.empty((self.nelem, nb), dtype=np.int32)
        grid_lib.phg_dof_map(self._p, order, _ptr(m, _i32))
        return m

    def dof_coords(self, order: int) -> np.ndarray:
        x = np.empty((self.ndof(order), 3), dtype=np.float64)
        grid_lib.phg_dof_coords(self._p, order, _ptr(x, _dbl))
        return x

    def dof_bdry(self, order: int, mask: int = 1) -> np.ndarray:
        b = np.empty(self.ndof(order), dtype=np.uint8)
        grid_lib.phg_dof_bdry(self._p, order, mask, _ptr(b, C.c_uint8))
        return b


class Space:
    """Mesh + Lagrange Pk space uploaded to the device."""

    def __init__(self, grid: Grid, order: int, bdry_mask: int = 1,
                 bdry: np.ndarray | None = None):
        self.order = order
        self.nelem = grid.nelem
        self.verts = np.ascontiguousarray(grid.verts(), dtype=np.float64)
        self.elem_verts = np.ascontiguousarray(grid.elem_verts(), dtype=np.int32)
        self.map = grid.dof_map(order)
        self.ndof = grid.ndof(order)
        self.bdry = grid.dof_bdry(order, bdry_mask) if bdry is None else \
            np.ascontiguousarray(bdry, dtype=np.uint8)
        self.coords = grid.dof_coords(order)
        h = _vp()
        _check(hip_lib.phg_hip_space_create(
            order, grid.nvert, _ptr(self.verts, _dbl), self.nelem,
            _ptr(self.elem_verts, _i32), _ptr(self.map, _i32), self.ndof,
            _ptr(self.bdry, C.c_uint8), C.byref(h)), "phg_hip_space_create")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            hip_lib.phg_hip_space_destroy(self._h)
            self._h = None


class Vec:
    """FP64 device vector."""

    def __init__(self, n: int, data: np.ndarray | None = None):
        self.n = int(n)
        h = _vp()
        _check(hip_lib.phg_hip_vec_create(self.n, C.byref(h)), "phg_hip_vec_create")
        self._h = h
        if data is not None:
            self.upload(data)

    def upload(self, data):
        a = np.ascontiguousarray(data, dtype=np.float64)
        assert a.size == self.n
        _check(hip_lib.phg_hip_vec_upload(self._h, _ptr(a, _dbl)), "vec_upload")

    def numpy(self) -> np.ndarray:
        a = np.empty(self.n, dtype=np.float64)
        _check(hip_lib.phg_hip_vec_download(self._h, _ptr(a, _dbl)), "vec_download")
        return a

    def dump(self, fn: str) -> None:
        """phgVecDump text format (matvec.c:6245)."""
        _check(hip_lib.phg_hip_vec_dump(self._h, fn.encode()), "vec_dump")

    def dot(self, other: "Vec") -> float:
        r = _dbl()
        _check(hip_lib.phg_hip_vec_dot(self._h, other._h, C.byref(r)), "vec_dot")
        return r.value

    def norm2(self) -> float:
        r = _dbl()
        _check(hip_lib.phg_hip_vec_norm2(self._h, C.byref(r)), "vec_norm2")
        return r.value

    def axpby(self, a: float, x: "Vec", b: float) -> None:
        _check(hip_lib.phg_hip_vec_axpby(a, x._h, b, self._h), "vec_axpby")

    def copy_from(self, src: "Vec") -> None:
        _check(hip_lib.phg_hip_vec_copy(src._h, self._h), "vec_copy")

    def __del__(self):
        if getattr(self, "_h", None):
            hip_lib.phg_hip_vec_destroy(self._h)
            self._h = None


class Mat:
    """Assembled CSR system on the device."""

    def __init__(self, h):
        self._h = h

    def info(self):
        n, nnz = _i64(), _i64()
        _check(hip_lib.phg_hip_mat_info(self._h, C.byref(n), C.byref(nnz)), "mat_info")
        return n.value, nnz.value

    @property
    def n(self):
        return self.info()[0]

    @property
    def nnz(self):
        return self.info()[1]

    def eliminate(self) -> None:
        _check(hip_lib.phg_hip_mat_eliminate(self._h), "mat_eliminate")

    def download(self):
        n, nnz = self.info()
        rp = np.empty(n + 1, dtype=np.int64)
        ci = np.empty(nnz, dtype=np.int32)
        va = np.empty(nnz, dtype=np.float64)
        rhs = np.empty(n, dtype=np.float64)
        _check(hip_lib.phg_hip_mat_download(self._h, _ptr(rp, _i64), _ptr(ci, _i32),
                                            _ptr(va, _dbl), _ptr(rhs, _dbl)),
               "mat_download")
        return rp, ci, va, rhs

    def rhs(self) -> Vec:
        v = Vec(self.n)
        _check(hip_lib.phg_hip_mat_rhs(self._h, v._h), "mat_rhs")
        return v

    def format_bytes(self, fmt: int, fp32: bool) -> int:
        """Bytes one SpMV moves in layout `fmt` (matrix + x + y)."""
        b = _i64()
        _check(hip_lib.phg_hip_mat_format_bytes(self._h, fmt, int(fp32), C.byref(b)),
               "mat_format_bytes")
        return b.value

    def tune(self, reps: int = 10) -> dict:
        """Times the SpMV layouts and keeps the fastest per precision.
        Returns {layout: (fp64_us, fp32_us)} and the choice."""
        _check(hip_lib.phg_hip_mat_tune(self._h, reps), "mat_tune")
        out = {}
        for f, name in FMT_NAMES.items():
            t = []
            for fp32 in (0, 1):
                us = _dbl()
                _check(hip_lib.phg_hip_mat_tune_time(self._h, f, fp32, C.byref(us)),
                       "mat_tune_time")
                t.append(us.value)
            out[name] = tuple(t)
        out["chosen"] = self.format()
        return out

    def format(self):
        a, b = _i32(), _i32()
        _check(hip_lib.phg_hip_mat_get_format(self._h, C.byref(a), C.byref(b)),
               "mat_get_format")
        return FMT_NAMES[a.value], FMT_NAMES[b.value]

    def set_format(self, fmt64: int, fmt32: int) -> None:
        _check(hip_lib.phg_hip_mat_set_format(self._h, fmt64, fmt32), "mat_set_format")

    def matvec(self, x: Vec, y: Vec, alpha=1.0, beta=0.0) -> None:
        _check(hip_lib.phg_hip_matvec(alpha, self._h, x._h, beta, y._h), "matvec")

    def dump_csr(self, fn: str) -> None:
        """phgMatDumpCSR text format (matvec.c:6134)."""
        _check(hip_lib.phg_hip_mat_dump_csr(self._h, fn.encode()), "mat_dump_csr")

    def element_operator(self, sp: "Space", pb: "Problem", kappa=None) -> None:
        """Builds the matrix-free FP32 operator (MFMA element products) that
        the reliable-update and batched solvers then use for their FP32
        iterations (include/phg_hip.h phg_hip_mat_element_operator)."""
        kap = _opt(kappa)
        _check(hip_lib.phg_hip_mat_element_operator(self._h, sp._h, C.byref(pb),
                                                    _ptr(kap, _dbl)), "mat_element_operator")

    def use_element_operator(self, on: bool) -> None:
        _check(hip_lib.phg_hip_mat_use_element_operator(self._h, 1 if on else 0),
               "mat_use_element_operator")

    def element_apply(self, x: Vec, y: Vec) -> None:
        _check(hip_lib.phg_hip_element_apply(self._h, x._h, y._h), "element_apply")

    def element_bytes(self):
        a, b = _i64(), _i64()
        _check(hip_lib.phg_hip_element_bytes(self._h, C.byref(a), C.byref(b)),
               "element_bytes")
        return a.value, b.value

    def element_layout(self):
        """Block layout of the matrix-free operator: blocks, touched-row
        entries, interior rows, shared rows, partial sums, Dirichlet rows,
        and whether it was built on the device (1) or by the host (0)."""
        out = (_i64 * 7)()
        _check(hip_lib.phg_hip_element_layout(self._h, out), "element_layout")
        return {k: int(v) for k, v in zip(
            ("blocks", "entries", "interior", "shared", "partials", "dirichlet", "device"),
            out)}

    def element_time(self, reps: int = 20):
        """(element kernel us, row gather us) per launch."""
        a, b = _dbl(), _dbl()
        _check(hip_lib.phg_hip_element_time(self._h, reps, C.byref(a), C.byref(b)),
               "element_time")
        return a.value, b.value

    def __del__(self):
        if getattr(self, "_h", None):
            hip_lib.phg_hip_mat_destroy(self._h)
            self._h = None


def _opt(a, dtype=np.float64):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


def element_matrices(sp: Space, pb: Problem, kappa=None, f=None):
    M = {1: 4, 2: 10, 3: 20, 4: 35}[sp.order] * pb.dim
    kap = _opt(kappa)
    f = _opt(f) if f is not None else np.zeros(sp.ndof * pb.dim)
    K = np.empty((sp.nelem, M, M), dtype=np.float64)
    B = np.empty((sp.nelem, M), dtype=np.float64)
    _check(hip_lib.phg_hip_element_matrices(sp._h, C.byref(pb), _ptr(kap, _dbl),
                                            _ptr(f, _dbl), _ptr(K, _dbl), _ptr(B, _dbl)),
           "element_matrices")
    return K, B


def assemble(sp: Space, pb: Problem, kappa=None, f=None, g=None) -> Mat:
    n = sp.ndof * pb.dim
    kap = _opt(kappa)
    f = _opt(f) if f is not None else np.zeros(n)
    g = _opt(g) if g is not None else np.zeros(n)
    h = _vp()
    _check(hip_lib.phg_hip_assemble(sp._h, C.byref(pb), _ptr(kap, _dbl), _ptr(f, _dbl),
                                    _ptr(g, _dbl), C.byref(h)), "assemble")
    return Mat(h)


def dof_norms(sp: Space, v: Vec):
    """(||v_h||_L2, ||grad v_h||_L2) of a scalar DOF (phgDofNormL2)."""
    a, b = _dbl(), _dbl()
    _check(hip_lib.phg_hip_dof_norms(sp._h, v._h, C.byref(a), C.byref(b)), "dof_norms")
    return a.value, b.value


def adjoint_logkappa(sp: Space, pb: Problem, kappa, A: Mat, u: Vec, ud: Vec,
                     maxit=100000, rtol=1e-12, lam: Vec | None = None):
    """Adjoint gradient of J = 1/2 int (u_h - ud_h)^2 with respect to log kappa_e.
    Returns (J, grad[nelem], adjoint iterations); fills `lam` when given."""
    kap = _opt(kappa)
    J = _dbl()
    nits = _i32()
    grad = np.empty(sp.nelem)
    _check(hip_lib.phg_hip_adjoint_logkappa(
        sp._h, C.byref(pb), _ptr(kap, _dbl), A._h, u._h, ud._h, maxit, rtol,
        C.byref(J), _ptr(grad, _dbl), lam._h if lam is not None else None,
        C.byref(nits)), "adjoint_logkappa")
    return J.value, grad, nits.value


BLOCK_MAX = 16
SPMM_SHAPES = ((1, 1), (2, 1), (4, 1), (4, 4), (8, 4), (16, 4))


def pcg_block(A: "Mat", bs, xs, maxit=100000, rtol=1e-10, btol=1e-10, atol=0.0,
              mixed=False, delta=0.1):
    """Jacobi-PCG for up to 16 right-hand sides sharing one SpMM per
    iteration.  Returns (nits[k], residual[k])."""
    k = len(bs)
    if k != len(xs) or not 1 <= k <= BLOCK_MAX:
        raise ValueError("1 <= len(bs) == len(xs) <= 16 required")
    barr = (_vp * k)(*[b._h for b in bs])
    xarr = (_vp * k)(*[x._h for x in xs])
    nits = np.zeros(k, dtype=np.int32)
    res = np.zeros(k)
    _check(hip_lib.phg_hip_pcg_block(A._h, k, barr, xarr, maxit, rtol, btol, atol,
                                     int(mixed), delta, _ptr(nits, _i32), _ptr(res, _dbl)),
           "pcg_block")
    return nits, res


def pcg_multi(As, bs, xs, maxit=100000, rtol=1e-10, btol=1e-10, atol=0.0, delta=0.1,
              use_graph=True):
    """Mixed-precision PCG on several independent systems at once (streams,
    optionally one HIP graph replay per chunk).  Returns (nits[k], residual[k])."""
    k = len(As)
    if not (k == len(bs) == len(xs)) or not 1 <= k <= 16:
        raise ValueError("1 <= len(As) == len(bs) == len(xs) <= 16 required")
    aarr = (_vp * k)(*[A._h for A in As])
    barr = (_vp * k)(*[b._h for b in bs])
    xarr = (_vp * k)(*[x._h for x in xs])
    nits = np.zeros(k, dtype=np.int32)
    res = np.zeros(k)
    _check(hip_lib.phg_hip_pcg_multi(k, aarr, barr, xarr, maxit, rtol, btol, atol, delta,
                                     int(use_graph), _ptr(nits, _i32), _ptr(res, _dbl)),
           "pcg_multi")
    return nits, res


class Batch:
    """Re-assembly + mixed-precision solve of several systems in captured HIP
    graphs (phg_hip_batch_*).  systems: [(space, problem, A, kappa, f, g, x)]
    with A the eliminated matrix assembled on space for problem and kappa
    (or None), f, g, x device Vecs; run() assembles every A again from the
    vectors' current contents and solves A x = rhs(A) from the current x.
    Returns (nits[k], residual[k])."""

    def __init__(self, systems, maxit=100000, rtol=1e-10, btol=1e-10, atol=0.0,
                 delta=0.1):
        k = len(systems)
        if not 1 <= k <= 16:
            raise ValueError("1 <= number of systems <= 16 required")
        self._keep = systems
        self._pbs = [pb for _, pb, _, _, _, _, _ in systems]
        null = _vp()
        spa = (_vp * k)(*[sp._h for sp, _, _, _, _, _, _ in systems])
        pba = (_P(Problem) * k)(*[C.pointer(pb) for pb in self._pbs])
        aa = (_vp * k)(*[A._h for _, _, A, _, _, _, _ in systems])
        ka = (_vp * k)(*[(kp._h if kp is not None else null)
                         for _, _, _, kp, _, _, _ in systems])
        fa = (_vp * k)(*[f._h for _, _, _, _, f, _, _ in systems])
        ga = (_vp * k)(*[g._h for _, _, _, _, _, g, _ in systems])
        xa = (_vp * k)(*[x._h for _, _, _, _, _, _, x in systems])
        h = _vp()
        _check(hip_lib.phg_hip_batch_create(k, spa, pba, aa, ka, fa, ga, xa, maxit, rtol,
                                            btol, atol, delta, C.byref(h)),
               "phg_hip_batch_create")
        self._h = h
        self.k = k

    def run(self):
        nits = np.zeros(self.k, dtype=np.int32)
        res = np.zeros(self.k)
        _check(hip_lib.phg_hip_batch_run(self._h, _ptr(nits, _i32), _ptr(res, _dbl)),
               "phg_hip_batch_run")
        return nits, res

    def __del__(self):
        if getattr(self, "_h", None):
            hip_lib.phg_hip_batch_destroy(self._h)
            self._h = None


def spmm_tune(A: "Mat", reps: int = 10) -> dict:
    """Times the SpMM wavefront shapes; returns {shape: (fp64_us, fp32_us)}."""
    _check(hip_lib.phg_hip_spmm_tune(A._h, reps), "spmm_tune")
    out = {}
    for i, sh in enumerate(SPMM_SHAPES):
        t = []
        for fp32 in (0, 1):
            us = _dbl()
            _check(hip_lib.phg_hip_spmm_tune_time(A._h, i, fp32, C.byref(us)),
                   "spmm_tune_time")
            t.append(us.value)
        out[sh] = tuple(t)
    return out


def spmm_set_shape(A: "Mat", shape64: int, shape32: int) -> None:
    """Selects SpMM shapes by index into SPMM_SHAPES."""
    _check(hip_lib.phg_hip_spmm_set_shape(A._h, shape64, shape32), "spmm_set_shape")


def spmm_bytes(A: "Mat", fp32: bool) -> int:
    b = _i64()
    _check(hip_lib.phg_hip_spmm_bytes(A._h, int(fp32), C.byref(b)), "spmm_bytes")
    return b.value


def mat_from_csr(rowptr, cols, vals, rhs, bdry, eliminated: bool) -> "Mat":
    """Uploads a host CSR system (e.g. PHG's packed matrix)."""
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    ci = np.ascontiguousarray(cols, dtype=np.int32)
    va = np.ascontiguousarray(vals, dtype=np.float64)
    rh = np.ascontiguousarray(rhs, dtype=np.float64)
    bd = np.ascontiguousarray(bdry, dtype=np.uint8)
    h = _vp()
    _check(hip_lib.phg_hip_mat_from_csr(len(rp) - 1, _ptr(rp, _i64), _ptr(ci, _i32),
                                        _ptr(va, _dbl), _ptr(rh, _dbl), _ptr(bd, C.c_uint8),
                                        int(eliminated), C.byref(h)), "mat_from_csr")
    return Mat(h)


def stream_copy(n: int, reps: int = 20) -> float:
    """Average microseconds of a plain copy of n doubles (16 n bytes)."""
    us = _dbl()
    _check(hip_lib.phg_hip_stream_copy(n, reps, C.byref(us)), "stream_copy")
    return us.value


def profile(on: bool) -> None:
    """Turns per-kernel HIP-event timing on/off (resets the totals)."""
    _check(hip_lib.phg_hip_profile(int(on)), "phg_hip_profile")


def profile_read(kind: int):
    """(launches, total milliseconds) recorded for a kernel class."""
    n, ms = _i64(), _dbl()
    _check(hip_lib.phg_hip_profile_read(kind, C.byref(n), C.byref(ms)), "profile_read")
    return n.value, ms.value


def pcg(A: Mat, b: Vec, x: Vec, maxit=10000, rtol=1e-10, btol=1e-10, atol=0.0):
    """FP64 Jacobi-PCG (solver-phg.c:485). Returns (nits, residual); the
    residual starts at PCG_Solve's 1.0 (solver-phg.c:2433), which a zero
    right-hand side leaves in place."""
    res = _dbl(1.0)
    nits = _check(hip_lib.phg_hip_pcg(A._h, b._h, x._h, maxit, rtol, btol, atol,
                                      C.byref(res)), "pcg")
    return nits, res.value


def bicgstab(A: Mat, b: Vec, x: Vec, maxit=10000, rtol=1e-10, btol=1e-10, atol=0.0):
    """FP64 Jacobi-BiCGstab (solver-phg.c:607). Returns (nits, residual)."""
    res = _dbl()
    nits = _check(hip_lib.phg_hip_bicgstab(A._h, b._h, x._h, maxit, rtol, btol, atol,
                                           C.byref(res)), "bicgstab")
    return nits, res.value


GMRES_TYPES = ("gmres_l", "gmres_r", "fgmres_r", "lgmres_l", "lgmres_r", "flgmres_r")


def gmres(A: Mat, b: Vec, x: Vec, restart=10, augk=2, maxit=10000, rtol=1e-10, btol=1e-10,
          atol=0.0, gmres_type="flgmres_r"):
    """FP64 Jacobi-GMRES of the -gmres_type named (solver-phg.c:71-77; PHG's
    default flgmres_r, solver-phg.c:2063).  Returns (nits, residual)."""
    res = _dbl()
    if gmres_type not in GMRES_TYPES:
        raise ValueError(f"gmres_type must be one of {GMRES_TYPES}")
    nits = _check(hip_lib.phg_hip_gmres_type(A._h, b._h, x._h, GMRES_TYPES.index(gmres_type),
                                             restart, augk, maxit, rtol, btol, atol,
                                             C.byref(res)), "gmres")
    return nits, res.value


def pcg_mixed(A: Mat, b: Vec, x: Vec, maxit=100, rtol=1e-10, btol=1e-10, atol=0.0,
              inner_rtol=1e-4):
    """FP32 PCG under FP64 iterative refinement.
    Returns (outer_steps, inner_iterations, residual)."""
    res = _dbl()
    inner = _i32()
    outer = _check(hip_lib.phg_hip_pcg_mixed(A._h, b._h, x._h, maxit, rtol, btol, atol,
                                             inner_rtol, C.byref(inner), C.byref(res)),
                   "pcg_mixed")
    return outer, inner.value, res.value


def pcg_reliable(A: Mat, b: Vec, x: Vec, maxit=100000, rtol=1e-10, btol=1e-10,
                 atol=0.0, delta=0.1):
    """FP32 PCG with reliable FP64 residual updates.
    Returns (fp32_iterations, fp64_residual_updates, residual)."""
    res = _dbl()
    upd = _i32()
    nits = _check(hip_lib.phg_hip_pcg_reliable(A._h, b._h, x._h, maxit, rtol, btol,
                                               atol, delta, C.byref(upd),
                                               C.byref(res)), "pcg_reliable")
    return nits, upd.value, res.value


def pc_jacobi(A: Mat, b: Vec, x: Vec) -> None:
    """solver-phg.c:207 jacobi_pc: x = b ./ diag(A) (PhgHipError with code
    PHG_HIP_E_ZERO_DIAG on a zero diagonal entry)."""
    _check(hip_lib.phg_hip_pc_jacobi(A._h, b._h, x._h), "pc_jacobi")
