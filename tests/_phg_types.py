"""ctypes mirrors of the PHG structs re-declared in include/phg_oem_hip.h
(serial PHG build: map.h:83-140, matvec.h:33-170, solver.h:26-103), so the
tests can act as PHG's solver driver and hand the plugin PHG-shaped objects."""
from __future__ import annotations

import ctypes as C
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OEM_LIB_PATH = os.path.join(ROOT, "phg_amd", "lib", "libphg_oem_hip.so")

FLOAT, INT, BOOLEAN = C.c_double, C.c_int, C.c_int
MAT_OP_N, MAT_OP_T, MAT_OP_D = 0, 1, 2
PHG_UNPACKED, PHG_PACKED, PHG_MATRIX_FREE = 0, 1, 3
_P = C.POINTER


class MAP(C.Structure):
    _fields_ = [("lif", FLOAT), ("dofs", C.c_void_p), ("map_nb", C.c_void_p),
                ("D2Lmap", _P(INT)), ("L2Vmap", _P(INT)), ("O2Gmap", _P(INT)),
                ("ordering", _P(INT)), ("partition", _P(INT)), ("offset", _P(INT)),
                ("bdry", _P(INT)), ("bdry_partition", _P(INT)), ("bdry_nlocal", INT),
                ("bdry_localsize", INT), ("bdry_nglobal", INT), ("D2Lmap_size", INT),
                ("L2Vmap_size", INT), ("nlocal", INT), ("nglobal", INT),
                ("localsize", INT), ("comm", C.c_int), ("rank", C.c_int),
                ("nprocs", C.c_int), ("halo_type", C.c_int), ("refcount", C.c_int),
                ("ndof", C.c_ubyte), ("bogus_Lmap", BOOLEAN)]


class VEC(C.Structure):
    _fields_ = [("magic", C.c_int), ("map", _P(MAP)), ("data", _P(FLOAT)),
                ("offp_data", _P(FLOAT)), ("mat", C.c_void_p), ("nvec", C.c_int),
                ("assembled", BOOLEAN)]


class MAT_ROW(C.Structure):
    _fields_ = [("data", _P(FLOAT)), ("cols", _P(INT)), ("ncols", INT), ("alloc", INT)]


class MAT(C.Structure):
    _fields_ = [("magic", C.c_int), ("rmap", _P(MAP)), ("cmap", _P(MAP)),
                ("rows", _P(MAT_ROW)), ("row_buffer", MAT_ROW), ("diag", _P(FLOAT)),
                ("diag1", _P(FLOAT)), ("type", C.c_int), ("packed_cols", _P(INT)),
                ("packed_data", _P(FLOAT)), ("packed_ind", _P(C.c_size_t)),
                ("nnz_d", C.c_size_t), ("nnz_o", C.c_size_t), ("localsize", INT),
                ("handle_bdry_eqns", BOOLEAN), ("mv_a", FLOAT), ("mv_b", FLOAT),
                ("mv_x", C.c_void_p), ("mv_func", C.c_void_p), ("mv_data", _P(C.c_void_p)),
                ("bdry_localsize", INT), ("bdry", _P(INT)), ("bdry_eqns", C.c_void_p),
                ("bdry_eqns_count", INT), ("bdry_eqns_allocated", INT),
                ("bdry_rows", C.c_void_p), ("blocks", C.c_void_p), ("refcount", C.c_int),
                ("mode", C.c_int), ("assembled", BOOLEAN)]


class SOLVER(C.Structure):
    _fields_ = [("residual", FLOAT), ("rtol", FLOAT), ("atol", FLOAT), ("btol", FLOAT),
                ("cond", C.c_double), ("mat", _P(MAT)), ("mat_bak", _P(MAT)),
                ("rhs", _P(VEC)), ("D", _P(FLOAT)), ("oem_solver", C.c_void_p),
                ("oem_data", C.c_void_p), ("pc", C.c_void_p), ("St", C.c_void_p),
                ("symmetry", C.c_int), ("matrix_free", C.c_int), ("nits", C.c_int),
                ("maxit", C.c_int), ("monitor", BOOLEAN), ("scaling", BOOLEAN),
                ("warn_maxit", BOOLEAN), ("pc_option_flag", BOOLEAN),
                ("rhs_updated", BOOLEAN), ("sent_to_oem", BOOLEAN),
                ("oem_created", BOOLEAN), ("assembled", BOOLEAN)]


_SP = _P(SOLVER)
FN_SOLVER = C.CFUNCTYPE(C.c_int, _SP)
FN_ADDMAT = C.CFUNCTYPE(C.c_int, _SP, INT, _P(INT), INT, _P(INT), _P(FLOAT))
FN_SOLVE = C.CFUNCTYPE(C.c_int, _SP, _P(VEC), BOOLEAN)


class OEM_SOLVER(C.Structure):
    _fields_ = [("name", C.c_char_p), ("RegisterOptions", C.c_void_p),
                ("Initialize", C.c_void_p), ("Finalize", C.c_void_p),
                ("Init", FN_SOLVER), ("Create", FN_SOLVER), ("Destroy", FN_SOLVER),
                ("AddMatrixEntries", FN_ADDMAT), ("AddRHSEntries", C.c_void_p),
                ("Assemble", FN_SOLVER), ("SetPC", C.c_void_p), ("Solve", FN_SOLVE),
                ("GetSolver", C.c_void_p), ("GetMatrix", C.c_void_p), ("GetPC", C.c_void_p),
                ("req_symmetry", C.c_int), ("iterative", BOOLEAN), ("parallel", BOOLEAN),
                ("allow_matrix_free", BOOLEAN), ("allow_extern_pc", BOOLEAN)]


def load():
    lib = C.CDLL(OEM_LIB_PATH)
    sig = [("phgMapCreateSimpleMap", _P(MAP), [C.c_int, INT, INT]),
           ("phgMapCreateVec", _P(VEC), [_P(MAP), C.c_int]),
           ("phgVecDestroy", None, [_P(_P(VEC))]),
           ("phgSolverMat2Solver", _SP, [C.c_void_p, _P(MAT)]),
           ("phgSolverVecSolve", C.c_int, [_SP, BOOLEAN, _P(VEC)]),
           ("phgSolverAssemble", C.c_int, [_SP]),
           ("phgSolverDestroy", C.c_int, [_P(_SP)]),
           ("phgSolverSetMaxIt", None, [_SP, C.c_int]),
           ("phgSolverSetTol", None, [_SP, FLOAT]),
           ("phgSolverHipSetOptions", C.c_int, [_SP, C.c_int, C.c_int, C.c_int, C.c_int]),
           ("phgSolverHipSetGMRESType", C.c_int, [_SP, C.c_char_p]),
           ("phgMatCreateHip", _P(MAT), [C.c_void_p, BOOLEAN]),
           ("phgMatDestroyHip", None, [_P(_P(MAT))]),
           ("phgHipMatVec", C.c_int, [C.c_int, _P(MAT), _P(VEC), _P(VEC)]),
           ("phgHipJacobiPC", None, [C.c_void_p, _P(VEC), _P(_P(VEC))])]
    for name, res, args in sig:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def vtable(lib, name):
    return OEM_SOLVER.in_dll(lib, name)
