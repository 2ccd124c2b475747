"""The C-ABI library loads and exports every symbol include/phg_hip.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(phg_hip_[a-z0-9_]+)\s*\(", text)))


def test_hip_library_exports_header_symbols():
    import phg_amd
    lib = ctypes.CDLL(phg_amd.HIP_LIB_PATH)
    names = _declared("phg_hip.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_uninitialised_calls_fail_cleanly():
    import phg_amd
    lib = phg_amd.hip_lib
    # no phg_hip_init in this process: entry points report the state error
    v = ctypes.c_void_p()
    rc = lib.phg_hip_vec_create(10, ctypes.byref(v))
    assert rc == -5
    assert b"phg_hip_init" in lib.phg_hip_last_error()


def test_oem_library_exports_header_symbols():
    """libphg_oem_hip.so exports the plugin vtables, the MV_FUNC / PC_PROC
    hooks and the PHG stand-ins include/phg_oem_hip.h declares."""
    import phg_amd  # noqa: F401  (loads libphg_hip.so, which it links)
    lib = ctypes.CDLL(os.path.join(ROOT, "phg_amd", "lib", "libphg_oem_hip.so"))
    text = open(os.path.join(ROOT, "include", "phg_oem_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    funcs = set(re.findall(r"\b(phg[A-Z][A-Za-z0-9_]*)\s*\(", text))
    data = set(re.findall(r"extern\s+OEM_SOLVER\s+(\w+)\s*;", text))
    assert {"phgSolverHipPCG_", "phgSolverHipBiCGstab_", "phgSolverHipGMRES_"} <= data
    assert {"phgHipMatVec", "phgHipJacobiPC", "phgSolverVecSolve"} <= funcs
    missing = [n for n in sorted(funcs | data) if not hasattr(lib, n)]
    assert not missing, missing
