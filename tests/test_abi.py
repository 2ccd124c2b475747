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
