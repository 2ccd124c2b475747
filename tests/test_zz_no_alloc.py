"""phg_hip_dof_norms_noalloc (the entry point that allocates no device
memory) through the C ABI: the same norms as phg_hip_dof_norms bit for bit,
on two calls in a row, the second on a smaller mesh (the scratch array keeps
the first call's values beyond the second mesh's elements), and a clear
error for a mesh beyond the scratch array's size."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests._cases import poisson_case

pytestmark = pytest.mark.gpu

_dbl = C.c_double


def _entry(pa, name):
    f = getattr(pa.hip_lib, name)
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(_dbl), C.POINTER(_dbl)]
    return f


def _norms(pa, name, sp, v):
    l2, h1 = _dbl(), _dbl()
    rc = _entry(pa, name)(sp._h, v._h, C.byref(l2), C.byref(h1))
    return rc, l2.value, h1.value


def test_noalloc_norms_match_the_allocating_entry_point(hip):
    pa = hip
    rng = np.random.default_rng(17)
    for n, p in ((8, 2), (3, 1)):          # second call: fewer elements
        c = poisson_case(n, p, "sin")
        sp = pa.Space(c["grid"], p, bdry=c["bdry"])
        v = pa.Vec(c["ndof"], rng.standard_normal(c["ndof"]))
        rc, l2, h1 = _norms(pa, "phg_hip_dof_norms", sp, v)
        assert rc == 0
        rc, l2n, h1n = _norms(pa, "phg_hip_dof_norms_noalloc", sp, v)
        assert rc == 0, pa.hip_lib.phg_hip_last_error()
        print(f"P{p} {n}^3: l2 {l2n!r} (allocating {l2!r}), grad {h1n!r} ({h1!r})",
              flush=True)
        assert np.float64(l2n).view(np.uint64) == np.float64(l2).view(np.uint64)
        assert np.float64(h1n).view(np.uint64) == np.float64(h1).view(np.uint64)
        # v = x: the closed forms 1/3 and 1 on the unit cube
        vx = pa.Vec(c["ndof"], c["coords"][:, 0].copy())
        rc, l2x, h1x = _norms(pa, "phg_hip_dof_norms_noalloc", sp, vx)
        assert rc == 0
        assert abs(l2x * l2x - 1.0 / 3.0) < 1e-13 and abs(h1x * h1x - 1.0) < 1e-13


def test_noalloc_rejects_meshes_beyond_the_scratch(hip):
    pa = hip
    g = pa.Grid.kuhn(72)                    # 2,239,488 elements > 2^21
    assert g.nelem > 1 << 21
    sp = pa.Space(g, 1)
    v = pa.Vec(g.ndof(1))
    rc, _, _ = _norms(pa, "phg_hip_dof_norms_noalloc", sp, v)
    assert rc == -2
    assert b"at most" in pa.hip_lib.phg_hip_last_error()
    w = pa.Vec(g.ndof(1) + 1)               # not a DOF of the space
    rc, _, _ = _norms(pa, "phg_hip_dof_norms_noalloc", sp, w)
    assert rc == -2
