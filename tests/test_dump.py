"""phgMatDumpCSR / phgVecDump text formats (matvec.c:6134-6297)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as orc
from tests._cases import poisson_case

pytestmark = pytest.mark.gpu


def test_dump_matches_reference_sample_layout(hip, tmp_path):
    """The 4 x 4 sample of the matvec.c:6134 comment: size, ia, ja, a."""
    pa = hip
    rowptr = np.array([0, 3, 5, 8, 10], dtype=np.int64)
    cols = np.array([0, 1, 3, 1, 2, 0, 2, 3, 2, 3], dtype=np.int32)
    vals = np.array([4.0, -1.0, 0.1, 5.5, 1.0 / 3.0, 2e-300, 7.0, -2.5, 1e17, 3.0])
    A = pa.mat_from_csr(rowptr, cols, vals, np.zeros(4), np.zeros(4, np.uint8), True)
    fn = tmp_path / "A.csr"
    A.dump_csr(str(fn))
    lines = fn.read_text().splitlines()
    expect = (["4"] + [str(int(v)) for v in rowptr] + [str(int(c)) for c in cols] +
              ["%0.17g" % v for v in vals])
    assert lines == expect
    x = pa.Vec(4, vals[:4])
    fv = tmp_path / "x.vec"
    x.dump(str(fv))
    assert fv.read_text().splitlines() == ["4"] + ["%0.17g" % v for v in vals[:4]]


def test_dump_of_assembled_system_round_trips(hip, tmp_path):
    """An assembled, not yet eliminated system: the dump eliminates first
    (as phgMatDumpCSR assembles first) and holds exactly the oracle's
    eliminated CSR in PHG's storage order."""
    pa = hip
    c = poisson_case(3, 2, "kappa")
    pb = pa.Problem.poisson(use_kappa=True)
    sp = pa.Space(c["grid"], 2, bdry=c["bdry"])
    A = pa.assemble(sp, pb, kappa=c["kappa"], f=c["f"], g=c["g"])
    fn = tmp_path / "A.csr"
    A.dump_csr(str(fn))
    tok = fn.read_text().split()
    n = int(tok[0])
    rp = np.array(tok[1:n + 2], dtype=np.int64)
    nnz = int(rp[-1])
    ci = np.array(tok[n + 2:n + 2 + nnz], dtype=np.int64)
    va = np.array([float(t) for t in tok[n + 2 + nnz:]])
    assert len(va) == nnz
    Ao = orc.assemble(orc.problem(2, use_kappa=True), c["verts"], c["elem_verts"], c["map"],
                      c["ndof"], c["bdry"], c["g"], c["f"], c["kappa"])
    orc.eliminate(Ao, c["bdry"])
    assert n == Ao.n
    assert np.array_equal(rp, np.asarray(Ao.rowptr))
    assert np.array_equal(ci, np.asarray(Ao.cols))
    assert np.array_equal(va, np.asarray(Ao.vals))   # %.17g round-trips doubles
