"""Writes c2_64_oracle.npz: the oracle's FP64 solution of the P2
kappa-Poisson workload on the 64^3 x 6 mesh (test/cube.dat bisected 18 times
by PHG's refinement), at a fixed sample of DOFs.

Problem (the benchmark's headline system): -div(kappa grad u) = 1, u = 0 on
the boundary of the unit cube, 1,572,864 tetrahedra, 2,146,689 DOFs,
kappa_e = exp(g_e) with g_e ~ N(0, 0.5^2) from numpy's default_rng(2024),
element order.  The oracle assembles and eliminates it (the CPU restatement
of PHG's assembly, bit for bit the GPU CSR) and runs its Jacobi-PCG
(solver-phg.c:485 solver_pcg) from x0 = 0 to rtol = btol = 1e-12, far below
the 1e-8 the solvers are checked at.

Stored: `index` (4096 DOFs drawn from default_rng(7), sorted), `x` the
solution there, `norm2` the 2-norm of the whole solution, `nits` and
`residual` of the oracle PCG.  About 90 s of one CPU core:
    python tests/golden/make_c2_oracle.py
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main():
    import oracle as orc
    from tests._cases import poisson_case
    c = poisson_case(64, 2, "kappa", sigma=0.5, seed=2024, mesh="phg")
    t0 = time.perf_counter()
    A = orc.assemble(orc.problem(2, use_kappa=True), c["verts"], c["elem_verts"], c["map"],
                     c["ndof"], c["bdry"], c["g"], c["f"], c["kappa"])
    orc.eliminate(A, c["bdry"])
    t1 = time.perf_counter()
    x, nits, res = orc.pcg(A, A.rhs, rtol=1e-12, btol=1e-12, maxit=100000)
    t2 = time.perf_counter()
    idx = np.sort(np.random.default_rng(7).choice(c["ndof"], 4096, replace=False))
    out = os.path.join(HERE, "c2_64_oracle.npz")
    np.savez(out, index=idx.astype(np.int32), x=x[idx], norm2=np.linalg.norm(x),
             nits=nits, residual=res)
    print(f"assembly {t1 - t0:.1f}s, PCG {nits} its in {t2 - t1:.1f}s, residual "
          f"{res:.3e}, |x| {np.linalg.norm(x):.17g} -> {out}")


if __name__ == "__main__":
    main()
