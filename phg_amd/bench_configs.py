"""Benchmarks of the secondary workloads (one JSON line each).

    python -m phg_amd.bench_configs c1 c2asm c3 c4 c5

c1  P1 Poisson -Lap u = f, u = sin(pi x) sin(pi y) sin(pi z), 32^3 x 6 mesh
    (test/cube.dat + 15 PHG bisections, 35,937 DOFs): assembly + FP64 Jacobi-PCG to rtol 1e-10, error
    norms of u_h - I_h u.
c2asm  assembly + Dirichlet elimination of the headline P2 kappa-Poisson
    problem (64^3 x 6, 2,146,689 DOFs).
c3  P1 adjoint, 48^3 x 6 mesh read from an ALBERT file, log kappa ~
    N(0, 0.3^2): forward solve + adjoint gradient.
c4  vector P2 elasticity (lambda = mu = 1), x = 0 clamped, 40^3 x 6 mesh,
    16 body-force right-hand sides U(-1, 1): block PCG (FP64 and mixed
    precision, rtol 1e-8) against 16 sequential single-vector FP64 solves.
c5  batches of P3 kappa-Poisson problems on locally refined meshes (see
    phg_amd.multimesh).

All inputs are synthetic and seeded.  GPU timings only; the results are
checked against the CPU restatement by the test suite.
"""
from __future__ import annotations

import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _timed(fn, reps=3, warm=1):
    import phg_amd as pa
    for _ in range(warm):
        fn()
    pa.synchronize()
    t = time.perf_counter()
    out = None
    for _ in range(reps):
        out = fn()
    pa.synchronize()
    return (time.perf_counter() - t) / reps, out


def c1():
    import phg_amd as pa
    n, p = 32, 1
    g = pa.Grid.phg_cube(n)          # test/cube.dat + 15 bisections
    xyz = g.dof_coords(p)
    u_ex = np.sin(np.pi * xyz[:, 0]) * np.sin(np.pi * xyz[:, 1]) * np.sin(np.pi * xyz[:, 2])
    f = 3 * np.pi ** 2 * u_ex
    sp = pa.Space(g, p)
    pb = pa.Problem.poisson()

    def asm():
        A = pa.assemble(sp, pb, f=f, g=u_ex)
        A.eliminate()
        return A
    t_asm, A = _timed(asm)
    b = A.rhs()
    x = pa.Vec(A.n)

    def solve():
        x.axpby(0.0, x, 0.0)
        return pa.pcg(A, b, x, rtol=1e-10, btol=1e-10)
    t_sol, (nits, res) = _timed(solve, reps=5)
    e = pa.Vec(A.n, x.numpy() - u_ex)
    l2, h1 = pa.dof_norms(sp, e)
    return {"config": "c1", "workload": "P1 Poisson sin, 32^3x6 (35,937 DOFs)",
            "n": A.n, "nnz": A.info()[1], "gpu_assemble_s": round(t_asm, 5),
            "gpu_solve_s": round(t_sol, 5), "its": nits, "gpu_dofs_per_s": round(A.n / t_sol, 1),
            "err_l2": l2, "err_grad_l2": h1}


def c2asm():
    """Assembly and Dirichlet elimination (with the solver numbering) of the
    P2 kappa-Poisson problem on the 64^3 x 6 mesh (test/cube.dat + 18 bisections), inputs on the host as
    PHG hands them over."""
    import phg_amd as pa
    n, p = 64, 2
    g = pa.Grid.phg_cube(n)          # test/cube.dat + 18 bisections
    nd = g.ndof(p)
    bd = g.dof_bdry(p)
    kappa = np.exp(np.random.default_rng(2024).normal(0.0, 0.5, size=g.nelem))
    f, gv = np.ones(nd), np.zeros(nd)
    sp = pa.Space(g, p, bdry=bd)
    pb = pa.Problem.poisson(use_kappa=True)
    ta, te = [], []
    for _ in range(4):
        pa.synchronize()
        t0 = time.perf_counter()
        A = pa.assemble(sp, pb, kappa=kappa, f=f, g=gv)
        pa.synchronize()
        t1 = time.perf_counter()
        A.eliminate()
        pa.synchronize()
        ta.append(t1 - t0)
        te.append(time.perf_counter() - t1)
        nnz = A.info()[1]
        del A
    # the first pass loads the kernels' code objects
    return {"config": "c2asm", "workload": "P2 kappa-Poisson assembly, 64^3x6 "
            f"({g.nelem} tets, {nd} DOFs)", "nnz": nnz,
            "assemble_s": round(float(np.mean(ta[1:])), 5),
            "eliminate_s": round(float(np.mean(te[1:])), 5),
            "first_pass_s": round(ta[0] + te[0], 5)}


def c3():
    import phg_amd as pa
    n, p = 48, 1
    g0 = pa.Grid.kuhn(n)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "cube48.dat")
        g0.write_albert(path)
        g = pa.Grid.read_albert(path)
    rng = np.random.default_rng(31)
    kappa = np.exp(rng.normal(0.0, 0.3, size=g.nelem))
    rng = np.random.default_rng(32)
    kappa_d = np.exp(rng.normal(0.0, 0.3, size=g.nelem))
    nd = g.ndof(p)
    f, gz = np.ones(nd), np.zeros(nd)
    sp = pa.Space(g, p)
    pb = pa.Problem.poisson(use_kappa=True)

    def solve(k):
        A = pa.assemble(sp, pb, kappa=k, f=f, g=gz)
        A.eliminate()
        u = pa.Vec(A.n)
        pa.pcg(A, A.rhs(), u, rtol=1e-12, btol=1e-12)
        return A, u
    _, ud = solve(kappa_d)

    def run():
        A, u = solve(kappa)
        return pa.adjoint_logkappa(sp, pb, kappa, A, u, ud, rtol=1e-12)
    t_gpu, (J, grad, nits) = _timed(run)
    return {"config": "c3", "workload": "P1 adjoint, 48^3x6 from ALBERT file (117,649 DOFs)",
            "gpu_forward_adjoint_s": round(t_gpu, 5), "adjoint_its": nits, "J": J,
            "grad_abs_max": float(np.abs(grad).max())}


def c4(seq=16):
    import phg_amd as pa
    n, p, k = 40, 2, 16
    g = pa.Grid.kuhn(n, (1, -1, -1, -1, -1, -1))
    nd = g.ndof(p)
    sp = pa.Space(g, p)
    pb = pa.Problem.elasticity(E=2.5, nu=0.25)   # lambda = mu = 1
    rng = np.random.default_rng(40)
    A = None
    bs = []
    t0 = time.perf_counter()
    for j in range(k):
        fj = rng.uniform(-1.0, 1.0, size=3 * nd)
        Aj = pa.assemble(sp, pb, f=fj, g=np.zeros(3 * nd))
        Aj.eliminate()
        bs.append(Aj.rhs())
        if A is None:
            A = Aj
        else:
            del Aj
    t_setup = time.perf_counter() - t0
    n_, nnz = A.info()
    shapes = pa.spmm_tune(A, reps=10)
    xs = [pa.Vec(n_) for _ in range(k)]

    def block(mixed):
        def run():
            for x in xs:
                x.axpby(0.0, x, 0.0)
            return pa.pcg_block(A, bs, xs, rtol=1e-8, btol=1e-8, mixed=mixed)
        return run
    t_mp, (nits_mp, _) = _timed(block(True), reps=2)
    x_mp = [x.numpy() for x in xs[:2]]
    t_64, (nits_64, _) = _timed(block(False), reps=2)
    x_64 = [x.numpy() for x in xs[:2]]
    # the reference's way: one right-hand side after another
    A.tune(reps=10)
    x1 = pa.Vec(n_)
    pa.synchronize()
    t0 = time.perf_counter()
    its_seq = []
    for j in range(seq):
        x1.axpby(0.0, x1, 0.0)
        its_seq.append(pa.pcg(A, bs[j], x1, rtol=1e-8, btol=1e-8)[0])
    pa.synchronize()
    t_seq = (time.perf_counter() - t0) * k / seq
    err = max(np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(x_mp, x_64))
    best = min(shapes, key=lambda s: shapes[s][1])
    b32 = pa.spmm_bytes(A, True)
    b64 = pa.spmm_bytes(A, False)
    best64 = min(shapes, key=lambda s: shapes[s][0])
    return {"config": "c4",
            "workload": "vector P2 elasticity 40^3x6 (1,594,323 DOFs), 16 RHS, rtol 1e-8",
            "n": n_, "nnz": nnz, "setup_s_16_assemblies": round(t_setup, 3),
            "block_mixed_s": round(t_mp, 4), "block_fp64_s": round(t_64, 4),
            "sequential_fp64_s": round(t_seq, 4), "sequential_solves_timed": seq,
            "its_block_mixed_max": int(nits_mp.max()), "its_block_fp64_max": int(nits_64.max()),
            "its_sequential_max": int(max(its_seq)),
            "rhs_dofs_per_s_block_mixed": round(k * n_ / t_mp, 1),
            "rhs_dofs_per_s_block_fp64": round(k * n_ / t_64, 1),
            "rhs_dofs_per_s_sequential": round(k * n_ / t_seq, 1),
            "speedup_block_mixed_vs_sequential": round(t_seq / t_mp, 2),
            "mixed_vs_fp64_block_rel_diff": err,
            "spmm_shapes_us": {str(s): v for s, v in shapes.items()},
            "spmm_fp32_GBps": round(b32 / shapes[best][1] / 1e3, 1),
            "spmm_fp64_GBps": round(b64 / shapes[best64][0] / 1e3, 1),
            "spmm_bytes": [b64, b32]}


def c4spmm(reps=20):
    """The C4 system's SpMM alone (for kernel traces and counters): the
    elasticity matrix of c4, the SpMM shape tuner, then `reps` launches of
    the chosen shape in FP32 and in FP64 (16 right-hand sides)."""
    import phg_amd as pa
    n, p = 40, 2
    g = pa.Grid.kuhn(n, (1, -1, -1, -1, -1, -1))
    nd = g.ndof(p)
    sp = pa.Space(g, p)
    pb = pa.Problem.elasticity(E=2.5, nu=0.25)
    fj = np.random.default_rng(40).uniform(-1.0, 1.0, size=3 * nd)
    A = pa.assemble(sp, pb, f=fj, g=np.zeros(3 * nd))
    A.eliminate()
    shapes = pa.spmm_tune(A, reps=reps)
    best32 = min(shapes, key=lambda s: shapes[s][1])
    best64 = min(shapes, key=lambda s: shapes[s][0])
    b32, b64 = pa.spmm_bytes(A, True), pa.spmm_bytes(A, False)
    n_, nnz = A.info()
    # a plain copy of known size, the calibration of HBM counters
    ncopy = 1 << 24
    copy_us = pa.stream_copy(ncopy, reps)
    return {"config": "c4spmm", "workload": "vector P2 elasticity 40^3x6 SpMM, 16 columns",
            "n": n_, "nnz": nnz, "copy": {"bytes_each_way": 8 * ncopy, "us": copy_us},
            "shapes_us": {str(s): v for s, v in shapes.items()},
            "fp32": {"shape": best32, "us": shapes[best32][1], "bytes": b32,
                     "GBps": round(b32 / shapes[best32][1] / 1e3, 1)},
            "fp64": {"shape": best64, "us": shapes[best64][0], "bytes": b64,
                     "GBps": round(b64 / shapes[best64][0] / 1e3, 1)}}


def solvers(n=32, p=2, rtol=1e-8):
    """PHG's Krylov solvers on one kappa-Poisson system (Jacobi PC, FP64,
    x0 = 0, rtol = btol): PCG, BiCGstab and the default GMRES (flgmres_r,
    restart 10, augk 2)."""
    import phg_amd as pa
    g = pa.Grid.phg_cube(n) if n & (n - 1) == 0 else pa.Grid.kuhn(n)
    nd = g.ndof(p)
    rng = np.random.default_rng(2024)
    kappa = np.exp(rng.normal(0.0, 0.5, size=g.nelem))
    sp = pa.Space(g, p)
    A = pa.assemble(sp, pa.Problem.poisson(use_kappa=True), kappa=kappa, f=np.ones(nd),
                    g=np.zeros(nd))
    A.eliminate()
    A.tune(reps=10)
    b = A.rhs()
    out = {"config": "solvers",
           "workload": f"P{p} kappa-Poisson {n}^3x6 ({A.n} DOFs), Jacobi PC, rtol {rtol}",
           "fp64_layout": A.format()[0]}
    xs = {}
    for name, fn in (("pcg", lambda x: pa.pcg(A, b, x, rtol=rtol, btol=rtol)),
                     ("bicgstab", lambda x: pa.bicgstab(A, b, x, rtol=rtol, btol=rtol)),
                     ("gmres", lambda x: pa.gmres(A, b, x, maxit=20000, rtol=rtol,
                                                  btol=rtol))):
        x = pa.Vec(A.n)

        def run(x=x, fn=fn):
            x.axpby(0.0, x, 0.0)
            return fn(x)
        t, (nits, res) = _timed(run, reps=2)
        xs[name] = x.numpy()
        out[name] = {"s": round(t, 4), "its": int(nits), "residual": res,
                     "us_per_iteration": round(t / max(nits, 1) * 1e6, 1),
                     "dofs_per_s": round(A.n / t, 1)}
    ref = xs["pcg"]
    for name in ("bicgstab", "gmres"):
        out[name]["rel_diff_vs_pcg"] = float(np.linalg.norm(xs[name] - ref) /
                                             np.linalg.norm(ref))
    return out


def main(argv):
    import phg_amd as pa
    pa.init(0)
    which = argv or ["c1", "c3", "c4"]
    for w in which:
        if w == "c5":
            from phg_amd import multimesh
            out = multimesh.bench()
        else:
            out = globals()[w]()
        print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
