"""Batches of independent finite-element problems on unevenly refined meshes.

Each mesh starts from the n^3 x 6 mesh of the unit cube as PHG builds it
(test/cube.dat bisected 3 log2(n) times) and is refined toward a seeded
random corner of the cube with `levels` sweeps of PHG's bisection (host,
phg_grid_bisect); the problem on it is -div(kappa grad u) = 1,
u = 0 on the boundary, with a lognormal P0 coefficient.  The systems are
assembled on the GPU and solved together by phg_hip_pcg_multi (all systems
in flight on a stream pool; a chunk of iterations of every system captured
once in a HIP graph and replayed until all converge).

    python -m phg_amd.multimesh          # the 8-mesh P3 batch, one JSON line
    python -m phg_amd.multimesh batch    # assembly + solve for new kappa: sequential
                                         # against one phg_hip_batch_run
"""
from __future__ import annotations

import json
import sys
import time

import numpy as np

CORNERS = np.array([[i, j, k] for i in (0.0, 1.0) for j in (0.0, 1.0) for k in (0.0, 1.0)])


def batch_meshes(nmesh=8, n=16, seed=5, lmin=2, lmax=8, radius=0.4):
    """(grid, corner, levels) of each mesh: corners and level counts drawn
    from `seed`; the level counts are spread over [lmin, lmax]."""
    import phg_amd as pa
    rng = np.random.default_rng(seed)
    base = pa.Grid.phg_cube(n) if n & (n - 1) == 0 else pa.Grid.kuhn(n)
    levels = np.linspace(lmin, lmax, nmesh).round().astype(int)
    rng.shuffle(levels)
    out = []
    for i in range(nmesh):
        corner = CORNERS[rng.integers(0, 8)]
        out.append((base.refine_toward(corner, int(levels[i]), radius=radius), corner,
                    int(levels[i])))
    return out


def assemble_batch(meshes, p=3, sigma=0.5, seed=6, element_operator=False,
                   kappas=None):
    """Assembles and eliminates each problem; returns [(space, A, b)].  With
    element_operator, each matrix also gets the matrix-free FP32 operator
    (MFMA element products) for the FP32 iterations of the solvers; the
    coefficients drawn are appended to the list `kappas` when given."""
    import phg_amd as pa
    rng = np.random.default_rng(seed)
    pb = pa.Problem.poisson(use_kappa=True)
    out = []
    for g, _, _ in meshes:
        kappa = np.exp(rng.normal(0.0, sigma, size=g.nelem))
        if kappas is not None:
            kappas.append(kappa)
        nd = g.ndof(p)
        sp = pa.Space(g, p)
        A = pa.assemble(sp, pb, kappa=kappa, f=np.ones(nd), g=np.zeros(nd))
        A.eliminate()
        if element_operator:
            A.element_operator(sp, pb, kappa=kappa)
        out.append((sp, A, A.rhs()))
    return out


def A_fmt32(A):
    """Format code of the FP32 SpMV layout the tuner chose."""
    import phg_amd as pa
    name = A.format()[1]
    return next(k for k, v in pa.FMT_NAMES.items() if v == name)


def bench(nmesh=8, n=16, p=3, rtol=1e-8, reps=2):
    import phg_amd as pa
    t0 = time.perf_counter()
    meshes = batch_meshes(nmesh, n)
    t_mesh = time.perf_counter() - t0
    pa.synchronize()
    t0 = time.perf_counter()
    kappas = []
    sys_ = assemble_batch(meshes, p, kappas=kappas)
    pa.synchronize()
    t_asm = time.perf_counter() - t0
    As = [A for _, A, _ in sys_]
    bs = [b for _, _, b in sys_]
    ndofs = [A.n for A in As]
    for A in As:
        A.tune(reps=5)
    xs = [pa.Vec(A.n) for A in As]

    def zero():
        for x in xs:
            x.axpby(0.0, x, 0.0)

    def run_seq():
        zero()
        return [pa.pcg_reliable(A, b, x, rtol=rtol, btol=rtol)[0] for A, b, x in zip(As, bs, xs)]

    def run_multi(graph):
        def f():
            zero()
            return pa.pcg_multi(As, bs, xs, rtol=rtol, btol=rtol, use_graph=graph)[0]
        return f

    def timed(fn):
        fn()
        pa.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            out = fn()
        pa.synchronize()
        return (time.perf_counter() - t) / reps, out
    t_seq, its_seq = timed(run_seq)
    x_seq = [x.numpy() for x in xs]
    t_str, its_str = timed(run_multi(False))
    t_gr, its_gr = timed(run_multi(True))
    same = all(np.array_equal(a, x.numpy()) for a, x in zip(x_seq, xs))
    x_csr = [x.numpy() for x in xs]
    # the same batch with the FP32 iterations on the matrix-free operator
    pa.synchronize()
    t0 = time.perf_counter()
    pb = pa.Problem.poisson(use_kappa=True)
    for (sp, A, _), kap in zip(sys_, kappas):
        A.element_operator(sp, pb, kappa=kap)
    pa.synchronize()
    t_eop_build = time.perf_counter() - t0
    t_eseq, its_eseq = timed(run_seq)
    t_egr, its_egr = timed(run_multi(True))
    dev = max(float(np.linalg.norm(x.numpy() - a) / np.linalg.norm(a))
              for a, x in zip(x_csr, xs))
    elem_b = [A.element_bytes() for A in As]
    # kernel level on the largest system: FP32 CSR SpMV (tuned layout) against
    # the operator's block kernel (the solvers sum the shared rows in their
    # update kernel; the stand-alone fix-up is listed beside it)
    big = int(np.argmax(ndofs))
    Ab = As[big]
    t_el, t_ga = Ab.element_time(reps=20)
    t_csr = Ab.tune(reps=20)[Ab.format()[1]][1]
    kernel_level = {"system": big, "dofs": ndofs[big], "nelem": meshes[big][0].nelem,
                    "csr_fp32_layout": Ab.format()[1], "csr_fp32_spmv_us": round(t_csr, 2),
                    "csr_fp32_GBps": round(Ab.format_bytes(A_fmt32(Ab), True) / t_csr / 1e3, 1),
                    "block_kernel_us": round(t_el, 2), "fixup_kernel_us": round(t_ga, 2),
                    "block_kernel_GBps": round(elem_b[big][0] / t_el / 1e3, 1),
                    "fixup_kernel_GBps": round(elem_b[big][1] / t_ga / 1e3, 1),
                    "block_speedup_vs_csr": round(t_csr / t_el, 3)}
    tot = sum(ndofs)
    return {"config": "c5",
            "workload": (f"{nmesh} P{p} kappa-Poisson problems on {n}^3x6 meshes "
                         f"(test/cube.dat + PHG bisections) refined with PHG's bisection "
                         f"toward seeded corners (levels {[m[2] for m in meshes]}), "
                         f"mixed-precision PCG to rtol {rtol}"),
            "dofs": ndofs, "total_dofs": tot, "nnz": [A.info()[1] for A in As],
            "host_refine_s": round(t_mesh, 3), "gpu_assemble_s": round(t_asm, 3),
            "sequential_s": round(t_seq, 4), "streams_s": round(t_str, 4),
            "graph_s": round(t_gr, 4),
            "dofs_per_s_sequential": round(tot / t_seq, 1),
            "dofs_per_s_graph": round(tot / t_gr, 1),
            "speedup_graph_vs_sequential": round(t_seq / t_gr, 3),
            "its": [int(i) for i in its_seq], "its_graph": [int(i) for i in its_gr],
            "graph_solution_bitwise_equal_to_sequential": bool(same),
            "element_operator": {
                "build_s": round(t_eop_build, 3),
                "sequential_s": round(t_eseq, 4), "graph_s": round(t_egr, 4),
                "dofs_per_s_graph": round(tot / t_egr, 1),
                "speedup_vs_csr_graph": round(t_gr / t_egr, 3),
                "its": [int(i) for i in its_eseq], "its_graph": [int(i) for i in its_egr],
                "max_rel_diff_vs_csr_solution": dev,
                "block_kernel_bytes": [a for a, _ in elem_b],
                "csr_fp32_bytes_per_spmv": [A.format_bytes(A_fmt32(A), True) for A in As],
                "kernel_level_largest": kernel_level}}


def bench_batch(nmesh=8, n=16, p=3, rtol=1e-8, reps=3, seed=6, sigma=0.5):
    """Assembly + solve with new coefficients: the sequential path (per mesh
    phg_hip_assemble, phg_hip_mat_eliminate, the element operator's
    coefficients on the space's layout, phg_hip_pcg_reliable) against one
    phg_hip_batch_run (assembly and solve of all meshes in the captured
    graphs).  Each repetition draws new kappa; both legs start from the same
    host arrays."""
    import phg_amd as pa
    meshes = batch_meshes(nmesh, n)
    pb = pa.Problem.poisson(use_kappa=True)
    rng = np.random.default_rng(seed)
    draws = [[np.exp(rng.normal(0.0, sigma, size=g.nelem)) for g, _, _ in meshes]
             for _ in range(reps + 1)]
    spaces = [pa.Space(g, p) for g, _, _ in meshes]
    ones = [np.ones(g.ndof(p)) for g, _, _ in meshes]
    zeros = [np.zeros(g.ndof(p)) for g, _, _ in meshes]

    def sequential(ks):
        out = []
        for sp, k, f, z in zip(spaces, ks, ones, zeros):
            A = pa.assemble(sp, pb, kappa=k, f=f, g=z)
            A.eliminate()
            A.element_operator(sp, pb, kappa=k)
            x = pa.Vec(A.n)
            nit, _, _ = pa.pcg_reliable(A, A.rhs(), x, rtol=rtol, btol=rtol)
            out.append((A, x, nit))
        pa.synchronize()
        return out
    sequential(draws[0])            # warm: spaces keep their workspaces and layouts
    t = time.perf_counter()
    for r in range(reps):
        seq = sequential(draws[1 + r])
    t_seq = (time.perf_counter() - t) / reps
    x_seq = [x.numpy() for _, x, _ in seq]
    its_seq = [int(i) for _, _, i in seq]
    # the batch, made on matrices of the first draw
    t0 = time.perf_counter()
    kv = [pa.Vec(g.nelem, k) for (g, _, _), k in zip(meshes, draws[0])]
    fv = [pa.Vec(len(f), f) for f in ones]
    gv = [pa.Vec(len(z), z) for z in zeros]
    xv = [pa.Vec(len(z), z) for z in zeros]
    mats = []
    for sp, k, f, z in zip(spaces, draws[0], ones, zeros):
        A = pa.assemble(sp, pb, kappa=k, f=f, g=z)
        A.eliminate()
        A.element_operator(sp, pb, kappa=k)
        mats.append(A)
    B = pa.Batch([(sp, pb, A, k, f, g, x) for sp, A, k, f, g, x in
                  zip(spaces, mats, kv, fv, gv, xv)], rtol=rtol, btol=rtol)
    pa.synchronize()
    t_make = time.perf_counter() - t0
    B.run()
    pa.synchronize()
    t = time.perf_counter()
    for r in range(reps):
        for v, k in zip(kv, draws[1 + r]):
            v.upload(k)
        for x in xv:
            x.axpby(0.0, x, 0.0)
        its_b, _ = B.run()
    pa.synchronize()
    t_b = (time.perf_counter() - t) / reps
    dev = max(float(np.linalg.norm(x.numpy() - a) / np.linalg.norm(a))
              for a, x in zip(x_seq, xv))
    same_csr = all(all(np.array_equal(u, v) for u, v in zip(A.download(), A1.download()))
                   for A, (A1, _, _) in zip(mats, seq))
    tot = sum(m.n for m in mats)
    return {"sequential_assemble_solve_s": round(t_seq, 4),
            "batch_assemble_solve_s": round(t_b, 4),
            "speedup": round(t_seq / t_b, 3),
            "dofs_per_s_batch": round(tot / t_b, 1),
            "dofs_per_s_sequential": round(tot / t_seq, 1),
            "batch_make_s": round(t_make, 3),
            "its_sequential": its_seq, "its_batch": [int(i) for i in its_b],
            "max_rel_diff_vs_sequential_solution": dev,
            "csr_bitwise_equal_to_sequential_assembly": bool(same_csr),
            "nnz": [int(m.nnz) for m in mats]}


def main():
    import phg_amd as pa
    pa.init(0)
    if len(sys.argv) > 1 and sys.argv[1] == "batch":
        print(json.dumps(bench_batch()), flush=True)
        return 0
    print(json.dumps(bench()), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
