"""Benchmark of the MI355X finite-element hot path.

Workload (default): P2 variable-coefficient Poisson  -div(kappa grad u) = 1,
u = 0 on the boundary of the unit cube, on the 64^3 x 6 mesh that PHG's
refinement makes of test/cube.dat in 18 bisections (1,572,864 tetrahedra,
2,146,689 DOFs, PHG's numbering), kappa_e = exp(g_e) with g_e ~ N(0, 0.5^2)
i.i.d. per element in element-index order (seeded).  FP64 assembly on the GPU, then mixed-precision
Jacobi-PCG: FP32 SpMV and vectors, with the residual replaced by the FP64
residual b - A x (x accumulated in FP64) at reliable-update points
(--mixed reliable, the default), or restarted FP32 PCG solves under FP64
iterative refinement (--mixed refine).

A step is one solve from x0 = 0 to the stopping rule with the assembled system
already resident in HBM.  value = DOFs / (seconds per solve).  The same rule
is timed for the plain FP64 GPU PCG; every solution is checked against a
tightly converged FP64 reference solution (relative 2-norm error <= 1e-8).

Prints one JSON line (last line of stdout).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DOFs solved/s to 1e-8 rel err: mixed-prec vs FP64 PCG; SpMV HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(n, nnz, s_val, s_vec):
    """Algorithmic CSR SpMV bytes: values + int32 columns + int32 row pointers
    + one read of x + one write of y."""
    return nnz * (s_val + 4) + (n + 1) * 4 + n * s_vec + n * s_vec


def source_digest(files):
    """sha256 (16 hex digits) of the kernel sources a measurement was made on."""
    import hashlib
    h = hashlib.sha256()
    for f in files:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def measured_traffic(kernel_key, nnz):
    """HBM bytes per launch of a kernel from the committed PMC summary
    (profiles/hbm_traffic.json: 2 x FETCH_SIZE + WRITE_SIZE, scaled by a copy
    kernel of known size), when it was measured on this workload and on the
    kernel sources of this tree (their digest is stored with the numbers;
    after a change of those sources the figure is stale and None is
    returned)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel_key)
    if not k or k.get("nnz") != nnz:
        return None
    src = k.get("sources")
    if not src or source_digest(src) != k.get("source_sha16"):
        return None
    return k["bytes"]


def build_problem(n, p, sigma, seed):
    import phg_amd as pa
    t0 = time.perf_counter()
    # the mesh as PHG builds it: test/cube.dat bisected 3 log2(n) times
    g = pa.Grid.phg_cube(n) if n & (n - 1) == 0 else pa.Grid.kuhn(n)
    nd = g.ndof(p)
    bd = g.dof_bdry(p)
    rng = np.random.default_rng(seed)
    kappa = np.exp(rng.normal(0.0, sigma, size=g.nelem))
    f = np.ones(nd)
    gv = np.zeros(nd)
    t1 = time.perf_counter()
    return g, nd, bd, kappa, f, gv, t1 - t0


def cpu_baseline(g, p, kappa, its):
    """Oracle (sequential C restatement of PHG's assembly + Jacobi-PCG,
    solver-phg.c:485) on the benchmark's own system, one thread: the
    assembly and elimination in full, then a bounded sample of `its` PCG
    iterations from x0 = 0 (the solve to the stopping rule takes about a
    minute on one core; PCG iterations cost the same throughout)."""
    import oracle as orc
    nd = g.ndof(p)
    bd = g.dof_bdry(p)
    f, gv = np.ones(nd), np.zeros(nd)
    t0 = time.perf_counter()
    A = orc.assemble(orc.problem(p, use_kappa=True), g.verts(), g.elem_verts(),
                     g.dof_map(p), nd, bd, gv, f, kappa)
    orc.eliminate(A, bd)
    t1 = time.perf_counter()
    x, nits, res = orc.pcg(A, A.rhs, maxit=its, rtol=1e-30, btol=1e-30)
    t2 = time.perf_counter()
    return dict(ndof=nd, nelem=g.nelem, assemble_s=t1 - t0, sample_s=t2 - t1,
                sample_its=nits)


def oracle_reference_error(x, n, order, sigma, seed):
    """Relative 2-norm error of x against the oracle's FP64 solution of this
    system (tests/golden/c2_64_oracle.npz: the oracle's PCG to 1e-12 at a
    fixed sample of 4,096 DOFs, and the norm of the whole solution), when
    the workload is that system; None otherwise."""
    if (n, order, sigma, seed) != (64, 2, 0.5, 2024):
        return None
    ref = np.load(os.path.join(ROOT, "tests", "golden", "c2_64_oracle.npz"))
    idx, xs = ref["index"], ref["x"]
    return {"sampled_rel_err": float(np.linalg.norm(x[idx] - xs) / np.linalg.norm(xs)),
            "norm_rel_diff": float(abs(np.linalg.norm(x) - float(ref["norm2"])) /
                                   float(ref["norm2"])),
            "samples": int(len(idx)), "oracle_its": int(ref["nits"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=64, help="cubes per side")
    ap.add_argument("--order", type=int, default=2)
    ap.add_argument("--sigma", type=float, default=0.5)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--rtol", type=float, default=1e-8,
                    help="relative residual stopping rule of both solvers")
    ap.add_argument("--mixed", choices=("reliable", "refine"), default="reliable",
                    help="mixed-precision scheme: FP32 PCG with reliable FP64 "
                         "residual updates, or restarted FP32 PCG under FP64 "
                         "iterative refinement")
    ap.add_argument("--delta", type=float, default=0.1,
                    help="reliable-update threshold")
    ap.add_argument("--inner-rtol", type=float, default=1e-4)
    ap.add_argument("--cpu-its", type=int, default=150,
                    help="PCG iterations of the CPU baseline sample on the benchmark's "
                         "system (0: skip the CPU baseline)")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--operator", choices=("element", "csr"), default="element",
                    help="FP32 operator of the mixed solver: the matrix-free element "
                         "operator or the FP32 copy of the assembled CSR")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} "
              "requested); multi-GPU sharding is design-only.", file=sys.stderr)
        sys.exit(2)

    import phg_amd as pa
    pa.init(0)

    g, nd, bd, kappa, f, gv, t_mesh = build_problem(args.n, args.order, args.sigma,
                                                    args.seed)
    log(f"mesh {args.n}^3x6: {g.nelem} tets, {nd} DOFs (host setup {t_mesh:.2f}s)")
    sp = pa.Space(g, args.order, bdry=bd)
    pb = pa.Problem.poisson(use_kappa=True)
    pa.synchronize()
    t0 = time.perf_counter()
    A = pa.assemble(sp, pb, kappa=kappa, f=f, g=gv)
    A.eliminate()
    pa.synchronize()
    t_asm = time.perf_counter() - t0
    n, nnz = A.info()
    log(f"assembled: n={n} nnz={nnz} ({nnz / n:.1f}/row) in {t_asm:.3f}s")
    t0 = time.perf_counter()
    tune = A.tune(reps=20)
    log(f"SpMV layouts (us fp64/fp32): " + ", ".join(
        f"{k} {v[0]:.1f}/{v[1]:.1f}" for k, v in tune.items() if k != "chosen")
        + f"; chosen {tune['chosen']} ({time.perf_counter() - t0:.2f}s)")

    eop = None
    if args.operator == "element":
        t0 = time.perf_counter()
        A.element_operator(sp, pb, kappa=kappa)
        pa.synchronize()
        t_eop = time.perf_counter() - t0
        eb, fb = A.element_bytes()
        tb, tf = A.element_time(reps=20)
        eop = {"setup_s": round(t_eop, 4), "layout": A.element_layout(),
               "block_kernel": {"bytes": eb, "avg_us": round(tb, 2),
                                "GBps": round(eb / tb / 1e3, 1)},
               "fixup_kernel": {"bytes": fb, "avg_us": round(tf, 2),
                                "GBps": round(fb / tf / 1e3, 1)}}
        log(f"element operator: set-up {t_eop:.3f}s, {eop['layout']}, block kernel "
            f"{tb:.1f} us ({eb / tb / 1e3:.0f} GB/s), fix-up {tf:.1f} us")

    b = A.rhs()
    # FP64 reference solution (untimed, converged far below the target)
    xref = pa.Vec(n)
    nits_ref, res_ref = pa.pcg(A, b, xref, maxit=100000, rtol=1e-13, btol=1e-13)
    xr = xref.numpy()
    nxr = np.linalg.norm(xr)
    log(f"reference FP64 PCG: {nits_ref} its, residual {res_ref:.3e}")

    x = pa.Vec(n)

    def run_refine():
        x.axpby(0.0, x, 0.0)
        return pa.pcg_mixed(A, b, x, maxit=50, rtol=args.rtol, btol=args.rtol,
                            inner_rtol=args.inner_rtol)

    def run_reliable():
        x.axpby(0.0, x, 0.0)
        return pa.pcg_reliable(A, b, x, maxit=100000, rtol=args.rtol,
                               btol=args.rtol, delta=args.delta)

    run_mixed = run_reliable if args.mixed == "reliable" else run_refine
    run_other = run_refine if args.mixed == "reliable" else run_reliable

    def run_fp64():
        x.axpby(0.0, x, 0.0)
        return pa.pcg(A, b, x, maxit=100000, rtol=args.rtol, btol=args.rtol)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        pa.synchronize()
        t = time.perf_counter()
        out = None
        for _ in range(steps):
            out = fn()
        pa.synchronize()
        return (time.perf_counter() - t) / steps, out

    # mixed precision (the headline)
    sec_mp, out_mp = timed(run_mixed, args.steps, args.warmup)
    x_mp = x.numpy()
    err_mp = np.linalg.norm(x_mp - xr) / nxr
    err_oracle = oracle_reference_error(x_mp, args.n, args.order, args.sigma, args.seed)
    if err_oracle is not None:
        log(f"mixed vs the oracle's FP64 solution: sampled rel.err "
            f"{err_oracle['sampled_rel_err']:.2e}, |x| rel.diff {err_oracle['norm_rel_diff']:.2e}")
    if args.mixed == "reliable":
        inner, outer, res_mp = out_mp
    else:
        outer, inner, res_mp = out_mp
    log(f"mixed ({args.mixed}): {sec_mp * 1e3:.2f} ms/solve, {outer} FP64 residuals, "
        f"{inner} FP32 its, rel.err {err_mp:.2e}")
    # the other mixed-precision scheme, for comparison
    sec_o, out_o = timed(run_other, max(1, args.steps), 1)
    err_o = np.linalg.norm(x.numpy() - xr) / nxr
    log(f"mixed (other scheme): {sec_o * 1e3:.2f} ms/solve, {out_o}, rel.err {err_o:.2e}")
    # the same mixed solver on the FP32 copy of the assembled CSR
    csr_mixed = None
    if eop is not None:
        A.use_element_operator(False)
        sec_c, out_c = timed(run_mixed, max(1, args.steps), 1)
        err_c = np.linalg.norm(x.numpy() - xr) / nxr
        A.use_element_operator(True)
        csr_mixed = {"ms_per_solve": round(sec_c * 1e3, 3), "value": round(nd / sec_c, 1),
                     "fp32_its": out_c[0], "rel_err": float(f"{err_c:.3e}")}
        log(f"mixed on the FP32 CSR: {sec_c * 1e3:.2f} ms/solve, {out_c}, rel.err {err_c:.2e}")
    # plain FP64 GPU path with the same stopping rule
    sec_64, (nits64, res64) = timed(run_fp64, max(1, args.steps), 1)
    err_64 = np.linalg.norm(x.numpy() - xr) / nxr
    log(f"fp64:  {sec_64 * 1e3:.2f} ms/solve, {nits64} its, rel.err {err_64:.2e}")

    roof = None
    spmv64 = None
    if not args.no_profile:
        pa.profile(True)
        run_mixed()
        l32, ms32 = pa.profile_read(pa.K_SPMV32)
        pa.profile(True)
        run_fp64()
        l64, ms64 = pa.profile_read(pa.K_SPMV64)
        pa.profile(False)
        inv = {v: k for k, v in pa.FMT_NAMES.items()}
        f64, f32 = inv[tune["chosen"][0]], inv[tune["chosen"][1]]
        if l32 > 0 and eop is not None:
            # the FP32 operator apply in the solver: the element-block kernel
            # (interior and Dirichlet rows, partials of shared rows, (p,Ap));
            # k_ru_update sums the shared rows' partials as it reads Ap
            t32 = ms32 / l32 * 1e-3
            b32 = eop["block_kernel"]["bytes"]
            ach = b32 / t32 / 1e9
            roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic("k_eop_block<2>", nnz),
                    "kernel": "k_eop_block<2> (FP32 element operator apply + (p,Ap))",
                    "bytes_per_launch": b32, "avg_us": round(t32 * 1e6, 2),
                    "launches": l32,
                    "csr_equivalent_GBps": round(spmv_bytes(n, nnz, 4, 4) / t32 / 1e9, 1)}
        elif l32 > 0:
            t32 = ms32 / l32 * 1e-3
            b32 = A.format_bytes(f32, True)
            ach = b32 / t32 / 1e9
            tkey = (f"k_ru_spmv<{tune['chosen'][1]}>" if args.mixed == "reliable"
                    else f"k_pcg_spmv<float,{tune['chosen'][1]}>")
            roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic(tkey, nnz),
                    "kernel": (f"k_ru_spmv<{tune['chosen'][1]}> (FP32 SpMV + (p,Ap))"
                               if args.mixed == "reliable" else
                               f"k_pcg_spmv<float,{tune['chosen'][1]}> (FP32 SpMV + (p,Ap))"),
                    "bytes_per_launch": b32, "avg_us": round(t32 * 1e6, 2),
                    "launches": l32,
                    "csr_equivalent_GBps": round(spmv_bytes(n, nnz, 4, 4) / t32 / 1e9, 1)}
        if l64 > 0:
            t64 = ms64 / l64 * 1e-3
            b64 = A.format_bytes(f64, False)
            spmv64 = {"layout": tune["chosen"][0],
                      "achieved_GBps": round(b64 / t64 / 1e9, 1),
                      "csr_equivalent_GBps": round(spmv_bytes(n, nnz, 8, 8) / t64 / 1e9, 1),
                      "frac": round(b64 / t64 / 1e9 / HBM_PEAK_GBS, 4),
                      "avg_us": round(t64 * 1e6, 2), "bytes_per_launch": b64,
                      "launches": l64}

    # achievable HBM rate: a plain copy kernel moving the FP64 SpMV's bytes
    copy_n = spmv_bytes(n, nnz, 8, 8) // 16
    copy_us = pa.stream_copy(copy_n, 20)
    copy_gbs = 16 * copy_n / copy_us / 1e3
    log(f"copy kernel: {16 * copy_n / 1e6:.1f} MB in {copy_us:.1f} us = {copy_gbs:.0f} GB/s")

    # assembly + elimination again, warm (the first pass above loads the
    # kernels' code objects), for the assemble+solve rates
    pa.synchronize()
    t0 = time.perf_counter()
    A2 = pa.assemble(sp, pb, kappa=kappa, f=f, g=gv)
    A2.eliminate()
    pa.synchronize()
    t_asm_warm = time.perf_counter() - t0
    # the operator of a further matrix on the same mesh (a new kappa: time
    # steps, adjoint or finite-difference loops) takes the space's block
    # layout and only computes its coefficients
    t_setup_next = None
    if eop is not None:
        t0 = time.perf_counter()
        A2.element_operator(sp, pb, kappa=kappa)
        pa.synchronize()
        t_setup_next = time.perf_counter() - t0
    del A2
    t_setup = eop["setup_s"] if eop else 0.0
    # one solve per assembly: the first matrix of a mesh pays the layout
    # set-up, the next ones the coefficients only
    asm_solve = {"assemble_s": round(t_asm_warm, 4), "element_operator_setup_s": t_setup,
                 "mixed_dofs_per_s": round(nd / (t_asm_warm + t_setup + sec_mp), 1),
                 "element_operator_setup_next_s": (round(t_setup_next, 4)
                                                   if t_setup_next is not None else None),
                 "mixed_next_matrix_dofs_per_s": (
                     round(nd / (t_asm_warm + t_setup_next + sec_mp), 1)
                     if t_setup_next is not None else None),
                 "mixed_fp32_csr_dofs_per_s": (round(nd / (t_asm_warm + sec_c), 1)
                                               if csr_mixed else None),
                 "fp64_dofs_per_s": round(nd / (t_asm_warm + sec_64), 1)}
    log(f"assemble+solve: assembly {t_asm_warm * 1e3:.1f} ms (warm), operator set-up "
        f"{t_setup * 1e3:.1f} ms (first matrix of the mesh), "
        f"{(t_setup_next or 0) * 1e3:.1f} ms (next); mixed "
        f"{asm_solve['mixed_dofs_per_s'] / 1e6:.1f} M DOFs/s (next matrix "
        f"{(asm_solve['mixed_next_matrix_dofs_per_s'] or 0) / 1e6:.1f} M, FP32 CSR "
        f"{(asm_solve['mixed_fp32_csr_dofs_per_s'] or 0) / 1e6:.1f} M), "
        f"FP64 {asm_solve['fp64_dofs_per_s'] / 1e6:.1f} M DOFs/s")

    cpu = None
    if args.cpu_its > 0:
        c = cpu_baseline(g, args.order, kappa, args.cpu_its)
        # the solve to the stopping rule takes the FP64 iteration count (the
        # GPU FP64 PCG follows the oracle's recurrence, within 2 iterations)
        t_solve = c["sample_s"] / c["sample_its"] * nits64
        cpu = {"value": round(c["ndof"] / t_solve, 1), "unit": "DOFs/s",
               "cores": 1, "kind": "port",
               "sample": (f"oracle FP64 Jacobi-PCG (sequential C restatement of "
                          f"solver-phg.c:485) on this system ({c['ndof']} DOFs): "
                          f"{c['sample_its']} iterations in {c['sample_s']:.2f}s, "
                          f"scaled to the {nits64} iterations of rtol {args.rtol} "
                          f"({t_solve:.1f}s per solve); oracle assembly + elimination "
                          f"{c['assemble_s']:.2f}s"),
               "assemble_s": round(c["assemble_s"], 3)}
        log("cpu baseline:", cpu["sample"])

    value = nd / sec_mp
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "DOFs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec_mp * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": (f"P2 -div(kappa grad u)=1 on {args.n}^3x6 mesh "
                         + ("(test/cube.dat + PHG bisections) " if args.n & (args.n - 1) == 0
                            else "(structured Kuhn) ") +
                         f"({g.nelem} tets, {nd} DOFs), kappa=exp(N(0,{args.sigma}^2)) "
                         f"seed {args.seed}; " + (
                             f"FP32 Jacobi-PCG with reliable FP64 residual updates "
                             f"(delta {args.delta})" + (
                                 ", FP32 iterations on the matrix-free element operator "
                                 "(FP32 MFMA), FP64 residuals on the assembled CSR"
                                 if args.operator == "element" else
                                 ", FP32 iterations on the FP32 CSR")
                             if args.mixed == "reliable" else
                             f"FP32 Jacobi-PCG (inner rtol {args.inner_rtol}) under "
                             f"FP64 iterative refinement") +
                         f" to FP64 relative residual {args.rtol}"),
            "nnz": nnz, "parallelism": "single GPU",
        },
        "rel_err_vs_fp64": float(f"{err_mp:.3e}"),
        "rel_err_vs_oracle": err_oracle,
        "fp64_pcg": {"value": round(nd / sec_64, 1), "ms_per_solve": round(sec_64 * 1e3, 3),
                     "its": nits64, "rel_err": float(f"{err_64:.3e}")},
        "mixed_vs_fp64": round(sec_64 / sec_mp, 3),
        "mixed_detail": {"scheme": args.mixed, "fp64_residuals": outer,
                         "fp32_its": inner},
        "mixed_other_scheme": {"ms_per_solve": round(sec_o * 1e3, 3),
                               "value": round(nd / sec_o, 1),
                               "rel_err": float(f"{err_o:.3e}")},
        "assemble_s": round(t_asm, 4),
        "assemble_plus_solve": asm_solve,
        "spmv_layout_us": {k: v for k, v in tune.items() if k != "chosen"},
        "roofline": roof,
        "operator": args.operator,
        "element_operator": eop,
        "mixed_csr": csr_mixed,
        "spmv_fp64": spmv64,
        "copy_kernel": {"bytes": 16 * copy_n, "avg_us": round(copy_us, 2),
                        "GBps": round(copy_gbs, 1)},
        "cpu_baseline": cpu,
    }
    if err_mp > 1e-8:
        log(f"WARNING: mixed-precision error {err_mp:.2e} above 1e-8")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
