"""Kernel-level driver for counter runs (rocprofv3 --pmc / --kernel-trace).

Assembles the P2 variable-coefficient Poisson system on an n^3 x 6 Kuhn mesh
(default 64: 2,146,689 DOFs), runs the SpMV layout tuner (every layout, FP64
and FP32, `--reps` launches each) and a plain copy of the same byte count as
the FP64 CSR SpMV, so that counter totals per kernel can be compared with
the algorithmic bytes.  Prints the per-launch algorithmic bytes.  With
--element-operator it also builds the matrix-free operator and launches its
block and fix-up kernels `--reps` times each.

    python -m phg_amd.profile_spmv --n 64 --reps 20 [--element-operator]
"""
import argparse
import json
import sys

import numpy as np


def _layout_bytes(pa, A):
    """Bytes per SpMV of the layouts still held by the matrix."""
    out = {}
    for f, name in pa.FMT_NAMES.items():
        try:
            out[name] = [A.format_bytes(f, False), A.format_bytes(f, True)]
        except pa.PhgHipError:
            pass
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--element-operator", action="store_true")
    a = ap.parse_args()
    import phg_amd as pa
    pa.init(0)
    g = pa.Grid.kuhn(a.n)
    nd = g.ndof(2)
    rng = np.random.default_rng(a.seed)
    kappa = np.exp(rng.normal(0.0, 0.5, size=g.nelem))
    sp = pa.Space(g, 2)
    pb = pa.Problem.poisson(use_kappa=True)
    A = pa.assemble(sp, pb, kappa=kappa, f=np.ones(nd), g=np.zeros(nd))
    A.eliminate()
    n, nnz = A.info()
    t = A.tune(reps=a.reps)
    b64 = nnz * 12 + (n + 1) * 4 + 16 * n
    b32 = nnz * 8 + (n + 1) * 4 + 8 * n
    copy_n = b64 // 16
    us_copy = pa.stream_copy(copy_n, a.reps)
    out = {"n": n, "nnz": nnz, "spmv_bytes_fp64": b64, "spmv_bytes_fp32": b32,
           "copy_doubles": copy_n, "copy_bytes": 16 * copy_n,
           "copy_us": us_copy, "copy_GBps": 16 * copy_n / us_copy / 1e3,
           "layouts_us": {k: v for k, v in t.items() if k != "chosen"},
           "layout_bytes": _layout_bytes(pa, A),
           "chosen": t["chosen"]}
    if a.element_operator:
        A.element_operator(sp, pb, kappa=kappa)
        tb, tf = A.element_time(reps=a.reps)
        eb, fb = A.element_bytes()
        out["element_operator"] = {"layout": A.element_layout(), "nelem": g.nelem,
                                   "block_bytes": eb, "block_us": tb,
                                   "fixup_bytes": fb, "fixup_us": tf}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
