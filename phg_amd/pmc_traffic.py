"""HBM traffic per dispatch from two rocprofv3 counter runs of bench.py.

    cd /tmp && export TMPDIR=/tmp
    rocprofv3 --pmc FETCH_SIZE -d OUT/pf -o run -- python3 -u REPO/bench.py \
        --steps 1 --warmup 0 --cpu-its 0 --no-profile
    rocprofv3 --pmc WRITE_SIZE -d OUT/pw -o run -- python3 -u REPO/bench.py \
        --steps 1 --warmup 0 --cpu-its 0 --no-profile
    python3 -m phg_amd.pmc_traffic OUT/pf/run_results.db OUT/pw/run_results.db \
        --nnz 57377401 --copy-bytes 365731296 > profiles/hbm_traffic.json

The counters are scaled by what they read for bench.py's copy kernel
(`k_copy`, `--copy-bytes` read and as many written per launch; the largest
`k_copy` dispatches, since the refinement solver also copies vectors).  Each
kernel's figure is the average over its dispatches that fetched more than
1 MiB (the working launches; slots of the reliable-update loop that only
read the solver state fetch almost nothing), stored with the digest of the
sources it was measured on, so bench.py reports it only for those sources.
"""
from __future__ import annotations

import argparse
import json
import os
import sqlite3
import sys
from collections import defaultdict

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

# bench.py's roofline keys: kernel-name prefix -> (key, sources)
KERNELS = {
    "void (anonymous namespace)::k_eop_block<2, false>": (
        "k_eop_block<2>", ["phg_amd/csrc/eop.hip", "phg_amd/csrc/phg_hip_internal.h"]),
    "void (anonymous namespace)::k_ru_spmv<(anonymous namespace)::XsViewT<1> >": (
        "k_ru_spmv<csr_xstage_pf>", ["phg_amd/csrc/krylov.hip", "phg_amd/csrc/dictcsr.hip"]),
    "void (anonymous namespace)::k_ru_spmv<(anonymous namespace)::XsViewT<0> >": (
        "k_ru_spmv<csr_xstage>", ["phg_amd/csrc/krylov.hip", "phg_amd/csrc/dictcsr.hip"]),
    "void (anonymous namespace)::k_pcg_spmv<double, (anonymous namespace)::XsViewT<1> >": (
        "k_pcg_spmv<double,csr_xstage_pf>",
        ["phg_amd/csrc/krylov.hip", "phg_amd/csrc/dictcsr.hip"]),
}


def per_dispatch(db, counter):
    con = sqlite3.connect(db)
    cols = [r[1] for r in con.execute("pragma table_info(counters_collection)")]
    ki, ci, vi, di = (cols.index(c) for c in ("kernel_name", "counter_name", "value",
                                              "dispatch_id"))
    tot, name = defaultdict(float), {}
    for r in con.execute("select * from counters_collection"):
        if r[ci] == counter:
            tot[r[di]] += r[vi]
            name[r[di]] = r[ki]
    out = defaultdict(list)
    for d, v in tot.items():
        out[name[d]].append(v)
    return out


def working_avg(vals):
    big = [v for v in vals if v > 1024.0]
    return (sum(big) / len(big), len(big)) if big else (0.0, 0)


def main(argv):
    ap = argparse.ArgumentParser()
    ap.add_argument("fetch_db")
    ap.add_argument("write_db")
    ap.add_argument("--nnz", type=int, required=True)
    ap.add_argument("--copy-bytes", type=int, required=True,
                    help="bytes k_copy reads (and writes) per launch")
    a = ap.parse_args(argv)
    from bench import source_digest
    fe, wr = per_dispatch(a.fetch_db, "FETCH_SIZE"), per_dispatch(a.write_db, "WRITE_SIZE")
    copy = next(k for k in fe if "k_copy" in k)
    cf, cw = max(fe[copy]), max(wr[copy])
    fscale = a.copy_bytes / (cf * 1024.0)
    wscale = a.copy_bytes / (cw * 1024.0)
    out = {"command": __doc__.split("\n\n")[1].strip(),
           "fetch_scale": round(fscale, 3), "write_scale": round(wscale, 3),
           "copy_kernel": {"fetch_kib": round(cf), "write_kib": round(cw),
                           "bytes_each_way": a.copy_bytes},
           "kernels": {}}
    for kname, (key, srcs) in KERNELS.items():
        match = [k for k in fe if k.startswith(kname)]
        if not match:
            continue
        f, nf = working_avg(fe[match[0]])
        w, _ = working_avg(wr.get(match[0], []))
        out["kernels"][key] = {"nnz": a.nnz, "dispatches": nf, "fetch_kib": round(f),
                               "write_kib": round(w),
                               "bytes": int(round((f * fscale + w * wscale) * 1024.0)),
                               "sources": srcs, "source_sha16": source_digest(srcs)}
    json.dump(out, sys.stdout, indent=1)
    print()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
