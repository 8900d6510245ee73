#!/usr/bin/env python3
"""Mixed-precision form against the all-FP64 form on one weighted graph.

Workload (default): R-MAT scale 25, edge factor 16, (a,b,c,d) = (0.57, 0.19,
0.19, 0.05), seed 5, symmetrised, self-loops and duplicates removed,
lognormal(0, 1) edge weights stored as FP32.

Two full multi-level runs on the GPU with the graph resident in HBM:
  FP32_HASH : FP32 hash sums, FP64 for 2m, the sigma_tot term of the gain and
              the final modularity (an FP64 reduction over the input graph)
  FP64      : every sum in FP64
Reported: both final modularities, |dQ|, the fraction of vertices whose
community agrees after matching each FP32 community with its most frequent
FP64 counterpart (and the reverse), level-0 pass times and whole-run times.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "mg-alphagcd_amd")


def load_pkg():
    spec = importlib.util.spec_from_file_location(
        "mg_alphagcd_amd", os.path.join(PKG, "__init__.py"), submodule_search_locations=[PKG])
    m = importlib.util.module_from_spec(spec)
    sys.modules["mg_alphagcd_amd"] = m
    spec.loader.exec_module(m)
    return m


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def agreement(a, b):
    """Fraction of vertices whose community in a agrees with b after mapping
    each community of a to its most frequent counterpart in b."""
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    key = a * (int(b.max()) + 1) + b
    uk, cnt = np.unique(key, return_counts=True)
    ua = uk // (int(b.max()) + 1)
    # uk is sorted, so equal ua are contiguous: max count per group
    starts = np.flatnonzero(np.r_[True, ua[1:] != ua[:-1]])
    return float(np.maximum.reduceat(cnt, starts).sum()) / len(a)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=25)
    ap.add_argument("--edge-factor", type=float, default=16)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--passes", type=int, default=5, help="timed level-0 passes per mode")
    args = ap.parse_args()
    mgl = load_pkg()
    t0 = time.perf_counter()
    g = mgl.gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, seed=args.seed, weights=2)
    log(f"graph: R-MAT scale {args.scale} ef {args.edge_factor}: n={g.n} nnz={g.nnz} "
        f"({time.perf_counter() - t0:.1f} s)")
    ctx = mgl.Context(0)
    out = {"metric": "FP32-hash vs FP64 Louvain: modularity difference and agreement",
           "workload": f"R-MAT scale {args.scale} edge factor {args.edge_factor:g} "
                       f"(0.57,0.19,0.19,0.05) seed {args.seed}, lognormal(0,1) FP32 weights",
           "vertices": g.n, "directed_entries": g.nnz, "data": "synthetic"}
    runs = {}
    for name, prec in (("fp32_hash", mgl.PREC_FP32_HASH), ("fp64", mgl.PREC_FP64)):
        s = ctx.session(g, precision=prec)
        s.reset()
        s.pass_()
        ms = []
        for _ in range(args.passes):
            s.pass_()
            ms.append(s.stats()["pass_ms"])
        s.close()
        r = ctx.louvain(g, precision=prec)
        runs[name] = r
        out[name] = {"final_q": r.final_q, "levels": len(r.levels),
                     "level_iters": [lv.iters for lv in r.levels], "seconds": r.seconds,
                     "level0_pass_ms_median": float(np.median(ms)),
                     "level0_edges_per_s": g.nnz / (float(np.median(ms)) * 1e-3)}
        log(f"{name}: Q={r.final_q:.12f} levels={len(r.levels)} {r.seconds:.3f} s, "
            f"level-0 pass {np.median(ms):.3f} ms")
    a32, a64 = runs["fp32_hash"].final_assign, runs["fp64"].final_assign
    out["abs_dq"] = abs(runs["fp32_hash"].final_q - runs["fp64"].final_q)
    out["rel_dq"] = out["abs_dq"] / abs(runs["fp64"].final_q)
    out["agreement_fp32_to_fp64"] = agreement(a32, a64)
    out["agreement_fp64_to_fp32"] = agreement(a64, a32)
    out["identical_labels"] = float(np.mean(a32 == a64))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
