#!/usr/bin/env python3
"""Batched form against the plain form on many independent graphs.

Workload: graphs i = 0..G-1 with vertex counts log-uniform in [1,000, 200,000]
(seed 4,000,000 + i). Even i: planted partition with 8-256 communities
(power-law sizes, exponent 2.1, min 8), mean intra / inter degree 16 / 3. Odd i:
R-MAT (0.57, 0.19, 0.19, 0.05) with edge factor uniform in [8, 32], scale
ceil(log2 n), vertices >= n dropped. Weights uniform (0.5, 1.5] as float.

Batched: mgl_louvain_batched over groups of --batch graphs (one launch sequence
per group, inner loops captured as hipGraphs). Plain: mgl_louvain per graph.
Both time the device run with the inputs already in HBM (the library's own
clock starts after upload). Every batched result is checked against the plain
one (levels, modularity, final assignment).
"""
import argparse
import importlib.util
import json
import math
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


def make_graph(mgl, i):
    rng = np.random.default_rng(4_000_000 + i)
    n = int(round(math.exp(rng.uniform(math.log(1_000), math.log(200_000)))))
    seed = 4_000_000 + i
    if i % 2 == 0:
        k = int(rng.integers(8, 257))
        k = max(1, min(k, n // 8))
        return mgl.gen_sbm(n, k, exponent=2.1, min_size=8, deg_in=16, deg_out=3, seed=seed,
                           weights=1)
    ef = float(rng.uniform(8, 32))
    scale = max(1, math.ceil(math.log2(n)))
    return mgl.gen_rmat(scale, ef, 0.57, 0.19, 0.19, seed=seed, weights=1, n_limit=n)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", type=int, default=4096)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--plain-sample", type=int, default=0,
                    help="time the plain form on every k-th graph only (0 = all)")
    args = ap.parse_args()
    mgl = load_pkg()
    ctx = mgl.Context(0)
    t0 = time.perf_counter()
    total_v = total_e = 0
    bat_s = plain_s = 0.0
    plain_n = 0
    mismatches = 0
    for b0 in range(0, args.graphs, args.batch):
        graphs = [make_graph(mgl, i) for i in range(b0, min(args.graphs, b0 + args.batch))]
        total_v += sum(g.n for g in graphs)
        total_e += sum(g.nnz for g in graphs)
        res = ctx.louvain_batched(graphs)
        bat_s += res[0].seconds
        for j, (g, r) in enumerate(zip(graphs, res)):
            if args.plain_sample and (b0 + j) % args.plain_sample:
                continue
            p = ctx.louvain(g)
            plain_s += p.seconds
            plain_n += 1
            if (p.final_q != r.final_q or len(p.levels) != len(r.levels)
                    or not np.array_equal(p.final_assign, r.final_assign)):
                mismatches += 1
        print(f"graphs {b0}..{b0 + len(graphs) - 1}: batched {res[0].seconds:.3f} s, "
              f"plain so far {plain_s:.3f} s over {plain_n} graphs "
              f"({time.perf_counter() - t0:.0f} s elapsed)", file=sys.stderr, flush=True)
    gps_b = args.graphs / bat_s
    gps_p = plain_n / plain_s
    out = {
        "metric": "graphs/s, batched vs plain Louvain",
        "value": gps_b,
        "unit": "graphs/s",
        "plain_graphs_per_s": gps_p,
        "speedup": gps_b / gps_p,
        "graphs": args.graphs,
        "batch": args.batch,
        "plain_timed_graphs": plain_n,
        "vertices": total_v,
        "directed_entries": total_e,
        "batched_seconds": bat_s,
        "plain_seconds": plain_s,
        "mismatches": mismatches,
        "precision": "exact (dyadic float weights)",
        "data": "synthetic",
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
