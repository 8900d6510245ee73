#!/usr/bin/env python3
"""Benchmark of the Louvain pass on one MI355X.

Workload (default): R-MAT scale 24, edge factor 32, (a,b,c,d) = (0.57, 0.19,
0.19, 0.05), seed 3, symmetrised, self-loops and duplicates removed, unit
weights, int32 columns, int64 community sums (exact integer mode).

A step is one level-0 Louvain pass (inner-loop iteration): binned move kernels,
community-weight update, modularity, commit-or-reject, with the graph resident
in HBM. W warm-up passes run from the singleton state, then K passes are timed.
value = directed CSR entries x K / wall time of the K passes (edges/s).

roofline: the dominant move-kernel launch of each timed pass, timed with HIP
events on the library's stream; algorithmic bytes per launch =
entries x 8 (column + neighbour community) + vertices x 32 (offsets, c, k,
a[c_v], c_new) + distinct (vertex, community) pairs x 8 (a[c_j]).

cpu_baseline: the CPU oracle (C++/OpenMP restatement of the same rule) timing
one level-0 pass of the same graph on the host cores.
"""
import argparse
import ctypes as C
import importlib.util
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "mg-alphagcd_amd")
HBM_PEAK_GBPS = 8000.0
METRIC = "edges/s per Louvain pass + HBM GB/s on 1 MI355X; batched vs plain graphs/s"


def load_pkg():
    spec = importlib.util.spec_from_file_location(
        "mg_alphagcd_amd", os.path.join(PKG, "__init__.py"), submodule_search_locations=[PKG])
    m = importlib.util.module_from_spec(spec)
    sys.modules["mg_alphagcd_amd"] = m
    spec.loader.exec_module(m)
    return m


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(g, passes):
    """The oracle timing `passes` level-0 passes from singletons (host cores)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    n = g.n
    k = orc.vertex_weights(g.offsets)
    c = np.arange(n, dtype=np.int32)
    a = k.copy()
    cn = np.empty(n, np.int32)
    P = lambda x, t: x.ctypes.data_as(C.POINTER(t))
    q = -1.0
    t0 = time.perf_counter()
    for i in range(passes):
        q_new = orc.lib.orc_pass_i64(n, P(g.offsets, C.c_int64), P(g.cols, C.c_int32), None,
                                     P(k, C.c_int64), P(c, C.c_int32), P(a, C.c_int64), g.nnz,
                                     1 if i % 2 == 0 else 0, 2, P(cn, C.c_int32), q, 1e-6)
        q = q_new
    dt = time.perf_counter() - t0
    return dict(value=g.nnz * passes / dt, unit="edges/s", cores=int(orc.lib.orc_num_threads()),
                kind="port",
                sample=f"{passes} level-0 pass(es) from singletons of the same graph "
                       f"({g.nnz} entries), CPU oracle C++/OpenMP, {cpu_model()}, "
                       f"{dt:.2f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=float, default=32)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--version", type=int, default=2)
    ap.add_argument("--no-fuse", action="store_true")
    ap.add_argument("--cpu-passes", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--full", action="store_true", help="also time the full multi-level run")
    ap.add_argument("--tuning-file", default=None)
    ap.add_argument("--stats-out", default=None, help="write per-pass statistics (JSON)")
    ap.add_argument("--membench", type=int, default=0,
                    help="also time a streaming copy of this many bytes (HBM calibration)")
    ap.add_argument("--autotune", default=None,
                    help="run the bin-layout autotuner on this graph, write the layout to "
                         "this path and time with it")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design note, not a code path)", file=sys.stderr)
        sys.exit(2)

    mgl = load_pkg()
    t0 = time.perf_counter()
    g = mgl.gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, seed=args.seed)
    log(f"graph: R-MAT scale {args.scale} ef {args.edge_factor}: n={g.n} nnz={g.nnz} "
        f"({time.perf_counter() - t0:.1f} s to generate)")

    ctx = mgl.Context(0)
    log(f"device LDS per workgroup: {ctx.lds_bytes()} bytes; bins: {ctx.bin_bounds()}")
    copy_gbps = None
    if args.membench:
        copy_gbps = ctx.membench(args.membench, 5)
        log(f"streaming copy of {args.membench} bytes: {copy_gbps:.1f} GB/s (read + write)")
    if args.autotune:
        t0 = time.perf_counter()
        ctx.autotune(g, args.autotune)
        args.tuning_file = args.autotune
        log(f"autotune ({time.perf_counter() - t0:.1f} s): " + open(args.autotune).read())
    sess = ctx.session(g, version=args.version, fuse_modularity=not args.no_fuse,
                       tuning_file=args.tuning_file)
    bins = sess.bins()
    log(f"bins: {bins}")
    q0 = sess.reset()
    log(f"Q0 = {q0:.9f}")
    for i in range(args.warmup):
        q, acc = sess.pass_()
        log(f"warmup pass {i + 1}: Q={q:.9f} accepted={acc} {sess.stats()['pass_ms']:.3f} ms")
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        q, acc = sess.pass_()
        stats.append(sess.stats())
    wall = time.perf_counter() - t0
    for i, s in enumerate(stats):
        log(f"pass {args.warmup + i + 1}: pass {s['pass_ms']:.3f} ms move {s['move_ms']:.3f} "
            f"update {s['update_ms']:.3f} mod {s['modularity_ms']:.3f} | dominant bin "
            f"{s['dominant_bin']} {s['dominant_ms']:.3f} ms ({s['dominant_vertices']} v, "
            f"{s['dominant_entries']} e) moved {s['moved']} distinct {s['distinct']} Q={s['q']:.9f}")
    if args.stats_out:
        with open(args.stats_out, "w") as f:
            json.dump(stats, f)
    ms_per_step = wall * 1e3 / args.steps
    value = g.nnz * args.steps / wall

    dom_ms = float(np.mean([s["dominant_ms"] for s in stats]))
    dom_bytes = float(np.mean([8 * s["dominant_entries"] + 32 * s["dominant_vertices"]
                               + 8 * s["dominant_distinct"] for s in stats]))
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    pass_bytes = float(np.mean([8 * s["nnz"] + 32 * s["n"] + 8 * s["distinct"]
                                + 16 * s["n"] + 16 * s["moved"] for s in stats]))
    move_ms = float(np.mean([s["move_ms"] for s in stats]))

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "move_rmat24_traffic.json")
    if os.path.exists(tfile):
        try:
            with open(tfile) as f:
                t = json.load(f)
            if t.get("bin") == stats[-1]["dominant_bin"]:
                traffic = t.get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

    full = None
    if args.full:
        t0 = time.perf_counter()
        res = ctx.louvain(g, version=args.version, fuse_modularity=not args.no_fuse)
        full = dict(seconds=time.perf_counter() - t0, levels=len(res.levels), q=res.final_q,
                    level_iters=[l.iters for l in res.levels])
        log(f"full run: {full}")

    sess.close()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(g, args.cpu_passes)
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {
            "workload": f"Louvain level-0 pass, R-MAT scale {args.scale} edge factor "
                        f"{args.edge_factor:g} (0.57,0.19,0.19,0.05) seed {args.seed}, "
                        "symmetrised, no self-loops/duplicates, unit weights",
            "vertices": g.n,
            "directed_entries": g.nnz,
            "version": args.version,
            "fused_modularity": not args.no_fuse,
            "precision": "exact int64 sums, FP64 gain",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"move bin {stats[-1]['dominant_bin']}",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": traffic,
        },
        "pass_algorithmic_GBps": pass_bytes / (ms_per_step * 1e-3) / 1e9,
        "move_ms": move_ms,
        "cpu_baseline": cpu,
    }
    if copy_gbps:
        out["copy_kernel_GBps"] = copy_gbps
    if full:
        out["full_run"] = full
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
