"""Benchmark of the DistGER embedding hot path on one MI355X.

Workload (2,238,731-vertex config): undirected Chung-Lu graph, 2,238,731
vertices, ~29M undirected edges, degree exponent 2.2, seed 1, 10% of the edges
held out for link prediction; the walker runs on the training graph with the
reference defaults (--min_L 20, mu 0.999, no length cap).

One step = one walk round of V walks (the size of every round after the
first, include/walk.hpp:764-775) in one persistent launch, followed by the
corpus packing and the relative-entropy update.  Inputs are resident in HBM
before the timed region; the timed region is bracketed by device syncs.

Prints one JSON line.
"""

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "gdistger-re_amd"))
sys.path.insert(0, os.path.join(REPO, "tests"))

import numpy as np  # noqa: E402

import distger  # noqa: E402

V_C2 = 2238731
E_C2 = 29_000_000
MI355X_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(args):
    t0 = time.time()
    edges = distger.gen_chung_lu(args.vertices, args.edges, 2.2, 1, 1)
    train, pos, neg = distger.lp_split(args.vertices, edges, 0.1, 1)
    log(f"graph: {args.vertices} vertices, {len(edges)} edges, train {len(train)}, "
        f"held out {len(pos)} + {len(neg)} non-edges ({time.time() - t0:.1f} s)")
    return train, pos, neg


def walker_bytes(stats_list):
    """Algorithmic HBM bytes of the walker kernel: 16 B per row-offset pair
    read (walk start and every move), 8 B per adjacency unit drawn, 4 B per
    vertex written, 8 B per visit-count update (read-modify-write of a
    uint32), 12 B per walk record (offset + length)."""
    walks = sum(s.walkers for s in stats_list)
    tokens = sum(s.tokens for s in stats_list)
    draws = sum(s.draws for s in stats_list)
    moves = sum(s.moves for s in stats_list)
    return 16 * moves + 8 * draws + 4 * tokens + 8 * tokens + 12 * walks


def cpu_baseline(ctx, seed, sample_walks, threads):
    import dgoracle

    rp, nbr, cn = ctx.graph_csr()
    g = dgoracle.OracleGraph.__new__(dgoracle.OracleGraph)
    g.L = dgoracle.load("dg")
    g.v_num = len(rp) - 1
    g.row_ptr, g.nbr, g.cn, g.m = rp, nbr, cn, len(nbr)
    g._g = dgoracle._Graph(g.v_num, g.m, rp.ctypes.data, nbr.ctypes.data, cn.ctypes.data)
    cfg = g.cfg(seed, 20, 0, 0.999)
    t0 = time.time()
    tok, off = g.walk_round(cfg, 2, 0, sample_walks, threads=threads)
    dt = time.time() - t0
    steps = len(tok) - sample_walks
    return steps / dt, dt, steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--vertices", type=int, default=V_C2)
    ap.add_argument("--edges", type=int, default=E_C2)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=400_000)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; multi-GPU runs are not part of this build "
              "(see DESIGN.md, 'How the path would shard')", file=sys.stderr)
        sys.exit(2)

    train, pos, neg = build_workload(args)
    V = args.vertices
    with distger.Context(0) as ctx:
        t0 = time.time()
        ctx.graph_from_edges(V, train, undirected=True, cn_mode=distger.ffi.DG_CN_COMPUTE)
        log(f"graph resident in HBM ({time.time() - t0:.1f} s incl. common neighbours)")
        p = ctx.default_walk_params(seed=args.seed)
        ctx.walk_reset()
        # round 1 (V * init_round walks) seeds the cumulative counts, untimed
        st1 = ctx.walk_round(p, 1, 0, V * p.init_round)
        log(f"round 1: {st1.walkers} walks, {st1.tokens} tokens, walk kernel {st1.walk_ms:.2f} ms")
        for w in range(args.warmup):
            ctx.walk_round(p, 2 + w, 0, V)
        ctx.sync()
        ctx.timing_reset()
        stats = []
        t0 = time.perf_counter()
        for k in range(args.steps):
            stats.append(ctx.walk_round(p, 2 + args.warmup + k, 0, V))
        ctx.sync()
        wall = time.perf_counter() - t0
        n_walk, walk_ms = ctx.timing("walk")
        _, pack_ms = ctx.timing("pack")
        _, re_ms = ctx.timing("re")
        steps = sum(s.steps for s in stats)
        value = steps / wall
        # algorithmic bytes of the dominant kernel, per launch
        per_launch_bytes = walker_bytes(stats) / max(n_walk, 1)
        avg_walk_s = walk_ms / 1e3 / max(n_walk, 1)
        achieved = per_launch_bytes / avg_walk_s / 1e9
        cpu = None
        if not args.no_cpu:
            threads = min(16, os.cpu_count() or 1)
            rate, dt, cs = cpu_baseline(ctx, args.seed, args.cpu_sample, threads)
            cpu = {"value": rate, "unit": "walk steps/s", "cores": threads, "kind": "port",
                   "sample": f"{args.cpu_sample} walks of one round on the same graph "
                             f"({cs} steps, {dt:.1f} s)"}
    out = {
        "metric": "walk steps/s + SGNS words/s at equal AUC; BF16-MFMA vs FP32 plain, 1 GPU",
        "value": value,
        "unit": "walk steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Chung-Lu V={V} E~{args.edges} gamma=2.2 seed=1, 90% train; "
                               f"walk rounds of V walks, min_L 20, mu 0.999",
                   "global_batch": V, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": MI355X_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / MI355X_HBM_GBS, "traffic": None},
        "kernels_ms": {"walk": walk_ms / max(n_walk, 1), "pack": pack_ms / max(n_walk, 1),
                       "re": re_ms / max(n_walk, 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
