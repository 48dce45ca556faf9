"""Benchmark of the DistGER embedding hot path on one MI355X.

Workload (the 2,238,731-vertex configuration): undirected Chung-Lu graph with
2,238,731 vertices and ~29M undirected edges, degree exponent 2.2, seed 1;
10% of the edges held out (with as many uniform non-edges) for link
prediction; walks on the training graph with the reference defaults
(--min_L 20, mu 0.999, --min_R 5, no length cap); SGNS dim 128, window 10, 5
negatives, sample 1e-3, alpha 0.01 linear decay.

One step = one round of the embedding pipeline: V walks in one persistent
walker launch (plus corpus packing and the relative-entropy update), then the
BF16-MFMA batched SGNS trainer over that round's corpus.  `value` is corpus
words walked and trained per second over the timed steps (inputs resident in
HBM, timed region bracketed by device syncs).

After the timed steps, the plain FP32 form (word2vec.cu skip_gram_kernel
semantics) and the BF16-MFMA form are each trained from the same initial
tables on the same accumulated corpus, and both embeddings are scored by
link-prediction AUC on the held-out edges.

Prints one JSON line on stdout (progress on stderr).
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
from distger import ffi  # noqa: E402

V_C2 = 2238731
E_C2 = 29_000_000
HBM_GBS = 8000.0
BF16_TFLOPS = 2500.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def walker_bytes(stats_list):
    """Algorithmic bytes of the walker kernel by the per-step figure of the
    precomputed-common-neighbour data path (the reference's, walk.hpp:920):
    two dependent random 64-B transactions per walk step (the drawn adjacency
    unit and the candidate's row offsets) plus 4 B per emitted vertex."""
    tokens = sum(s.tokens for s in stats_list)
    steps = sum(s.steps for s in stats_list)
    return 2 * 64 * steps + 4 * tokens


def walker_transaction_bytes(stats_list):
    """Random 64-B transactions the kernel actually issues (a bounce-back
    append reads nothing): one per drawn adjacency unit and one per row-offset
    pair read (walk start and every move), plus 4 B per emitted vertex."""
    tokens = sum(s.tokens for s in stats_list)
    draws = sum(s.draws for s in stats_list)
    moves = sum(s.moves for s in stats_list)
    return 64 * (draws + moves) + 4 * tokens


def sgns_sent_bytes(kept, pool_rows, dim, negative):
    """Algorithmic HBM bytes of the sentence kernel: per position its K + 1
    output rows (target, negatives) are read and updated in place (FP32: 4 B
    read + 4 B read and 4 B written by the atomic add, per element); per
    sentence each distinct input row is read once and updated once the same
    way; per position 4 B token, 1 B window shrink and 4 B per negative id."""
    rows = (negative + 1) * kept + pool_rows
    return rows * dim * 12 + kept * (5 + 4 * negative)


def measured_profile(shape, dim):
    """From the committed rocprofv3 summary of the BF16 sentence kernel: HBM
    bytes per kept position (FETCH_SIZE and WRITE_SIZE over the timed
    launches, calibrated against a copy kernel) and the MFMA utilisation, for
    a launch shape and dim the summary was taken at; None where the summary
    is absent or holds no measurement at this shape and dim.  The summary
    ends in one block per measured shape, each opened by its `shape:` line."""
    path = os.path.join(REPO, "profiles", "sg_sent_kernel_bf16_c2_hbm_counters.txt")
    keys = ("shape", "dim", "traffic_bytes_per_kept_position", "mfma_util_pct")
    blocks = []
    try:
        with open(path) as f:
            for line in f:
                k, sep, v = line.partition(":")
                if not sep or k not in keys:
                    continue
                if k == "shape":
                    blocks.append({})
                if blocks:
                    blocks[-1][k] = v.strip()
    except OSError:
        return None
    want = f"{shape.get('minibatch')},{shape.get('block_waves')},{shape.get('pool_split')}"
    for b in blocks:
        if b.get("shape") != want or b.get("dim") != str(dim):
            continue
        try:
            return {"traffic_per_kept": float(b["traffic_bytes_per_kept_position"]),
                    "mfma_util_pct": float(b.get("mfma_util_pct", "nan"))}
        except (KeyError, ValueError):
            return None
    return None


def auc_of(ctx, pos, neg):
    ps = ctx.sgns_score_pairs(pos["src"], pos["dst"])
    ns = ctx.sgns_score_pairs(neg["src"], neg["dst"])
    return distger.auc(ps, ns)


def tuple_workload(ctx, iters=20, warmup=3):
    """Differentiable interface: minibatches of 65,536 (center, context, 5
    negatives) tuples, ids Zipf(1.1) over 2,238,731 rows, dim 128, rows
    N(0, 0.1^2).  One step = forward + backward + apply.  Algorithmic bytes
    per step: the unique rows U = U_in + U_out read once at storage
    precision s and the FP32 master rows read and written once,
    2 U d s + 2 U d 4."""
    rows, n, K, dim = 2238731, 65536, 5, 128
    t = distger.gen_zipf(rows, 1.1, n * (K + 2), 2024).reshape(n, K + 2)
    s0 = distger.gen_normal(rows * dim, 0.1, 99).reshape(rows, dim)
    s1 = distger.gen_normal(rows * dim, 0.1, 100).reshape(rows, dim)
    u_in = len(np.unique(t[:, 1]))
    u_out = len(np.unique(np.concatenate([t[:, :1], t[:, 2:]], axis=1)))
    ctx.sgns_tables_create(s0, s1)
    del s0, s1
    res = {"tuples": n, "negatives": K, "rows": rows, "dim": dim, "unique_in": u_in, "unique_out": u_out}
    # the update (apply) runs inside the segmented-sum launches ("seg")
    phases = ("tuple_fwd", "tuple_prep", "tuple_sort", "tuple_seg")
    for bf16 in (False, True):
        for _ in range(warmup):
            ctx.tuples_forward(t, K, bf16=bf16)
            ctx.tuples_backward(t, K, bf16=bf16, lr=0.0, apply=True)
        ctx.timing_reset()
        for _ in range(iters):
            ctx.tuples_forward(t, K, bf16=bf16)
            ctx.tuples_backward(t, K, bf16=bf16, lr=0.0, apply=True)
        us = {ph: 1e3 * ctx.timing(ph)[1] / iters for ph in phases}
        step_us = sum(us.values())
        alg = 2 * (u_in + u_out) * dim * (2 if bf16 else 4) + 2 * (u_in + u_out) * dim * 4
        gbs = alg / (step_us * 1e-6) / 1e9
        res["bf16" if bf16 else "fp32"] = {
            "step_us": step_us, "tuples_per_s": n / (step_us * 1e-6),
            "phase_us": {k.replace("tuple_", ""): v for k, v in us.items()},
            "algorithmic_bytes_per_step": alg, "achieved_gbs": gbs, "frac_hbm": gbs / HBM_GBS}
        log(f"tuples {'bf16' if bf16 else 'fp32'}: {step_us:.1f} us per forward+backward+apply "
            f"of {n} tuples ({n / (step_us * 1e-6):.3e} tuples/s)")
        # the same step as one call (dg_sgns_tuples_step: one gather pass for
        # the logits, loss and gradients)
        for _ in range(warmup):
            ctx.tuples_step(t, K, bf16=bf16, lr=0.0, apply=True)
        ctx.timing_reset()
        for _ in range(iters):
            ctx.tuples_step(t, K, bf16=bf16, lr=0.0, apply=True)
        fused_us = sum(1e3 * ctx.timing(ph)[1] / iters for ph in phases)
        res["bf16" if bf16 else "fp32"]["fused_step_us"] = fused_us
        log(f"tuples {'bf16' if bf16 else 'fp32'}: {fused_us:.1f} us per fused step "
            f"(dg_sgns_tuples_step)")
    return res


def rmat25(args):
    """Weighted information-oriented walks on R-MAT scale 25 (33,554,432
    vertices, edge factor 16, (0.57, 0.19, 0.19, 0.05), weights U(0, 1], seed
    4, deduplicated; acceptance tanh(pa * w)), then BF16-MFMA SGNS dim 128,
    window 10, 5 negatives on the last round's corpus.  One step = one round
    of V walks followed by training on it."""
    scale = 25
    V = 1 << scale
    t0 = time.time()
    edges = distger.gen_rmat(scale, 16, 0.57, 0.19, 0.19, 4)
    log(f"R-MAT scale {scale}: {len(edges)} edges ({time.time() - t0:.1f} s)")
    out = {}
    with distger.Context(0) as ctx:
        t0 = time.time()
        ctx.graph_from_edges(V, edges, undirected=True, cn_mode=ffi.DG_CN_COMPUTE_WEIGHTED)
        _, cn_ms = ctx.timing("cn")
        _, th_ms = ctx.timing("threshold")
        m = ctx.graph_info()[1]
        del edges
        log(f"graph resident: {m} entries, common neighbours {cn_ms:.0f} ms ({time.time() - t0:.1f} s)")
        wp = ctx.default_walk_params(seed=4)
        ctx.walk_reset()

        def walk(rnd, n):
            try:
                return ctx.walk_round(wp, rnd, 0, n)
            except distger.DgError as e:  # isolated vertices: relative entropy undefined
                if e.code != -7:
                    raise
                return None
        walk(1, 5 * V)
        bf = ctx.default_sgns_params(form=ffi.DG_SGNS_BATCHED_BF16, dim=128, window=10, negative=5,
                                     sample=1e-3, alpha=0.01, alpha_schedule=ffi.DG_ALPHA_LINEAR,
                                     table_size=args.table_size, seed=4)
        t0 = time.time()
        ctx.sgns_init(bf)
        log(f"SGNS tables for {V} vertices ({time.time() - t0:.1f} s)")
        shape = {}
        if not args.no_tune:
            t0 = time.time()
            bf, d_rate, t_rate = ctx.sgns_autotune(bf, corpus=ffi.DG_CORPUS_ROUND,
                                                   sample_walks=args.tune_walks)
            shape = {"minibatch": bf.minibatch, "block_waves": bf.block_waves,
                     "pool_split": bf.pool_split}
            log(f"autotune ({time.time() - t0:.1f} s): {shape}, {d_rate:.3e} -> {t_rate:.3e} words/s")

        def step(rnd):  # one round of V walks and its training, as one HIP graph
            return ctx.pipeline_step(wp, rnd, 0, V, bf, allow_undefined_re=True)
        step(2)  # warmup
        ctx.sync()
        ctx.timing_reset()
        words = 0
        t0 = time.perf_counter()
        for r in range(args.steps):
            _, st = step(3 + r)
            words += st.words
        ctx.sync()
        wall = time.perf_counter() - t0
        n_walk, walk_ms = ctx.timing("walk")
        _, bf_ms = ctx.timing("sgns_bf16")
        tok, off = ctx.corpus()
        steps_last = len(tok) - (len(off) - 1)
        log(f"{args.steps} rounds: walker {walk_ms / n_walk:.2f} ms/round, SGNS {bf_ms / args.steps:.1f} ms/round")
        walk_bytes = 2 * 64 * steps_last + 4 * len(tok)
        walk_gbs = walk_bytes / (walk_ms / n_walk * 1e-3) / 1e9
        resident = 8 * (V + 1) + 16 * m + 2 * V * 128 * 4 + 4 * args.table_size
        out = {"walk": {"walk_steps_per_s": steps_last / (walk_ms / n_walk * 1e-3),
                        "kernel_ms": walk_ms / n_walk, "algorithmic_bytes_per_launch": walk_bytes,
                        "achieved_gbs": walk_gbs, "frac_hbm": walk_gbs / HBM_GBS},
               "sgns_bf16_words_per_s": words / (bf_ms * 1e-3),
               "graph": {"vertices": V, "entries": m, "cn_ms": cn_ms, "threshold_ms": th_ms},
               "launch_shape": shape,
               "hbm_resident_bytes": resident, "hbm_resident_frac": resident / 288e9}
    return {
        "metric": "walk steps/s + SGNS words/s at equal AUC; BF16-MFMA vs FP32 plain, 1 GPU",
        "value": words / wall, "unit": "words/s (weighted walk + BF16-MFMA SGNS pipeline)",
        "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": "R-MAT scale 25 (33,554,432 vertices), edge factor 16, "
                               "(0.57,0.19,0.19,0.05), weights U(0,1], seed 4; per step one round "
                               "of V weighted walks + BF16-MFMA SGNS dim 128 w10 K5 on it, "
                               "captured as one HIP graph",
                   "global_batch": V, "parallelism": "single GPU"},
        **out}


def _held_out_auc(syn0, rank_of_vertex, pos, neg):
    """Link-prediction AUC of host tables: dot(syn0[u], syn0[v]) per pair."""
    def scores(e):
        out = np.zeros(len(e), np.float32)
        for i in range(0, len(e), 1 << 20):
            ru = rank_of_vertex[e["src"][i:i + (1 << 20)]].astype(np.int64)
            rv = rank_of_vertex[e["dst"][i:i + (1 << 20)]].astype(np.int64)
            out[i:i + len(ru)] = np.einsum("ij,ij->i", syn0[ru], syn0[rv])
        return out
    return distger.auc(scores(pos), scores(neg))


def cpu_baseline(ctx, rnd, V, pos, neg, threads, seed, table_size):
    """The reference's CPU path restated on the host's cores, on one full
    round of the workload: the walker oracle walks round `rnd` (V walks, the
    device's current round: the same corpus bit for bit) on `threads`
    threads, then DSGL, the reference's CPU trainer (include/dsgl.hpp:783-1427,
    restated in FP32 in oracle/dsgl_oracle.c, optimised build), trains that
    corpus with one compute thread per host thread (Hogwild on shared
    tables, as the reference's threads), vocabulary from the walk visit
    counts; its embeddings are scored on the held-out edges.  The GPU's DSGL
    form trains the same corpus from the same tables for comparison."""
    import dgoracle
    import dgsgns_oracle

    rp, nbr, cn = ctx.graph_csr()
    g = dgoracle.OracleGraph.from_csr(rp, nbr, cn)
    del nbr, cn
    t0 = time.time()
    tok, off = g.walk_round(g.cfg(seed, 20, 0, 0.999), rnd, 0, V, threads=threads, tok_cap=V * 24 + 4096)
    walk_s = time.time() - t0
    dtok, doff = ctx.corpus()
    same = bool(np.array_equal(dtok, tok) and np.array_equal(doff, off))
    del dtok, doff
    walk_rate = len(tok) / walk_s
    freq = ctx.vertex_freq().astype(np.uint64)
    tr = dgsgns_oracle.Trainer(freq, 128, table_size=table_size, sample=1e-3, float_init=True)
    tr.dsgl_setup(sample=1e-3, table_size=table_size)
    t0 = time.time()
    tr.dsgl_f32(tok, off, 10, 5, 0.01, threads, seed, threads=threads, baseline=True)
    sgns_s = time.time() - t0
    sgns_rate = len(tok) / sgns_s
    cpu_auc = _held_out_auc(tr.syn0, tr.rank_of_vertex, pos, neg)
    log(f"CPU: walk {walk_rate:.3e} vertices/s, DSGL {sgns_rate:.3e} words/s on {threads} threads, "
        f"AUC {cpu_auc:.4f}")
    # the GPU's DSGL form on the same round from the same initial tables
    p = ctx.default_sgns_params(form=ffi.DG_SGNS_DSGL_FP32, dim=128, window=10, negative=5,
                                sample=1e-3, alpha=0.01, alpha_schedule=ffi.DG_ALPHA_LINEAR,
                                table_size=table_size, seed=seed, vocab_source=ffi.DG_VOCAB_VISITS,
                                init_float_ops=1)
    ctx.sgns_init(p)
    st = ctx.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
    gpu_auc = auc_of(ctx, pos, neg)
    gpu_rate = st.words / (st.ms * 1e-3)
    log(f"GPU DSGL form on the same round: {gpu_rate:.3e} words/s, AUC {gpu_auc:.4f}")
    rate = 1.0 / (1.0 / walk_rate + 1.0 / sgns_rate)
    base = {"value": rate, "unit": "words/s", "cores": threads, "kind": "port",
            "sample": (f"one full round of the workload ({V} walks, {len(tok)} words): walker "
                       f"oracle on {threads} threads ({walk_rate:.3e} vertices/s, {walk_s:.1f} s), "
                       f"then DSGL (dsgl.hpp semantics, FP32, vectorised) with {threads} compute "
                       f"threads on {threads} host threads ({sgns_rate:.3e} words/s, {sgns_s:.1f} s); "
                       f"value = walk + train words/s of the two legs in sequence"),
            "walk_vertices_per_s": walk_rate, "dsgl_words_per_s": sgns_rate, "auc": cpu_auc,
            "corpus_equals_device_round": same}
    gpu = {"words_per_s": gpu_rate, "auc": gpu_auc, "kernel_ms": st.ms, "words": st.words,
           "speedup_vs_cpu_dsgl": gpu_rate / sgns_rate, "auc_gap_vs_cpu_dsgl": gpu_auc - cpu_auc}
    return base, gpu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--vertices", type=int, default=V_C2)
    ap.add_argument("--edges", type=int, default=E_C2)
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--table-size", type=int, default=100_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the plain-vs-BF16 AUC run")
    ap.add_argument("--workload", default="chung-lu-2.2m", choices=["chung-lu-2.2m", "rmat25"],
                    help="rmat25: weighted walks and BF16 SGNS on R-MAT scale 25")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shape instead of the autotuned one")
    ap.add_argument("--tune-walks", type=int, default=524288,
                    help="sentences of the round-1 corpus the autotuner times each shape on")
    ap.add_argument("--shape", default=None,
                    help="fixed batched launch shape MINIBATCH,BLOCK_WAVES,POOL_SPLIT instead of "
                         "the autotuned one (profiling runs)")
    ap.add_argument("--no-graph", action="store_true",
                    help="time direct calls (walk round, then SGNS) instead of the captured step")
    ap.add_argument("--direct-steps", type=int, default=2,
                    help="direct-call steps timed after the captured ones, for comparison")
    ap.add_argument("--no-tuples", action="store_true",
                    help="skip the differentiable tuple workload (65,536 Zipf tuples)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the walk -> train path runs on a single device, "
              "so --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "rmat25":
        print(json.dumps(rmat25(args)))
        return
    t0 = time.time()
    V = args.vertices
    edges = distger.gen_chung_lu(V, args.edges, 2.2, 1, 1)
    train, pos, neg = distger.lp_split(V, edges, 0.1, 1)
    log(f"graph: {V} vertices, {len(edges)} edges, train {len(train)}, held out {len(pos)} "
        f"+ {len(neg)} non-edges ({time.time() - t0:.1f} s)")

    sp = dict(dim=args.dim, window=10, negative=5, sample=1e-3, alpha=0.01,
              alpha_schedule=ffi.DG_ALPHA_LINEAR, table_size=args.table_size, seed=args.seed,
              min_count=0, iter=1)
    out = {}
    with distger.Context(0) as ctx:
        t0 = time.time()
        ctx.graph_from_edges(V, train, undirected=True, cn_mode=ffi.DG_CN_COMPUTE)
        _, cn_ms = ctx.timing("cn")
        log(f"graph resident in HBM ({time.time() - t0:.1f} s; common neighbours {cn_ms:.1f} ms)")
        rp, _, _ = ctx.graph_csr()
        counts = 2 * np.diff(rp).astype(np.uint64)
        wp = ctx.default_walk_params(seed=args.seed)
        ctx.walk_reset()
        ctx.train_corpus_clear()
        bf = ctx.default_sgns_params(form=ffi.DG_SGNS_BATCHED_BF16, **sp)
        t0 = time.time()
        ctx.sgns_init(bf)
        log(f"SGNS tables initialised ({time.time() - t0:.1f} s)")
        # round 1: V * init_round walks, trained once, untimed
        st1 = ctx.walk_round(wp, 1, 0, V * wp.init_round)
        ctx.corpus_append()
        ctx.sgns_train(bf, corpus=ffi.DG_CORPUS_ROUND)
        log(f"round 1: {st1.walkers} walks, {st1.tokens} words, walker {st1.walk_ms:.2f} ms")
        shape = {}
        if args.shape:
            mb, bw, split = (int(x) for x in args.shape.split(","))
            shape = {"minibatch": mb, "block_waves": bw, "pool_split": split}
            bf = ctx.default_sgns_params(form=ffi.DG_SGNS_BATCHED_BF16, **sp, **shape)
            out["launch_shape"] = dict(shape, source="--shape")
        elif not args.no_tune:
            t0 = time.time()
            tuned, d_rate, t_rate = ctx.sgns_autotune(bf, corpus=ffi.DG_CORPUS_ROUND,
                                                      sample_walks=args.tune_walks)
            shape = {"minibatch": tuned.minibatch, "block_waves": tuned.block_waves,
                     "pool_split": tuned.pool_split}
            out["autotune"] = dict(shape, default_words_per_s=d_rate, tuned_words_per_s=t_rate,
                                   sample_walks=min(args.tune_walks, st1.walkers),
                                   seconds=time.time() - t0)
            bf = tuned
            log(f"autotune ({time.time() - t0:.1f} s): {shape}, {d_rate:.3e} -> {t_rate:.3e} words/s")
        rnd = 2

        def step(r, graph):
            if graph:  # walk -> train as one HIP graph
                return ctx.pipeline_step(wp, r, 0, V, bf)
            return ctx.walk_round(wp, r, 0, V), ctx.sgns_train(bf, corpus=ffi.DG_CORPUS_ROUND)

        graph = not args.no_graph
        for _ in range(args.warmup):
            step(rnd, graph)
            ctx.corpus_append()
            rnd += 1
        ctx.sync()
        ctx.timing_reset()
        wstats, sstats = [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ws, ss = step(rnd, graph)
            wstats.append(ws)
            sstats.append(ss)
            rnd += 1
        ctx.sync()
        wall = time.perf_counter() - t0
        n_walk, walk_ms = ctx.timing("walk")
        _, pack_ms = ctx.timing("pack")
        _, re_ms = ctx.timing("re")
        n_bf, bf_ms = ctx.timing("sgns_bf16")
        _, prep_ms = ctx.timing("sgns_prep")
        words = sum(s.words for s in sstats)
        steps = sum(s.steps for s in wstats)
        value = words / wall
        log(f"timed: {args.steps} rounds, {words} words in {wall * 1e3:.1f} ms; walker "
            f"{walk_ms / n_walk:.2f} ms/round, SGNS BF16 {bf_ms / args.steps:.2f} ms/round")
        # walker roofline (per launch)
        walk_bytes = walker_bytes(wstats) / n_walk
        walk_gbs = walk_bytes / (walk_ms / n_walk * 1e-3) / 1e9
        walk_tbytes = walker_transaction_bytes(wstats) / n_walk
        walk_tgbs = walk_tbytes / (walk_ms / n_walk * 1e-3) / 1e9
        # SGNS BF16 roofline (per launch of the sentence kernels)
        kept = sum(s.kept for s in sstats)
        pool_rows = sum(s.pool_rows for s in sstats)
        bf_bytes = sgns_sent_bytes(kept, pool_rows, args.dim, 5) / n_bf
        bf_gbs = bf_bytes / (bf_ms / n_bf * 1e-3) / 1e9
        # the walker's fraction of HBM: the 64-B transactions it issues (one per
        # drawn adjacency unit, one per row-offset pair read; a bounce-back
        # append reads nothing) plus 4 B per emitted vertex, over its time; the
        # survey's per-step figure (two transactions for every step) is kept
        # beside it, labelled, as it overstates what the kernel reads
        out["walk"] = {"walk_steps_per_s": steps / (walk_ms * 1e-3),
                       "round_steps_per_s": steps / ((walk_ms + pack_ms + re_ms) * 1e-3),
                       "kernel_ms": walk_ms / n_walk, "pack_ms": pack_ms / n_walk,
                       "re_ms": re_ms / n_walk,
                       "bytes_per_launch": walk_tbytes,
                       "achieved_gbs": walk_tgbs, "frac_hbm": walk_tgbs / HBM_GBS,
                       "per_step_figure_bytes_per_launch": walk_bytes,
                       "per_step_figure_frac_hbm": walk_gbs / HBM_GBS}
        out["sgns_bf16_pipeline"] = {"words_per_s": words / ((bf_ms + prep_ms) * 1e-3),
                                     "kernel_ms_per_launch": bf_ms / n_bf, "launches": n_bf,
                                     "prep_ms_per_launch": prep_ms / n_bf,
                                     "kept": kept, "pool_rows": pool_rows,
                                     "bytes_per_launch": bf_bytes, "achieved_gbs": bf_gbs}

        if graph:
            inst, upd, launches = ctx.pipeline_info()
            cmp = {"captured_ms_per_step": wall * 1e3 / args.steps, "instantiations": inst,
                   "in_place_updates": upd, "graph_launches": launches,
                   "kernel_timing": "device wall-clock stamps inside the graph"}
            if args.direct_steps > 0:
                ctx.timing_reset()
                t1 = time.perf_counter()
                for _ in range(args.direct_steps):
                    step(rnd, False)
                    rnd += 1
                ctx.sync()
                cmp["direct_ms_per_step"] = (time.perf_counter() - t1) * 1e3 / args.direct_steps
                n_d, d_ms = ctx.timing("sgns_bf16")
                cmp["sg_sent_ms_per_launch_hip_events"] = d_ms / n_d
                cmp["sg_sent_ms_per_launch_stamps"] = bf_ms / n_bf
            out["walk_train_graph"] = cmp
            log(f"captured step {cmp['captured_ms_per_step']:.1f} ms vs direct calls "
                f"{cmp.get('direct_ms_per_step', float('nan')):.1f} ms; graph {inst} instantiation(s), "
                f"{upd} update(s)")

        if not args.no_compare:
            # plain FP32 vs BF16 MFMA on the same accumulated corpus (round 1 + warmup)
            nw, nt = ctx.train_corpus_info()
            res = {}
            forms = [("bf16_mfma", ffi.DG_SGNS_BATCHED_BF16), ("fp32_batched", ffi.DG_SGNS_BATCHED_FP32),
                     ("fp32_plain", ffi.DG_SGNS_PLAIN_FP32)]
            if args.dim > 128:  # the FP32 batched pools of larger dims exceed the register file
                forms = [f for f in forms if f[0] != "fp32_batched"]
            for name, form in forms:
                p = ctx.default_sgns_params(form=form, **sp,
                                            **(shape if form != ffi.DG_SGNS_PLAIN_FP32 else {}))
                ctx.sgns_init(p)
                ctx.timing_reset()
                t1 = time.perf_counter()
                st = ctx.sgns_train(p, corpus=ffi.DG_CORPUS_TRAIN)
                ctx.sync()
                w = time.perf_counter() - t1
                a = auc_of(ctx, pos, neg)
                res[name] = {"words_per_s": st.words / (st.ms * 1e-3), "wall_words_per_s": st.words / w,
                             "auc": a, "words": st.words, "kernel_ms": st.ms}
                log(f"{name}: {st.words} words, {st.words / (st.ms * 1e-3):.3e} words/s, AUC {a:.4f}")
            res["bf16_vs_plain_speedup"] = res["bf16_mfma"]["words_per_s"] / res["fp32_plain"]["words_per_s"]
            res["auc_gap_bf16_vs_plain"] = res["bf16_mfma"]["auc"] - res["fp32_plain"]["auc"]
            out["sgns_forms"] = res
            out["sgns_forms"]["corpus"] = f"{nw} walks, {nt} words (round 1 + {args.warmup} round(s))"
        cpu = None
        if not args.no_cpu:
            cpu, out["dsgl_gpu_same_corpus"] = cpu_baseline(
                ctx, rnd - 1, V, pos, neg, min(16, os.cpu_count() or 1), args.seed, args.table_size)
        copy_gbs = ctx.copy_bench(1 << 30, 3)
        out["copy_kernel_gbs"] = copy_gbs
        if not args.no_tuples:
            out["tuples"] = tuple_workload(ctx)

    shape_used = {"minibatch": bf.minibatch or 131072, "block_waves": bf.block_waves or 2,
                  "pool_split": bf.pool_split or 16}
    prof = measured_profile(shape_used, args.dim)
    line = {
        "metric": "walk steps/s + SGNS words/s at equal AUC; BF16-MFMA vs FP32 plain, 1 GPU",
        "value": value,
        "unit": "words/s (walk + BF16-MFMA SGNS pipeline)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": (f"Chung-Lu V={V} E~{args.edges} gamma=2.2 seed=1, 90% train; "
                                f"per step one walk round of V walks (min_L 20, mu 0.999) + SGNS "
                                f"BF16-MFMA dim {args.dim} w10 K5 sample 1e-3 on that round"
                                + ("" if args.no_graph else ", captured as one HIP graph")),
                   "global_batch": V, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm",
                     "kernel": (f"sg_sent_kernel<bf16, {args.dim}, 16|32, 8, "
                                f"{shape_used.get('block_waves')}>, "
                                f"{shape_used.get('minibatch')} sentences per launch"),
                     "achieved": bf_gbs, "peak": HBM_GBS, "unit": "GB/s",
                     "frac": bf_gbs / HBM_GBS,
                     "traffic": (prof["traffic_per_kept"] * kept / n_bf) if prof else None,
                     "mfma_util_pct": prof["mfma_util_pct"] if prof else None},
        "cpu_baseline": cpu,
    }
    line.update(out)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
