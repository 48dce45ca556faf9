"""GPU walker parity: the HIP walker's corpus, visit counts and relative
entropy against the CPU oracle driven by the same Philox streams.  Integer
outputs must be bit-exact; the FP64 relative entropy must be the same double."""

import os

import numpy as np
import pytest

import dgoracle
import distger
from distger import ffi

pytestmark = pytest.mark.gpu


def oracle_graph_from_ctx(ctx):
    """Oracle view of the device CSR (same neighbour order and CN values)."""
    rp, nbr, cn = ctx.graph_csr()
    g = dgoracle.OracleGraph.__new__(dgoracle.OracleGraph)
    L = dgoracle.load("dg")
    g.L = L
    g.v_num = len(rp) - 1
    g.row_ptr = rp
    g.nbr = nbr
    g.cn = cn
    g.m = len(nbr)
    g.data = cn.astype(np.float32)
    g._g = dgoracle._Graph(g.v_num, g.m, rp.ctypes.data, nbr.ctypes.data, cn.ctypes.data)
    return g


@pytest.fixture
def c1_ctx(ctx, c1_edges):
    """The shared context with the 10,000-vertex Chung-Lu graph, loaded for each test (other
    tests load their own graphs into the same context)."""
    print(f"[walker-gpu] loading the 10,000-vertex Chung-Lu graph: {len(c1_edges)} edges", flush=True)
    ctx.graph_from_edges(10000, c1_edges, undirected=True, cn_mode=distger.ffi.DG_CN_COMPUTE)
    return ctx


def test_device_csr_matches_oracle(c1_ctx, c1_edges):
    g = dgoracle.OracleGraph(10000, c1_edges)
    rp, nbr, cn = c1_ctx.graph_csr()
    assert np.array_equal(rp, g.row_ptr)
    assert np.array_equal(nbr, g.nbr[: g.m])
    assert np.array_equal(cn, g.cn[: g.m])


@pytest.mark.parametrize("minl,mu,cap,seed", [
    (20, 0.999, 80, 7),    # reference termination, max length 80
    (20, 0.999, 0, 8),     # no cap (reference)
    (3, 0.8, 200, 9),      # uneven lengths, LDS spill beyond 47 vertices
    (4, 0.7, 120, 10),     # many walks reach the cap
])
def test_round_bit_exact(c1_ctx, minl, mu, cap, seed):
    print(f"[walker-gpu] round parity minl={minl} mu={mu} cap={cap} seed={seed}", flush=True)
    g = oracle_graph_from_ctx(c1_ctx)
    W = 50000
    p = c1_ctx.default_walk_params(seed=seed, min_length=minl, max_length=cap, mu=mu)
    c1_ctx.walk_reset()
    st = c1_ctx.walk_round(p, 3, 0, W)
    tok, off = c1_ctx.corpus()
    otok, ooff = g.walk_round(g.cfg(seed, minl, cap, mu), 3, 0, W, threads=8,
                              tok_cap=W * max(cap, 64) + 16)
    assert np.array_equal(off, ooff)
    assert np.array_equal(tok, otok)
    assert st.tokens == len(otok) and st.steps == len(otok) - W
    assert st.max_walk_len == int(np.diff(ooff).max())
    freq = c1_ctx.vertex_freq()
    assert np.array_equal(freq, np.bincount(otok, minlength=10000).astype(np.uint32))
    h, status = g.relative_entropy(freq)
    assert status == 0
    assert st.re == h  # same double


def test_walker_begin_offset(c1_ctx):
    g = oracle_graph_from_ctx(c1_ctx)
    p = c1_ctx.default_walk_params(seed=3, min_length=3, max_length=200, mu=0.8)
    c1_ctx.walk_reset()
    # 777 walks leave most vertices unvisited: the corpus is produced, the
    # relative entropy is reported as undefined (the reference asserts there)
    with pytest.raises(distger.DgError) as ei:
        c1_ctx.walk_round(p, 2, 12345, 777)
    assert ei.value.code == -7
    tok, off = c1_ctx.corpus()
    otok, ooff = g.walk_round(g.cfg(3, 3, 200, 0.8), 2, 12345, 777, threads=4)
    assert np.array_equal(off, ooff) and np.array_equal(tok, otok)


def test_walk_run_rounds_match_oracle(c1_ctx, tmp_path):
    """The whole round loop (walk.hpp:608-782): 30 rounds, RE per round."""
    print("[walker-gpu] 30-round run", flush=True)
    g = oracle_graph_from_ctx(c1_ctx)
    p = c1_ctx.default_walk_params(seed=42, max_length=80)
    stats = c1_ctx.walk_run(p, keep_corpus=True)
    ref = dgoracle.walk_run(g, 42, 10000, init_round=5, max_length=80, threads=8)
    assert len(stats) == len(ref) == 30
    for s, r in zip(stats, ref):
        assert s.round == r["round"]
        assert s.tokens == len(r["tokens"])
        assert s.re == r["re"] and s.delta_re == r["delta_re"]
    tok, off = c1_ctx.corpus()  # last round
    assert np.array_equal(tok, ref[-1]["tokens"])
    nw, nt = c1_ctx.train_corpus_info()
    assert nw == 50000 + 29 * 10000 and nt == sum(len(r["tokens"]) for r in ref)
    # walk file bytes (path.hpp:67-119), id order and 3-thread segment order
    for threads in (1, 3):
        a = tmp_path / f"gpu-{threads}.txt"
        b = tmp_path / f"ora-{threads}.txt"
        c1_ctx.corpus_write_text(str(a), threads)
        dgoracle.write_walks(str(b), ref[-1]["tokens"], ref[-1]["offsets"], threads)
        assert a.read_bytes() == b.read_bytes()


def test_converge_mode_stops_on_delta(c1_ctx):
    g = oracle_graph_from_ctx(c1_ctx)
    p = c1_ctx.default_walk_params(seed=5, max_length=80, rounds_mode=1, delta=0.01)
    stats = c1_ctx.walk_run(p)
    ref = dgoracle.walk_run(g, 5, 10000, max_length=80, delta=0.01, converge=True, threads=8)
    assert [s.re for s in stats] == [r["re"] for r in ref]
    assert stats[-1].stop == 1 and len(stats) < 30


def test_degree_zero_vertex_is_an_error(ctx):
    e = np.zeros(3, distger.EDGE_DTYPE)
    e[0] = (0, 1, 0)
    e[1] = (1, 2, 0)
    e[2] = (2, 0, 0)
    ctx.graph_from_edges(4, e, cn_mode=distger.ffi.DG_CN_COMPUTE)  # vertex 3 isolated
    p = ctx.default_walk_params(seed=1, max_length=10)
    ctx.walk_reset()
    with pytest.raises(distger.DgError) as ei:
        ctx.walk_round(p, 1, 0, 8)
    assert ei.value.code == -7  # DG_ERR_RE_UNDEFINED
    tok, off = ctx.corpus()
    assert list(tok[off[3]:off[4]]) == [3]  # a walk from the isolated vertex is one vertex


def test_length_limit_is_an_error(ctx, c1_edges):
    ctx.graph_from_edges(10000, c1_edges, cn_mode=distger.ffi.DG_CN_COMPUTE)
    p = ctx.default_walk_params(seed=1, min_length=2, mu=0.6, max_length=0, length_limit=64)
    ctx.walk_reset()
    with pytest.raises(distger.DgError) as ei:
        ctx.walk_round(p, 1, 0, 5000)
    assert ei.value.code == -6  # DG_ERR_WALK_TOO_LONG


def test_bin_roundtrip_and_load(ctx, c1_edges, tmp_path):
    path = str(tmp_path / "g.bin")
    e = c1_edges.copy()
    g = dgoracle.OracleGraph(10000, e)
    # store the CN in the edge data, as gcn -> gconverter does for the reference
    cn_by_edge = {}
    for u in range(10000):
        for k in range(int(g.row_ptr[u]), int(g.row_ptr[u + 1])):
            cn_by_edge[(u, int(g.nbr[k]))] = int(g.cn[k])
    e["data"] = [cn_by_edge[(int(a), int(b))] for a, b in zip(e["src"], e["dst"])]
    distger.write_edges_bin(path, e)
    assert os.path.getsize(path) == 12 * len(e)
    ctx.graph_load_bin(path, 10000, undirected=True, cn_mode=distger.ffi.DG_CN_FROM_DATA)
    rp, nbr, cn = ctx.graph_csr()
    assert np.array_equal(cn, g.cn[: g.m]) and np.array_equal(nbr, g.nbr[: g.m])


def _hub_graph(n_hub_nbrs, v, extra, seed):
    rng = np.random.default_rng(seed)
    pairs = set((0, i) for i in range(1, n_hub_nbrs + 1))
    while len(pairs) < n_hub_nbrs + extra:
        a, b = rng.integers(0, v, 2)
        if a != b:
            pairs.add((int(min(a, b)), int(max(a, b))))
    pairs = sorted(pairs)
    rng.shuffle(pairs)
    e = np.zeros(len(pairs), distger.EDGE_DTYPE)
    e["src"] = [p[0] for p in pairs]
    e["dst"] = [p[1] for p in pairs]
    return e


@pytest.mark.parametrize("hub,v,extra", [(3000, 5000, 20000), (9000, 12000, 30000)])
def test_common_neighbours_with_hubs(ctx, hub, v, extra):
    """Hub rows above 1024 entries take the workgroup path; above 8192 the
    chunked one.  Counts must equal the oracle's sorted-merge counts."""
    e = _hub_graph(hub, v, extra, hub)
    ctx.graph_from_edges(v, e, cn_mode=distger.ffi.DG_CN_COMPUTE)
    g = dgoracle.OracleGraph(v, e)
    rp, nbr, cn = ctx.graph_csr()
    assert np.array_equal(nbr, g.nbr[: g.m])
    assert np.array_equal(cn, g.cn[: g.m])


@pytest.mark.parametrize("undirected", [True, False])
def test_common_neighbours_multigraph(ctx, undirected):
    """Duplicate edges and self-loops: every entry gets the distinct-neighbour
    count of its pair, |N(u) & N(v) \\ {u, v}|, as the oracle's sort-unique
    merge computes it (also on a 1500-neighbour hub)."""
    rng = np.random.default_rng(11)
    v = 3000
    pairs = [(0, int(w)) for w in rng.integers(1, v, 1500)]       # hub with repeats
    pairs += [(int(a), int(b)) for a, b in rng.integers(0, v, size=(12000, 2))]
    pairs += [(int(a), int(a)) for a in rng.integers(0, v, 40)]  # self-loops
    pairs += pairs[100:400]                                        # duplicated edges
    e = np.zeros(len(pairs), distger.EDGE_DTYPE)
    e["src"] = [p[0] for p in pairs]
    e["dst"] = [p[1] for p in pairs]
    ctx.graph_from_edges(v, e, undirected=undirected, cn_mode=distger.ffi.DG_CN_COMPUTE)
    g = dgoracle.OracleGraph(v, e, undirected=undirected)
    rp, nbr, cn = ctx.graph_csr()
    assert np.array_equal(nbr, g.nbr[: g.m])
    assert np.array_equal(cn, g.cn[: g.m])


def test_bounce_thresholds_definition(c1_ctx):
    """thr = min{u : tanh(pa) <= u / (2^31-1)} with the oracle's tanh."""
    L = dgoracle.load("dg")
    rp, nbr, cn = c1_ctx.graph_csr()
    thr = c1_ctx.graph_thresholds()
    deg = np.diff(rp).astype(np.int64)
    src = np.repeat(np.arange(len(deg)), deg)
    M = 2147483647.0
    rng = np.random.default_rng(0)
    for e in rng.choice(len(nbr), 4000, replace=False):
        sd, dd, c = int(deg[src[e]]), int(deg[nbr[e]]), int(cn[e])
        pa = (1.0 / float(sd - c)) * float(max(sd, dd)) / float(min(sd, dd))
        pn = L.dgo_tanh(pa)
        t = int(thr[e])
        assert pn <= t / M
        assert t == 0 or not (pn <= (t - 1) / M)


def test_device_math_matches_host(ctx):
    """The walker's FP64 functions round identically on the device and in the
    oracle (host): sqrt, log2, log and tanh on log-uniform and walker-range
    inputs, pow on the Correlation forms (S/(S+1))^S and ((f+1)/f)^f."""
    import ctypes
    L = dgoracle.load("dg")
    rng = np.random.default_rng(17)
    n = 200_000
    wide = np.exp(rng.uniform(np.log(1e-300), np.log(1e300), n))
    unit = rng.uniform(0.0, 4.0, n)
    cases = [
        (ffi.DG_MATH_SQRT, np.concatenate([wide[: n // 2], unit[: n // 2]]), None, L.dgo_sqrt),
        (ffi.DG_MATH_LOG2, np.concatenate([wide[: n // 2], unit[: n // 2] + 1e-9]), None, L.dgo_log2),
        (ffi.DG_MATH_LOG, np.concatenate([wide[: n // 2], unit[: n // 2] + 1e-9]), None, L.dgo_log),
        (ffi.DG_MATH_TANH, rng.uniform(0.0, 20.0, n), None, L.dgo_tanh),
    ]
    s = rng.integers(1, 100_000, n).astype(np.float64)
    cases.append((ffi.DG_MATH_POW, np.concatenate([s[: n // 2] / (s[: n // 2] + 1), (s[n // 2:] + 1) / s[n // 2:]]),
                  s, L.dgo_pow))
    for fn, x, y, host in cases:
        print(f"[walker-gpu] device math fn={fn}", flush=True)
        dev = ctx.math_eval(fn, x, y)
        if y is None:
            ref = np.array([host(ctypes.c_double(v)) for v in x])
        else:
            ref = np.array([host(ctypes.c_double(a), ctypes.c_double(b)) for a, b in zip(x, y)])
        bad = np.flatnonzero(dev.view(np.uint64) != ref.view(np.uint64))
        assert len(bad) == 0, f"fn {fn}: {len(bad)} of {len(x)} differ, e.g. x={x[bad[0]]!r}"


@pytest.fixture(scope="module")
def rmat_weighted():
    """R-MAT scale 12, edge factor 8, (0.57, 0.19, 0.19, 0.05), weights
    U(0, 1], seed 4 (the weighted-walk configuration at small scale)."""
    return distger.gen_rmat(12, 8, 0.57, 0.19, 0.19, 4)


def test_weighted_round_bit_exact(ctx, rmat_weighted):
    """Weighted walks (acceptance tanh(pa * w)) on the device against the
    oracle: corpus and visit counts bit for bit."""
    print("[walker-gpu] weighted R-MAT round parity", flush=True)
    V = 1 << 12
    ctx.graph_from_edges(V, rmat_weighted, undirected=True, cn_mode=ffi.DG_CN_COMPUTE_WEIGHTED)
    g = dgoracle.OracleGraph(V, rmat_weighted, weighted=True)
    rp, nbr, cn = ctx.graph_csr()
    assert np.array_equal(rp, g.row_ptr) and np.array_equal(nbr, g.nbr[: g.m])
    assert np.array_equal(cn, g.cn[: g.m])
    p = ctx.default_walk_params(seed=13)
    ctx.walk_reset()
    try:  # isolated R-MAT vertices are never visited: relative entropy undefined
        ctx.walk_round(p, 1, 0, 5 * V)
    except distger.DgError as e:
        assert e.code == -7
    tok, off = ctx.corpus()
    otok, ooff = g.walk_round(g.cfg(13, 20, 0, 0.999), 1, 0, 5 * V, threads=8)
    assert np.array_equal(off, ooff) and np.array_equal(tok, otok)
    assert np.array_equal(ctx.vertex_freq(), np.bincount(otok, minlength=V).astype(np.uint32))


def test_weighted_thresholds_definition(ctx, rmat_weighted):
    """thr = min{u : tanh(pa * w) <= u / (2^31-1)} with the oracle's tanh."""
    V = 1 << 12
    ctx.graph_from_edges(V, rmat_weighted, undirected=True, cn_mode=ffi.DG_CN_COMPUTE_WEIGHTED)
    g = dgoracle.OracleGraph(V, rmat_weighted, weighted=True)
    L = dgoracle.load("dg")
    rp, nbr, cn = ctx.graph_csr()
    thr = ctx.graph_thresholds()
    deg = np.diff(rp).astype(np.int64)
    src = np.repeat(np.arange(len(deg)), deg)
    M = 2147483647.0
    rng = np.random.default_rng(1)
    for e in rng.choice(len(nbr), 4000, replace=False):
        sd, dd, c = int(deg[src[e]]), int(deg[nbr[e]]), int(cn[e])
        pa = (1.0 / float(sd - c)) * float(max(sd, dd)) / float(min(sd, dd))
        pn = L.dgo_tanh(pa * float(g.w[e]))
        t = int(thr[e])
        assert pn <= t / M
        assert t == 0 or not (pn <= (t - 1) / M)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("scale", [20, 25])
def test_weighted_rmat_sampled_walks(ctx, scale):
    """Weighted walks on R-MAT with edge factor 16, (0.57, 0.19, 0.19, 0.05),
    weights U(0, 1], seed 4 (scale 25: 33,554,432 vertices, 1.05G entries): a
    round of 5 V walks; 300 sampled walks equal the oracle's, which walks the
    device's CSR (same neighbour order, the device's common-neighbour counts)
    with the weights placed in CSR order from the edge list; the visit counts
    add up to the corpus and every step follows an edge."""
    import time
    V = 1 << scale
    t0 = time.time()
    edges = distger.gen_rmat(scale, 16, 0.57, 0.19, 0.19, 4)
    print(f"[walker-gpu] weighted R-MAT scale {scale}: {len(edges)} edges ({time.time() - t0:.0f} s)",
          flush=True)
    ctx.graph_from_edges(V, edges, undirected=True, cn_mode=ffi.DG_CN_COMPUTE_WEIGHTED)
    rp_w, w = dgoracle.csr_edge_data(V, edges)
    del edges
    print(f"[walker-gpu] scale {scale}: graph resident ({time.time() - t0:.0f} s)", flush=True)
    p = ctx.default_walk_params(seed=4)
    ctx.walk_reset()
    try:
        ctx.walk_round(p, 1, 0, 5 * V)
    except distger.DgError as e:
        assert e.code == -7  # isolated vertices: relative entropy undefined
    tok, off = ctx.corpus()
    assert int(ctx.vertex_freq().sum(dtype=np.uint64)) == len(tok)
    rp, nbr, cn = ctx.graph_csr()
    assert np.array_equal(rp, rp_w)
    g = dgoracle.OracleGraph.from_csr(rp, nbr, cn, w)
    rng = np.random.default_rng(scale)
    cfg = g.cfg(4, 20, 0, 0.999)
    for wk in rng.integers(0, 5 * V, 300):
        path, _ = g.walk_one(cfg, 1, int(wk))
        dev = tok[off[wk]:off[wk + 1]]
        assert np.array_equal(path, dev), f"walk {wk} differs"
        for a, b in zip(dev[:-1], dev[1:]):  # consecutive vertices are adjacent
            assert b in nbr[rp[a]:rp[a + 1]]
    print(f"[walker-gpu] scale {scale}: 300 sampled walks equal ({time.time() - t0:.0f} s)", flush=True)
