"""GPU parity at the sizes of the benchmarked workloads.

The 2,238,731-vertex Chung-Lu graph (gamma 2.2, seed 1, ~29M undirected edges,
10% held out for link prediction), walked with the reference parameters
(--min_L 20, mu 0.999, --min_R 5):

* walker: round 1 (5 V walks) and round 2 (V walks) bit for bit against the
  oracle on the device CSR: tokens, offsets, visit counts, and the relative
  entropy as the same double;
* SGNS on the first 300 walks of round 2, deterministic (one wavefront,
  sentences in order), the whole 2.2M-word vocabulary and the reference's
  1e8-entry unigram table: the plain FP32 form bit for bit against the
  oracle's restatement of skip_gram_kernel; the batched forms (FP32 at dim 128,
  BF16 at dims 128 and 256) within the tolerance written in each case of the
  FP64 restatement of their evaluation;
* the captured walk -> train step at the autotuned launch shape: its round
  equals the oracle's, its training moves the tables and separates held-out
  edges from non-edges.
"""

import numpy as np
import pytest

import dgoracle
import dgsgns_oracle
import distger
from distger import ffi

pytestmark = pytest.mark.gpu

V = 2238731
E = 29_000_000
TABLE = 100_000_000
ORACLE_THREADS = 16


def log(msg):
    print(f"[configs-gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def c2(ctx):
    """The 2.2M-vertex training graph in HBM and the oracle view of its CSR."""
    edges = distger.gen_chung_lu(V, E, 2.2, 1, 1)
    train, pos, neg = distger.lp_split(V, edges, 0.1, 1)
    del edges
    log(f"2.2M-vertex Chung-Lu graph: {len(train)} training edges, {len(pos)} held out")
    ctx.graph_from_edges(V, train, undirected=True, cn_mode=ffi.DG_CN_COMPUTE)
    rp, nbr, cn = ctx.graph_csr()
    g = dgoracle.OracleGraph.from_csr(rp, nbr, cn)
    return dict(g=g, pos=pos, neg=neg, counts=2 * np.diff(rp).astype(np.uint64))


def _walk(ctx, p, rnd, n):
    try:
        return ctx.walk_round(p, rnd, 0, n)
    except distger.DgError as e:  # unvisited vertices: relative entropy undefined
        assert e.code == -7
        return None


def test_c2_rounds_bit_exact(ctx, c2):
    g = c2["g"]
    p = ctx.default_walk_params(seed=1)
    ctx.walk_reset()
    freq = np.zeros(V, np.uint64)
    for rnd, n in ((1, 5 * V), (2, V)):
        log(f"round {rnd}: {n} walks on the device and in the oracle")
        st = ctx.walk_round(p, rnd, 0, n)
        tok, off = ctx.corpus()
        otok, ooff = g.walk_round(g.cfg(1, 20, 0, 0.999), rnd, 0, n, threads=ORACLE_THREADS,
                                  tok_cap=n * 24 + 4096)
        assert np.array_equal(off, ooff), f"round {rnd}: walk lengths differ"
        assert np.array_equal(tok, otok), f"round {rnd}: tokens differ"
        assert st.tokens == len(otok) and st.steps == len(otok) - n
        freq += np.bincount(otok, minlength=V).astype(np.uint64)
        dev_freq = ctx.vertex_freq()
        assert np.array_equal(dev_freq, freq.astype(np.uint32)), f"round {rnd}: visit counts differ"
        h, status = g.relative_entropy(dev_freq)
        assert status == 0
        assert st.re == h, f"round {rnd}: relative entropy {st.re!r} vs {h!r}"
        log(f"round {rnd}: {len(otok)} tokens, relative entropy {h!r} equal")


@pytest.fixture(scope="module")
def c2_walks(ctx, c2):
    """The first 300 walks of round 2 (walk ids 0..299) as the device corpus."""
    p = ctx.default_walk_params(seed=1)
    ctx.walk_reset()
    _walk(ctx, p, 2, 300)
    tok, off = ctx.corpus()
    return tok, off


def _params(ctx, **kw):
    base = dict(dim=128, window=10, negative=5, sample=1e-3, table_size=TABLE, alpha=0.025,
                alpha_schedule=ffi.DG_ALPHA_FIXED, seed=11, deterministic=1)
    base.update(kw)
    return ctx.default_sgns_params(**base)


def test_c2_plain_bit_exact(ctx, c2, c2_walks):
    tok, off = c2_walks
    p = _params(ctx, form=ffi.DG_SGNS_PLAIN_FP32)
    ctx.sgns_init(p)
    st = ctx.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
    s0, s1 = ctx.sgns_tables(128)
    o = dgsgns_oracle.Trainer(c2["counts"], 128, table_size=TABLE)
    assert np.array_equal(ctx.sgns_vocab(), o.vertex_of_rank)
    sen, soff = o.subsample(tok, off, 11, 0)
    assert st.kept == len(sen)
    o.plain(sen, soff, 10, 5, 0.025)
    assert np.array_equal(s1, o.syn1) and np.array_equal(s0, o.syn0)
    log(f"plain FP32 on 300 walks ({len(sen)} kept words) bit-exact")


# measured on these walks: FP32 3.3e-7, BF16 3.4e-3 (dim 128) and 3.8e-3 (dim 256)
@pytest.mark.parametrize("form,dim,tol", [
    (ffi.DG_SGNS_BATCHED_FP32, 128, 1e-5),
    (ffi.DG_SGNS_BATCHED_BF16, 128, 1e-2),
    (ffi.DG_SGNS_BATCHED_BF16, 256, 1e-2),
])
def test_c2_batched_vs_fp64_oracle(ctx, c2, c2_walks, form, dim, tol):
    """Max error over every row the epoch touched, relative to the largest
    update of that table.  FP32: MFMA accumulation order only; BF16: operands
    rounded to BF16 (about 8 significant bits)."""
    tok, off = c2_walks
    p = _params(ctx, form=form, dim=dim, exact_sigmoid=1)
    ctx.sgns_init(p)
    init0, _ = ctx.sgns_tables(dim)
    ctx.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
    s0, s1 = ctx.sgns_tables(dim)
    o = dgsgns_oracle.Trainer(c2["counts"], dim, table_size=TABLE)
    sen, soff = o.subsample(tok, off, 11, 0)
    o.shared_first_order(sen, soff, 10, 5, 0.025, True, 11, 0)
    d0 = o.syn0.astype(np.float64) - init0
    e0 = np.abs((s0.astype(np.float64) - init0) - d0).max() / np.abs(d0).max()
    e1 = np.abs(s1.astype(np.float64) - o.syn1).max() / np.abs(o.syn1).max()
    log(f"form {form} dim {dim}: max relative error syn0 {e0:.3e}, syn1 {e1:.3e} (tol {tol:g})")
    assert np.abs(o.syn1).max() > 0
    assert e0 < tol and e1 < tol


def test_c2_captured_step_at_tuned_shape(ctx, c2):
    """Round 1 trained directly, the launch shape autotuned on its corpus, then
    round 2 as one captured walk -> train graph (Hogwild, BF16): the round's
    corpus equals the oracle's, the tables stay finite and held-out edges score
    above non-edges."""
    g = c2["g"]
    wp = ctx.default_walk_params(seed=1)
    ctx.walk_reset()
    sp = ctx.default_sgns_params(form=ffi.DG_SGNS_BATCHED_BF16, dim=128, window=10, negative=5,
                                 sample=1e-3, alpha=0.01, alpha_schedule=ffi.DG_ALPHA_LINEAR,
                                 table_size=TABLE, seed=1)
    ctx.sgns_init(sp)
    ctx.walk_round(wp, 1, 0, 5 * V)
    ctx.sgns_train(sp, corpus=ffi.DG_CORPUS_ROUND)
    tuned, d_rate, t_rate = ctx.sgns_autotune(sp, corpus=ffi.DG_CORPUS_ROUND)
    log(f"autotune: minibatch {tuned.minibatch}, {tuned.block_waves} waves per workgroup, "
        f"pool split {tuned.pool_split}: {d_rate:.3e} -> {t_rate:.3e} words/s")
    rs, ss = ctx.pipeline_step(wp, 2, 0, V, tuned)
    tok, off = ctx.corpus()
    otok, ooff = g.walk_round(g.cfg(1, 20, 0, 0.999), 2, 0, V, threads=ORACLE_THREADS,
                              tok_cap=V * 24 + 4096)
    assert np.array_equal(off, ooff) and np.array_equal(tok, otok)
    assert ss.words == len(tok) and 0 < ss.kept <= ss.words
    s0, s1 = ctx.sgns_tables(128)
    assert np.isfinite(s0).all() and np.isfinite(s1).all() and np.abs(s1).max() > 0
    pos, neg = c2["pos"], c2["neg"]
    a = distger.auc(ctx.sgns_score_pairs(pos["src"], pos["dst"]),
                    ctx.sgns_score_pairs(neg["src"], neg["dst"]))
    log(f"captured step: {ss.words} words, AUC after two rounds {a:.4f}")
    assert a > 0.75
