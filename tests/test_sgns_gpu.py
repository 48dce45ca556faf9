"""GPU SGNS trainer against the CPU oracle.

* set-up (vocabulary, LCG rows): bit-exact;
* plain FP32 form in its deterministic order (one wavefront, sentences in
  order): bit-exact against the oracle's restatement of skip_gram_kernel;
* batched forms in deterministic mode (one wavefront, sentences in order, the
  sequential order to first order in the updates, per-position shared
  negatives): against the FP64 restatement of that evaluation
  (dgo_sg_shared_first_order), with the tolerance written in each test;
  non-deterministic (Hogwild, atomic row updates) modes for sanity and AUC.
"""

import os

import numpy as np
import pytest

import dgoracle
import dgsgns_oracle
import distger
from distger import ffi

pytestmark = pytest.mark.gpu

TABLE = 1_000_000


@pytest.fixture(scope="module")
def walked(ctx, c1_edges):
    """The 10,000-vertex Chung-Lu graph in HBM and one round of 1000 walks (reference parameters)."""
    print("[sgns-gpu] walking 1000 walks on the 10,000-vertex Chung-Lu graph", flush=True)
    ctx.graph_from_edges(10000, c1_edges, cn_mode=ffi.DG_CN_COMPUTE)
    p = ctx.default_walk_params(seed=5, max_length=80)
    ctx.walk_reset()
    try:
        ctx.walk_round(p, 1, 0, 1000)
    except distger.DgError as e:  # most vertices unvisited: RE undefined, corpus fine
        assert e.code == -7
    tok, off = ctx.corpus()
    rp, _, _ = ctx.graph_csr()
    counts = 2 * np.diff(rp).astype(np.uint64)
    return tok, off, counts


def _params(ctx, **kw):
    base = dict(dim=128, window=10, negative=5, table_size=TABLE, alpha=0.025,
                alpha_schedule=ffi.DG_ALPHA_FIXED, seed=11)
    base.update(kw)
    return ctx.default_sgns_params(**base)


def test_init_matches_oracle(ctx, walked):
    tok, off, counts = walked
    p = _params(ctx)
    ctx.sgns_init(p)
    o = dgsgns_oracle.Trainer(counts, 128, table_size=TABLE)
    assert np.array_equal(ctx.sgns_vocab(), o.vertex_of_rank)
    s0, s1 = ctx.sgns_tables(128)
    assert np.array_equal(s0, o.syn0) and not s1.any()


def test_init_large_vocabulary_matches_oracle():
    """A vocabulary large enough that the LCG initialisation runs on several
    host threads from jump-ahead states: the rows equal the single-threaded
    sequence bit for bit."""
    V = 50000
    edges = distger.gen_chung_lu(V, 200000, 2.5, 2, 3)
    with distger.Context(0) as c:
        c.graph_from_edges(V, edges, undirected=True, cn_mode=ffi.DG_CN_COMPUTE)
        rp, _, _ = c.graph_csr()
        p = c.default_sgns_params(dim=128, table_size=TABLE)
        c.sgns_init(p)
        s0, s1 = c.sgns_tables(128)
    o = dgsgns_oracle.Trainer(2 * np.diff(rp).astype(np.uint64), 128, table_size=TABLE)
    assert s0.size > (1 << 22)
    assert np.array_equal(s0, o.syn0) and not s1.any()


def test_plain_deterministic_bit_exact(ctx, walked):
    print("[sgns-gpu] plain form, one wavefront", flush=True)
    tok, off, counts = walked
    p = _params(ctx, form=ffi.DG_SGNS_PLAIN_FP32, deterministic=1)
    ctx.sgns_init(p)
    st = ctx.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
    s0, s1 = ctx.sgns_tables(128)
    o = dgsgns_oracle.Trainer(counts, 128, table_size=TABLE)
    sen, soff = o.subsample(tok, off, 11, 0)
    assert st.kept == len(sen) and st.words == len(tok)
    o.plain(sen, soff, 10, 5, 0.025)
    assert np.array_equal(s1, o.syn1)
    assert np.array_equal(s0, o.syn0)


def _rel_err(a, b, ref):
    d = np.abs(a.astype(np.float64) - b.astype(np.float64)).max()
    return d / max(np.abs(ref).max(), 1e-30)


# measured: FP32 3.0e-7 to 6.4e-7, BF16 4.6e-3 to 7.3e-3 (operands rounded to
# 8 significant bits)
@pytest.mark.parametrize("form,dim,tol", [(ffi.DG_SGNS_BATCHED_FP32, 128, 1e-5),
                                          (ffi.DG_SGNS_BATCHED_BF16, 128, 1e-2),
                                          (ffi.DG_SGNS_BATCHED_FP32, 64, 1e-5),
                                          (ffi.DG_SGNS_BATCHED_BF16, 64, 1e-2),
                                          (ffi.DG_SGNS_BATCHED_BF16, 256, 1e-2)])
def test_batched_deterministic_vs_fp64_oracle(ctx, walked, form, dim, tol):
    """One epoch in one wavefront (sentences in order), exact sigmoid: the
    updated rows within `tol` of the FP64 restatement of the batched forms'
    evaluation (dgo_sg_shared_first_order), relative to the largest update.
    FP32: MFMA accumulation order only; BF16: operands rounded to BF16.
    That restatement is pinned to the sequential order by the CPU tests."""
    print(f"[sgns-gpu] batched form {form} dim {dim}", flush=True)
    tok, off, counts = walked
    p = _params(ctx, form=form, deterministic=1, exact_sigmoid=1, dim=dim)
    ctx.sgns_init(p)
    init0, _ = ctx.sgns_tables(dim)
    ctx.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
    s0, s1 = ctx.sgns_tables(dim)
    o = dgsgns_oracle.Trainer(counts, dim, table_size=TABLE)
    sen, soff = o.subsample(tok, off, 11, 0)
    o.shared_first_order(sen, soff, 10, 5, 0.025, True, 11, 0)
    d0 = o.syn0.astype(np.float64) - init0
    e0 = _rel_err(s0 - init0, o.syn0 - init0, d0)
    e1 = _rel_err(s1, o.syn1, o.syn1)
    print(f"[sgns-gpu] form {form} dim {dim}: max rel err syn0 {e0:.3e} syn1 {e1:.3e}", flush=True)
    assert np.abs(o.syn1).max() > 0
    assert e0 < tol and e1 < tol


@pytest.mark.parametrize("form", [ffi.DG_SGNS_PLAIN_FP32, ffi.DG_SGNS_BATCHED_FP32,
                                  ffi.DG_SGNS_BATCHED_BF16])
def test_hogwild_forms_train(ctx, walked, form):
    """Concurrent (atomic) updates: finite, moving tables, scores separate
    walk co-occurring pairs from random pairs."""
    tok, off, counts = walked
    p = _params(ctx, form=form, deterministic=0, alpha_schedule=ffi.DG_ALPHA_LINEAR, iter=3)
    ctx.sgns_init(p)
    st = ctx.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
    assert st.words == 3 * len(tok) and 0 < st.alpha_end < 0.025
    s0, s1 = ctx.sgns_tables(128)
    assert np.isfinite(s0).all() and np.isfinite(s1).all() and np.abs(s1).max() > 0
    # neighbours on walks vs random pairs
    a = tok[:-1][np.diff(off.astype(np.int64)).repeat(np.diff(off).astype(np.int64))[:-1] > 0]
    src = tok[:5000]
    dst = tok[1:5001]
    rng = np.random.default_rng(0)
    rnd = rng.choice(np.unique(tok), (5000, 2))
    pos = ctx.sgns_score_pairs(src, dst)
    neg = ctx.sgns_score_pairs(rnd[:, 0], rnd[:, 1])
    assert distger.auc(pos, neg) > 0.6


def test_score_pairs_and_embedding_file(ctx, walked, tmp_path):
    tok, off, counts = walked
    p = _params(ctx)
    ctx.sgns_init(p)
    ctx.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
    s0, _ = ctx.sgns_tables(128)
    vocab = ctx.sgns_vocab()
    rank = np.zeros(10000, np.int64)
    rank[vocab] = np.arange(len(vocab))
    u = np.array([vocab[0], vocab[5], vocab[77]], np.uint32)
    v = np.array([vocab[3], vocab[5], vocab[1]], np.uint32)
    sc = ctx.sgns_score_pairs(u, v)
    exp = (s0[rank[u]].astype(np.float64) * s0[rank[v]]).sum(1)
    assert np.allclose(sc, exp, rtol=1e-5, atol=1e-7)
    path = str(tmp_path / "emb.txt")
    ctx.emb_write(path)
    with open(path) as f:
        head = f.readline().split()
        first = f.readline().split()
    assert head == [str(len(vocab)), "128"]
    assert int(first[0]) == vocab[0] and len(first) == 129
    assert abs(float(first[1]) - s0[0, 0]) < 1e-6


@pytest.mark.parametrize("shape", [(1, 16), (1, 32), (2, 32), (4, 16), (4, 32)])
def test_launch_shapes_train_alike(ctx, walked, shape):
    """Wavefronts per workgroup and the pool split change only which
    wavefront trains which sentence: the co-occurrence separation and the
    table norms match the default shape's (Hogwild noise aside)."""
    tok, off, counts = walked
    rng = np.random.default_rng(0)
    rnd = rng.choice(np.unique(tok), (5000, 2))
    res = []
    for bw, split in ((2, 16), shape):
        p = _params(ctx, form=ffi.DG_SGNS_BATCHED_BF16, deterministic=0, iter=3,
                    block_waves=bw, pool_split=split)
        ctx.sgns_init(p)
        ctx.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
        s0, s1 = ctx.sgns_tables(128)
        assert np.isfinite(s0).all() and np.isfinite(s1).all()
        a = distger.auc(ctx.sgns_score_pairs(tok[:5000], tok[1:5001]),
                        ctx.sgns_score_pairs(rnd[:, 0], rnd[:, 1]))
        res.append((a, np.linalg.norm(s0), np.linalg.norm(s1)))
    (a0, n00, n10), (a1, n01, n11) = res
    print(f"[sgns-gpu] shape {shape}: auc {a1:.4f} vs {a0:.4f}, norms {n01:.4f}/{n00:.4f} "
          f"{n11:.4f}/{n10:.4f}", flush=True)
    assert abs(a1 - a0) < 0.02
    assert abs(n01 / n00 - 1) < 0.02 and abs(n11 / n10 - 1) < 0.02


def test_autotune_leaves_training_state_unchanged(ctx, walked):
    """Tuning trains the sample many times on saved tables: afterwards a
    deterministic pass gives the same bits as one on an untouched context
    state, and the chosen shape is one of the candidates."""
    p = _params(ctx, form=ffi.DG_SGNS_BATCHED_BF16, deterministic=1,
                alpha_schedule=ffi.DG_ALPHA_LINEAR)
    ctx.sgns_init(p)
    ctx.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
    ref0, ref1 = ctx.sgns_tables(128)
    ctx.sgns_init(p)
    before = ctx.sgns_tables(128)
    tuned, d_rate, t_rate = ctx.sgns_autotune(p, corpus=ffi.DG_CORPUS_ROUND, sample_walks=600)
    print(f"[sgns-gpu] autotune: minibatch {tuned.minibatch} block_waves {tuned.block_waves} "
          f"pool_split {tuned.pool_split}; {d_rate:.3e} -> {t_rate:.3e} words/s", flush=True)
    after = ctx.sgns_tables(128)
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    assert tuned.block_waves in (1, 2, 4) and tuned.pool_split in (16, 32)
    assert tuned.minibatch in (32768, 65536, 131072, 262144, 524288)
    assert d_rate > 0 and t_rate >= d_rate
    assert tuned.deterministic == 1 and tuned.dim == 128  # the rest is the caller's
    ctx.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
    a0, a1 = ctx.sgns_tables(128)
    assert np.array_equal(a0, ref0) and np.array_equal(a1, ref1)
    with pytest.raises(distger.DgError) as e:
        ctx.sgns_autotune(_params(ctx, form=ffi.DG_SGNS_PLAIN_FP32), corpus=ffi.DG_CORPUS_ROUND)
    assert e.value.code == -8
