"""GPU checks of the trainer forms and file formats added beside the plain and
batched forms:

* the DSGL form (include/dsgl.hpp:783-1427) in deterministic mode (one
  wavefront runs the compute threads one after another) against the FP64
  restatement of the same semantics (oracle/dsgl_oracle.c), which the CPU tests
  pin to an independent statement of the reference's thread loop; in Hogwild
  mode for sanity;
* the negative-reuse form (__sgNegReuse, src/word2vec.cu:231-303) bit for bit
  against its restatement in one-wavefront order;
* the vocabulary from walk visit counts (dsgl.hpp:314-350);
* the per-launch learning rates against dsgl.hpp:886-897's float expression;
* the embedding file (dsgl.hpp:1443-1461), text and binary, byte for byte.
"""

import struct

import numpy as np
import pytest

import dgsgns_oracle
import distger
from distger import ffi

pytestmark = pytest.mark.gpu

TABLE = 1_000_000


def log(msg):
    print(f"[forms-gpu] {msg}", flush=True)


@pytest.fixture(scope="module")
def walked(ctx, c1_edges):
    """The 10,000-vertex Chung-Lu graph and one round of 2000 walks (max length 80)."""
    ctx.graph_from_edges(10000, c1_edges, cn_mode=ffi.DG_CN_COMPUTE)
    p = ctx.default_walk_params(seed=6, max_length=80)
    ctx.walk_reset()
    try:
        ctx.walk_round(p, 1, 0, 2000)
    except distger.DgError as e:  # most vertices unvisited: relative entropy undefined
        assert e.code == -7
    tok, off = ctx.corpus()
    rp, _, _ = ctx.graph_csr()
    return tok, off, 2 * np.diff(rp).astype(np.uint64), ctx.vertex_freq()


def _params(ctx, **kw):
    base = dict(dim=128, window=10, negative=5, table_size=TABLE, alpha=0.025, sample=1e-3,
                alpha_schedule=ffi.DG_ALPHA_LINEAR, seed=13, deterministic=1)
    base.update(kw)
    return ctx.default_sgns_params(**base)


# ---- DSGL ---------------------------------------------------------------------

def _dsgl_oracle(counts, dim, tok, off, window, K, alpha, workers, seed, exact):
    o = dgsgns_oracle.Trainer(counts, dim, table_size=TABLE, sample=1e-3, float_init=True)
    o.dsgl_setup(sample=1e-3, table_size=TABLE)
    init0 = o.syn0.copy()
    o.dsgl(tok, off, window, K, alpha, workers, seed, exact=exact)
    return o, init0


@pytest.mark.parametrize("dim,workers,K,window", [(64, 1, 5, 10), (128, 7, 5, 10), (256, 3, 5, 10),
                                                 (128, 2, 10, 5)])
def test_dsgl_deterministic_vs_fp64_oracle(ctx, walked, dim, workers, K, window):
    """Exact sigmoid, every row the epoch moved within 1e-4 of the FP64
    restatement, relative to the largest update of its table (FP32 rows and
    MFMA accumulation against FP64).  Ten negatives take the 16 output rows
    of a step (up to six, eight rows are held)."""
    tok, off, counts, _ = walked
    p = _params(ctx, form=ffi.DG_SGNS_DSGL_FP32, dim=dim, exact_sigmoid=1, init_float_ops=1,
                workers=workers, negative=K, window=window)
    ctx.sgns_init(p)
    init0, _ = ctx.sgns_tables(dim)
    st = ctx.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
    s0, s1 = ctx.sgns_tables(dim)
    o, oinit0 = _dsgl_oracle(counts, dim, tok, off, window, K, 0.025, workers, 13, True)
    assert np.array_equal(ctx.sgns_vocab(), o.vertex_of_rank)
    assert np.array_equal(init0, oinit0)
    d0 = o.syn0.astype(np.float64) - init0
    e0 = np.abs((s0.astype(np.float64) - init0) - d0).max() / np.abs(d0).max()
    e1 = np.abs(s1.astype(np.float64) - o.syn1).max() / np.abs(o.syn1).max()
    log(f"DSGL dim {dim}, {workers} workers, K {K}, window {window}: {st.kept} kept of {st.words} words, max relative "
        f"error syn0 {e0:.3e} syn1 {e1:.3e}; rate at the end {st.alpha_end:.6f}")
    assert np.abs(o.syn1).max() > 0 and np.abs(d0).max() > 0
    assert e0 < 1e-4 and e1 < 1e-4


def test_dsgl_table_sigmoid_close_to_oracle(ctx, walked):
    """The reference's table sigmoid: an index computed in float on the
    device and in double in the oracle can fall on neighbouring entries for a
    few logits, so the match is looser than with the exact sigmoid."""
    tok, off, counts, _ = walked
    p = _params(ctx, form=ffi.DG_SGNS_DSGL_FP32, init_float_ops=1, workers=4)
    ctx.sgns_init(p)
    init0, _ = ctx.sgns_tables(128)
    ctx.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
    s0, s1 = ctx.sgns_tables(128)
    o, _ = _dsgl_oracle(counts, 128, tok, off, 10, 5, 0.025, 4, 13, False)
    d0 = o.syn0.astype(np.float64) - init0
    e0 = np.abs((s0.astype(np.float64) - init0) - d0).max() / np.abs(d0).max()
    e1 = np.abs(s1.astype(np.float64) - o.syn1).max() / np.abs(o.syn1).max()
    log(f"DSGL table sigmoid: max relative error syn0 {e0:.3e} syn1 {e1:.3e}")
    assert e0 < 2e-2 and e1 < 2e-2


def test_dsgl_hogwild_trains(ctx, walked):
    """64 compute threads on as many wavefronts, updates racing: finite
    tables, walk neighbours score above random pairs."""
    tok, off, counts, _ = walked
    p = _params(ctx, form=ffi.DG_SGNS_DSGL_FP32, deterministic=0, iter=3, init_float_ops=1,
                workers=64)
    ctx.sgns_init(p)
    st = ctx.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
    s0, s1 = ctx.sgns_tables(128)
    assert np.isfinite(s0).all() and np.isfinite(s1).all() and np.abs(s1).max() > 0
    assert st.words == 3 * len(tok) and st.launches == 3
    rng = np.random.default_rng(0)
    rnd = rng.choice(np.unique(tok), (5000, 2))
    a = distger.auc(ctx.sgns_score_pairs(tok[:5000], tok[1:5001]),
                    ctx.sgns_score_pairs(rnd[:, 0], rnd[:, 1]))
    log(f"DSGL Hogwild: {st.ms:.2f} ms for 3 epochs, co-occurrence AUC {a:.4f}")
    assert a > 0.6
    # the captured step rejects the DSGL form before it plans anything: the
    # round corpus stays as it was
    with pytest.raises(distger.DgError) as e:
        ctx.pipeline_step(ctx.default_walk_params(seed=1), 2, 0, 100, p)
    assert e.value.code == -8
    tok2, off2 = ctx.corpus()
    assert np.array_equal(tok2, tok) and np.array_equal(off2, off)


def test_dsgl_rejects_too_many_negatives(ctx, walked):
    with pytest.raises(distger.DgError) as e:
        ctx.sgns_init(_params(ctx, form=ffi.DG_SGNS_DSGL_FP32, negative=15))
    assert e.value.code == -8


# ---- negative reuse -------------------------------------------------------------

@pytest.mark.parametrize("dim", [64, 128])
def test_negreuse_bit_exact(ctx, walked, dim):
    tok, off, counts, _ = walked
    p = _params(ctx, form=ffi.DG_SGNS_NEGREUSE_FP32, dim=dim, alpha_schedule=ffi.DG_ALPHA_FIXED)
    ctx.sgns_init(p)
    st = ctx.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
    s0, s1 = ctx.sgns_tables(dim)
    o = dgsgns_oracle.Trainer(counts, dim, table_size=TABLE)
    sen, soff = o.subsample(tok, off, 13, 0)
    assert st.kept == len(sen)
    o.reuse(sen, soff, 10, 5, 0.025, 13, 0)
    assert np.array_equal(s1, o.syn1) and np.array_equal(s0, o.syn0)
    log(f"negative reuse dim {dim}: {len(sen)} kept words bit-exact")


# ---- vocabulary from visit counts ----------------------------------------------

def test_vocab_from_visit_counts(ctx, walked):
    """DG_VOCAB_VISITS ranks the vertices by their cumulative walk visits
    (vertex_cn, dsgl.hpp:314-350), ties by vertex id; the DSGL form keeps every
    vertex whatever min_count says, the others drop counts below it."""
    tok, off, counts, freq = walked
    p = _params(ctx, vocab_source=ffi.DG_VOCAB_VISITS, min_count=2)
    ctx.sgns_init(p)
    o = dgsgns_oracle.Trainer(freq.astype(np.uint64), 128, min_count=2, table_size=TABLE)
    assert np.array_equal(ctx.sgns_vocab(), o.vertex_of_rank)
    assert len(o.vertex_of_rank) < 10000
    p = _params(ctx, vocab_source=ffi.DG_VOCAB_VISITS, min_count=2, form=ffi.DG_SGNS_DSGL_FP32)
    ctx.sgns_init(p)
    v = ctx.sgns_vocab()
    o = dgsgns_oracle.Trainer(freq.astype(np.uint64), 128, min_count=0, table_size=TABLE)
    assert len(v) == 10000 and np.array_equal(v, o.vertex_of_rank)
    assert np.all(np.diff(freq[v].astype(np.int64)) <= 0)


# ---- learning rates ----------------------------------------------------------------

def _linear_alpha(alpha0, words, total):
    f = np.float32
    al = f(alpha0) * (f(1) - f(words) / f(total))
    return max(al, f(alpha0) * f(0.0001))


def test_per_launch_rates_follow_the_linear_schedule(ctx, walked):
    """alpha0 (1 - words / (iter * corpus_words + 1)) in float, clamped at
    alpha0 * 1e-4 (dsgl.hpp:886-897), at the first sentence of each launch:
    the batched form with 300 sentences per launch over 3 epochs, and the
    plain form's reference batches."""
    tok, off, counts, _ = walked
    n_tok = int(off[-1])
    p = _params(ctx, form=ffi.DG_SGNS_BATCHED_BF16, deterministic=0, minibatch=300, iter=3)
    ctx.sgns_init(p)
    st = ctx.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
    got = ctx.sgns_last_alphas()
    exp = [_linear_alpha(0.025, ep * n_tok + int(off[s]), 3 * n_tok + 1)
           for ep in range(3) for s in range(0, len(off) - 1, 300)]
    assert len(got) == st.launches == len(exp)
    assert np.array_equal(got, np.array(exp, np.float32))
    assert got[-1] == np.float32(st.alpha_end)
    p = _params(ctx, form=ffi.DG_SGNS_PLAIN_FP32, deterministic=0)
    ctx.sgns_init(p)
    ctx.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
    assert np.array_equal(ctx.sgns_last_alphas(), np.array([_linear_alpha(0.025, 0, n_tok + 1)], np.float32))


def test_captured_step_rates_match_direct_calls(ctx, walked):
    """The captured step evaluates the schedule on the device: the same floats
    as the host for the same launches."""
    p = _params(ctx, form=ffi.DG_SGNS_BATCHED_BF16, deterministic=0, minibatch=700, iter=2)
    wp = ctx.default_walk_params(seed=6, max_length=80)
    ctx.sgns_init(p)
    ctx.walk_reset()
    ctx.pipeline_step(wp, 1, 0, 2000, p, allow_undefined_re=True)
    got = ctx.sgns_last_alphas()
    tok, off = ctx.corpus()
    n_tok = int(off[-1])
    exp = [_linear_alpha(0.025, ep * n_tok + int(off[s]), 2 * n_tok + 1)
           for ep in range(2) for s in range(0, len(off) - 1, 700)]
    assert np.array_equal(got, np.array(exp, np.float32))


# ---- embedding file ----------------------------------------------------------------

def test_embedding_file_bytes(ctx, tmp_path):
    """dsgl.hpp:1443-1461: "%ld %ld\\n", then per rank "%u " (vertex id) and dim
    values "%f " (text) or dim raw floats (binary), then "\\n".  Values chosen to
    exercise %f rounding (ties, tiny, negative zero, large)."""
    rows, dim = 5, 64
    rng = np.random.default_rng(2)
    s0 = rng.standard_normal((rows, dim)).astype(np.float32)
    s0[0, :8] = np.array([0.0000005, -0.0000005, 0.0000015, -0.0, 1e6 + 0.5, -123.4567895, 2.5e-7, 0.1],
                         np.float32)
    s1 = np.zeros_like(s0)
    ctx.sgns_tables_create(s0, s1)
    txt, binp = str(tmp_path / "e.txt"), str(tmp_path / "e.bin")
    ctx.emb_write(txt)
    ctx.emb_write(binp, binary=True)
    exp = f"{rows} {dim}\n".encode()
    exp_b = exp
    for r in range(rows):
        exp += f"{r} ".encode() + "".join("%f " % float(x) for x in s0[r]).encode() + b"\n"
        exp_b += f"{r} ".encode() + struct.pack(f"<{dim}f", *s0[r]) + b"\n"
    assert open(txt, "rb").read() == exp
    assert open(binp, "rb").read() == exp_b
