"""The captured walk -> train step (dg_pipeline_step) against the same round
run by direct calls (dg_walk_round, then dg_sgns_train on the round).

Both paths run on contexts of their own from the same graph and seeds.  Per
round the corpus, the cumulative visit counts and the relative entropy must
be equal, and in deterministic mode the trained tables must have the same
bits.  The graph is replayed over several rounds, with direct work on the
device (corpus and table reads) between the replays, and each replay is
checked.  A round longer than the walker's token buffer exercises the
device-side guard and the second run of the step.
"""

import numpy as np
import pytest

import distger
from distger import ffi

pytestmark = pytest.mark.gpu

TABLE = 1_000_000


@pytest.fixture(scope="module")
def small_graph():
    """Chung-Lu, 2,000 vertices, ~10k edges, gamma 2.5, min degree 2, seed 3."""
    return distger.gen_chung_lu(2000, 10000, 2.5, 2, 3)


def _open(edges, v_num):
    c = distger.Context(0)
    c.graph_from_edges(v_num, edges, cn_mode=ffi.DG_CN_COMPUTE)
    return c


def _sgns(c, form, deterministic, **kw):
    base = dict(dim=128, window=10, negative=5, table_size=TABLE, alpha=0.025,
                alpha_schedule=ffi.DG_ALPHA_LINEAR, seed=11, form=form,
                deterministic=deterministic, minibatch=700)
    base.update(kw)
    return c.default_sgns_params(**base)


def _state(c, dim):
    tok, off = c.corpus()
    s0, s1 = c.sgns_tables(dim)
    return tok, off, c.vertex_freq(), s0, s1


def _run_both(edges, v_num, wkw, form, deterministic, rounds, skw=None):
    skw = skw or {}
    eager, graph = _open(edges, v_num), _open(edges, v_num)
    try:
        wp_e = eager.default_walk_params(seed=7, **wkw)
        wp_g = graph.default_walk_params(seed=7, **wkw)
        sp_e = _sgns(eager, form, deterministic, **skw)
        sp_g = _sgns(graph, form, deterministic, **skw)
        for c, sp in ((eager, sp_e), (graph, sp_g)):
            c.sgns_init(sp)
            c.walk_reset()
        out = []
        for r, (nw, alpha) in enumerate(rounds, start=1):
            print(f"[pipeline-gpu] form={form} det={deterministic} round {r}: {nw} walks", flush=True)
            rs_e = eager.walk_round(wp_e, r, 0, nw)
            ss_e = eager.sgns_train(sp_e, corpus=ffi.DG_CORPUS_ROUND, alpha=alpha)
            rs_g, ss_g = graph.pipeline_step(wp_g, r, 0, nw, sp_g, alpha)
            a, b = _state(eager, sp_e.dim), _state(graph, sp_g.dim)
            out.append((rs_e, ss_e, rs_g, ss_g, a, b))
        info = graph.pipeline_info()
        return out, info
    finally:
        eager.close()
        graph.close()


def _check_round(rs_e, ss_e, rs_g, ss_g, a, b, tables_equal):
    tok_e, off_e, f_e, s0_e, s1_e = a
    tok_g, off_g, f_g, s0_g, s1_g = b
    assert np.array_equal(tok_e, tok_g) and np.array_equal(off_e, off_g)
    assert np.array_equal(f_e, f_g)
    assert rs_e.tokens == rs_g.tokens and rs_e.max_walk_len == rs_g.max_walk_len
    assert rs_e.draws == rs_g.draws and rs_e.moves == rs_g.moves
    assert rs_e.re == rs_g.re and rs_e.delta_re == rs_g.delta_re
    assert ss_e.words == ss_g.words and ss_e.kept == ss_g.kept
    assert ss_e.sentences == ss_g.sentences and ss_e.launches == ss_g.launches
    assert ss_e.pool_rows == ss_g.pool_rows
    assert ss_e.alpha_end == ss_g.alpha_end
    if tables_equal:
        assert np.array_equal(s0_e, s0_g) and np.array_equal(s1_e, s1_g)
    else:
        assert np.isfinite(s0_g).all() and np.isfinite(s1_g).all() and s1_g.any()


@pytest.mark.parametrize("form", [ffi.DG_SGNS_BATCHED_FP32, ffi.DG_SGNS_BATCHED_BF16])
def test_captured_step_equals_direct_calls_deterministic(small_graph, form):
    rounds = [(2000, 0.025), (2000, 0.0226), (2000, 0.0205), (2000, 0.0185)]
    out, info = _run_both(small_graph, 2000, {}, form, 1, rounds, {"iter": 2})
    for rec in out:
        _check_round(*rec, tables_equal=True)
    inst, upd, launches = info
    print(f"[pipeline-gpu] graph: {inst} instantiations, {upd} updates, {launches} launches", flush=True)
    assert launches == len(rounds) and inst + upd == len(rounds) and inst >= 1


def test_captured_step_hogwild_matches_walks(small_graph):
    rounds = [(10000, 0.025), (2000, 0.0226), (2000, 0.0205)]
    out, _ = _run_both(small_graph, 2000, {}, ffi.DG_SGNS_BATCHED_BF16, 0, rounds)
    for rec in out:
        _check_round(*rec, tables_equal=False)


def test_captured_step_grows_token_buffer(small_graph):
    """min_length 38 and max_length 40: every walk has 40 vertices, more than
    its slot holds, so the first run of the step overflows the token buffer,
    the guard empties the round, and the step runs again with a larger
    buffer."""
    rounds = [(2000, 0.025), (2000, 0.02)]
    out, info = _run_both(small_graph, 2000, {"min_length": 38, "max_length": 40},
                          ffi.DG_SGNS_BATCHED_FP32, 1, rounds)
    for rec in out:
        _check_round(*rec, tables_equal=True)
        assert rec[0].max_walk_len == 40
    assert info[2] == len(rounds) + 1  # one extra run, in the first round


def test_plain_form_is_rejected(small_graph):
    c = _open(small_graph, 2000)
    try:
        sp = _sgns(c, ffi.DG_SGNS_PLAIN_FP32, 1)
        c.sgns_init(sp)
        c.walk_reset()
        with pytest.raises(distger.DgError) as e:
            c.pipeline_step(c.default_walk_params(seed=7), 1, 0, 2000, sp)
        assert e.value.code == -8
    finally:
        c.close()


def test_zero_epochs_walk_only(small_graph):
    """iter 0: the step walks the round and trains nothing."""
    c = _open(small_graph, 2000)
    try:
        sp = _sgns(c, ffi.DG_SGNS_BATCHED_BF16, 0, iter=0)
        c.sgns_init(sp)
        c.walk_reset()
        before = c.sgns_tables(128)
        rs, ss = c.pipeline_step(c.default_walk_params(seed=7), 1, 0, 2000, sp)
        after = c.sgns_tables(128)
        assert rs.walkers == 2000 and rs.tokens > 2000
        assert ss.words == 0 and ss.kept == 0 and ss.launches == 0
        assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    finally:
        c.close()
