"""Every sentence of a launch is trained once the batched kernel's wavefronts
take their last sentences from a counter (Hogwild mode), not only those of
the stride.

The graph is 131,072 disjoint edges (i, i + 131072); one round walks from
i = 0 .. 131071 only, so walk i holds exactly the two vertices of edge i and
no other walk holds them.  Without subsampling, a walk's sentence is the only
one that moves the input rows of its two vertices, so an input row left at
its initial value means a sentence was never trained.  Two launches of
65,536 sentences each leave, at 256 CUs, 4,096 sentences per launch to the
counter.
"""

import numpy as np
import pytest

import distger
from distger import ffi

pytestmark = pytest.mark.gpu

M = 131072


def _pairs():
    e = np.zeros(M, dtype=ffi.EDGE_DTYPE)
    e["src"] = np.arange(M, dtype=np.uint32)
    e["dst"] = np.arange(M, dtype=np.uint32) + M
    e["data"] = 1.0
    return e


@pytest.mark.parametrize("form", [ffi.DG_SGNS_BATCHED_BF16, ffi.DG_SGNS_BATCHED_FP32])
def test_every_sentence_trained(form):
    print(f"[balance-gpu] {M} two-vertex walks, form {form}", flush=True)
    with distger.Context(0) as c:
        c.graph_from_edges(2 * M, _pairs(), undirected=True, cn_mode=ffi.DG_CN_COMPUTE)
        wp = c.default_walk_params(seed=3, max_length=24)
        c.walk_reset()
        try:
            c.walk_round(wp, 1, 0, M)
        except distger.DgError as e:
            assert e.code == -7
        tok, off = c.corpus()
        assert len(off) == M + 1
        lens = np.diff(off.astype(np.int64))
        assert lens.min() >= 2
        # walk i holds i and i + M only
        walk_of = np.repeat(np.arange(M), lens)
        assert np.all((tok == walk_of) | (tok == walk_of + M))
        p = c.default_sgns_params(form=form, dim=128, window=5, negative=5, sample=0.0,
                                  min_count=0, alpha=0.025, table_size=1_000_000, seed=7,
                                  deterministic=0, iter=1, minibatch=65536, block_waves=4, pool_split=16)
        c.sgns_init(p)
        vocab = c.sgns_vocab()
        assert len(vocab) == 2 * M
        init0, _ = c.sgns_tables(128)
        st = c.sgns_train(p, corpus=ffi.DG_CORPUS_ROUND)
        assert st.kept == len(tok) and st.launches == 2
        s0, _ = c.sgns_tables(128)
        assert np.isfinite(s0).all()
        moved = np.any(s0 != init0, axis=1)
        print(f"[balance-gpu] rows moved {int(moved.sum())} of {len(moved)}", flush=True)
        assert moved.all(), f"{int((~moved).sum())} input rows never trained"
