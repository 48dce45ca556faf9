"""CPU checks of the SGNS half of the oracle (oracle/sgns_oracle.c).

Each set-up function is checked against an independent statement of what it
computes (sorting, big-integer arithmetic, proportions, the closed form of a
formula); the training restatements against hand-derived SGD steps on tiny
sentences and, for the batched forms' first-order evaluation, against the
sequential order it approximates.
"""

import math

import numpy as np

import dgsgns_oracle


def _trainer(counts, dim, **kw):
    return dgsgns_oracle.Trainer(np.asarray(counts, np.uint64), dim, **kw)


def test_vocab_ranked_by_count_then_vertex_id():
    rng = np.random.default_rng(1)
    counts = rng.integers(1, 6, 300)  # many ties
    tr = _trainer(counts, 8, table_size=1000)
    order = np.lexsort((np.arange(300), -counts))  # count descending, then id
    assert np.array_equal(tr.vertex_of_rank, order)
    assert tr.train_words == counts.sum()
    assert np.array_equal(tr.rank_of_vertex[order], np.arange(300))


def test_vocab_min_count_keeps_the_first_word():
    counts = np.array([1, 9, 3, 2, 9], np.uint64)
    tr = _trainer(counts, 8, min_count=3, table_size=100)
    assert list(tr.vertex_of_rank) == [1, 4, 2]
    assert tr.train_words == 21
    tr = _trainer(np.array([1, 1, 1], np.uint64), 8, min_count=5, table_size=100)
    assert tr.vocab == 1  # word2vec keeps rank 0 whatever its count


def test_init_rows_follow_the_lcg():
    """syn0[a*d+b] from r <- r * 25214903917 + 11 (mod 2^64), r0 = 1, as
    ((r & 0xFFFF) / 65536 - 0.5) / d; computed with Python integers."""
    dim = 16
    tr = _trainer(np.arange(10, 0, -1), dim, table_size=100)
    r = 1
    exp = []
    for _ in range(tr.vocab * dim):
        r = (r * 25214903917 + 11) % (1 << 64)
        exp.append(np.float32(((r & 0xFFFF) / 65536.0 - 0.5) / dim))
    assert np.array_equal(tr.syn0.ravel(), np.array(exp, np.float32))
    assert np.all(np.abs(tr.syn0) <= 0.5 / dim)
    trf = _trainer(np.arange(10, 0, -1), dim, table_size=100, float_init=True)
    assert np.abs(trf.syn0 - tr.syn0).max() <= 1e-9


def test_unigram_table_proportions():
    counts = np.array([400, 300, 100, 100, 50, 20, 10, 5, 2, 1], np.uint64)
    ts = 1_000_000
    tr = _trainer(counts, 8, table_size=ts)
    p = counts.astype(np.float64) ** 0.75
    p /= p.sum()
    got = np.bincount(tr.table, minlength=len(counts)) / ts
    assert np.all(np.abs(got - p) <= 2.0 / ts)
    assert np.all(np.diff(tr.table) >= 0) and tr.table[0] == 0


def test_sigmoid_table_closed_form():
    tr = _trainer(np.ones(4), 8, table_size=100)
    i = np.arange(1000)
    x = ((i / np.float32(1000)).astype(np.float32) * 2 - 1) * 6
    e = np.exp(x.astype(np.float64)).astype(np.float32)
    exp = e / (e + np.float32(1))
    assert np.abs(tr.expt[:1000] - exp).max() <= 1e-7
    assert tr.expt[1000] == tr.expt[999]
    assert abs(tr.expt[500] - 0.5) < 1e-6


def test_keep_limits_against_the_keep_rule():
    """A token of count c is kept iff (sqrt(c/(s T)) + 1) s T / c >= u / 65536
    for its 16-bit draw u (word2vec.cu:1015-1020); the oracle stores the
    largest kept u."""
    counts = np.array([100000, 50000, 5000, 500, 50, 5], np.uint64)
    tr = _trainer(counts, 8, table_size=1000, sample=1e-3)
    st = np.float32(1e-3) * np.float32(counts.sum())
    for w, c in enumerate(counts):
        ran = (np.sqrt(np.float32(c) / st) + np.float32(1)) * st / np.float32(c)
        u = np.arange(65536)
        kept = (u / np.float32(65536)) <= ran
        lim = int(tr.keep[w])
        assert np.array_equal(kept, u <= lim), w


def _sig(f):
    return 1.0 / (1.0 + math.exp(-f))


def test_shared_order_one_pair_by_hand():
    """Sentence [a, b], window 1: position 0 has context b, position 1 context
    a; each position one shared negative.  The textbook sequential SGD steps,
    written out by hand in float64."""
    rng = np.random.default_rng(3)
    tr = _trainer(np.arange(30, 0, -1), 8, table_size=5000)
    x0 = rng.normal(0, 0.2, (tr.vocab, 8)).astype(np.float32)
    y0 = rng.normal(0, 0.2, (tr.vocab, 8)).astype(np.float32)
    a, b = 4, 11
    sen = np.array([a, b], np.int32)
    off = np.array([0, 2], np.uint64)
    alpha = 0.05
    X = x0.astype(np.float64)
    Y = y0.astype(np.float64)
    for q, (center, ctx) in enumerate(((a, b), (b, a))):
        bq, negs = tr.position_draws(9, 0, q, 0, 1, 1)
        assert bq == 0  # window 1: b = draw % 1
        x = X[ctx].copy()
        neu1e = np.zeros(8)
        for row, lab in ((center, 1.0), (int(negs[0]), 0.0)):
            if lab == 0.0 and row == center:
                continue
            g = (lab - _sig(x @ Y[row])) * alpha
            neu1e += g * Y[row]
            Y[row] += g * x
        X[ctx] += neu1e
    tr.syn0, tr.syn1 = x0.copy(), y0.copy()
    tr.shared(sen, off, 1, 1, alpha, True, 9, 0)
    assert np.abs(tr.syn0 - X).max() < 1e-6 and np.abs(tr.syn1 - Y).max() < 1e-6


def test_first_order_is_exact_with_one_context_per_position():
    """Window 1 on two-token sentences and one negative: no pair of a position
    interacts with another, so the batched evaluation equals the sequential
    order (with more negatives, two equal ones of a position would)."""
    rng = np.random.default_rng(4)
    tr = _trainer(np.arange(60, 0, -1), 16, table_size=5000)
    x0 = rng.normal(0, 0.2, (tr.vocab, 16)).astype(np.float32)
    y0 = rng.normal(0, 0.2, (tr.vocab, 16)).astype(np.float32)
    toks = rng.integers(0, tr.vocab, 80).astype(np.int32)
    off = np.arange(0, 81, 2).astype(np.uint64)
    tr.syn0, tr.syn1 = x0.copy(), y0.copy()
    tr.shared(toks, off, 1, 1, 0.05, True, 2, 0)
    seq = (tr.syn0.copy(), tr.syn1.copy())
    tr.syn0, tr.syn1 = x0.copy(), y0.copy()
    tr.shared_first_order(toks, off, 1, 1, 0.05, True, 2, 0)
    assert np.abs(tr.syn0 - seq[0]).max() < 1e-6 and np.abs(tr.syn1 - seq[1]).max() < 1e-6


def _bounce_corpus(rng, vocab, n, length=22):
    sens = []
    for _ in range(n):
        hub = int(rng.integers(0, vocab))
        sens.append([hub if i % 2 == 0 else int(rng.integers(0, vocab)) for i in range(length)])
    sen = np.array(sum(sens, []), np.int32)
    off = np.cumsum([0] + [len(x) for x in sens]).astype(np.uint64)
    return sen, off


def test_first_order_converges_to_the_sequential_order():
    """Walk-like sentences (a vertex every other position, so contexts repeat
    inside a window): the batched evaluation drops only second-order terms, so
    its difference from the sequential order, relative to the update, shrinks
    faster than alpha.  Dropping the H X or M Y term (or a Gram term) leaves a
    first-order difference that does not."""
    rng = np.random.default_rng(0)
    tr = _trainer(rng.integers(2, 50, 200), 32, table_size=10000)
    sen, off = _bounce_corpus(rng, tr.vocab, 40)
    x0 = rng.normal(0, 0.01, (tr.vocab, 32)).astype(np.float32)
    y0 = rng.normal(0, 0.01, (tr.vocab, 32)).astype(np.float32)
    rel = []
    for alpha in (0.05, 0.0125):
        tr.syn0, tr.syn1 = x0.copy(), y0.copy()
        tr.shared(sen, off, 5, 5, alpha, True, 3, 0)
        s0, s1 = tr.syn0.copy(), tr.syn1.copy()
        tr.syn0, tr.syn1 = x0.copy(), y0.copy()
        tr.shared_first_order(sen, off, 5, 5, alpha, True, 3, 0)
        d0 = np.abs(tr.syn0 - s0).max() / np.abs(s0 - x0).max()
        d1 = np.abs(tr.syn1 - s1).max() / np.abs(s1 - y0).max()
        rel.append(max(d0, d1))
    assert rel[0] < 3e-2 and rel[1] < 1e-2
    assert rel[1] < rel[0] / 2.5


def test_shared_draw_stream_layout():
    """Position draws: word 0 of the (sentence, position) stream gives the
    shrink, words 1..K the negatives, independent of K."""
    tr = _trainer(np.arange(100, 0, -1), 8, table_size=10000)
    b3, n3 = tr.position_draws(7, 5, 4, 0, 10, 3)
    b5, n5 = tr.position_draws(7, 5, 4, 0, 10, 5)
    assert b3 == b5 and np.array_equal(n3, n5[:3])
    draws = [tr.position_draws(7, 5, q, 0, 10, 1)[0] for q in range(200)]
    assert set(draws) == set(range(10))  # every shrink value occurs
    other = [tr.position_draws(7, 5, q, 1, 10, 1)[0] for q in range(200)]
    assert draws != other  # the epoch is part of the counter
