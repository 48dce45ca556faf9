"""Checks of the DSGL restatement (oracle/dsgl_oracle.c, dsgl_worker.inc) of
the reference's CPU trainer (include/dsgl.hpp:783-1427).

The oracle is pinned three ways:
* an independent statement of one compute thread, written here from the
  reference's text (its variable names: sentence_length, neg_buffer,
  outputs.indices / meta, corrM), run on tiny corpora that reach every branch
  (ranges that start mid-walk and run out with sentences in hand, sentences
  of one token, negatives that repeat or equal the target, target2 equal to
  a negative, the 10000-word rate updates), against the FP64 oracle;
* closed forms: from zero output rows with the exact sigmoid, the first
  step's output updates are +-alpha/2 sums of the context rows with the
  reference's labels;
* the FP32 threaded form against the FP64 one.
A change to any labelled rule of the oracle fails a case here.
"""

import math

import numpy as np
import pytest

import dgoracle
import dgsgns_oracle

MAX_SENT = 1000


def philox_word(key, worker, c2, tag, j):
    o = dgoracle.philox((j >> 2, worker, c2, tag), (key & 0xFFFFFFFF, key >> 32))
    return int(o[j & 3])


TAG_SAMPLE = 0x53554253
TAG_WIN = 0x44535749
TAG_NEG = 0x4453474E


def epoch_key(seed, epoch):
    return seed ^ (epoch << 32)


class Ref:
    """One DSGL compute thread (dsgl.hpp:783-1427) with the documented
    deviations: Philox draws, write-backs that add live - snapshot, the
    rate from the worker's own count times the number of workers."""

    def __init__(self, tr, tokens, offsets, window, negative, alpha0, workers, seed, exact):
        self.tr, self.tok, self.off = tr, [int(x) for x in tokens], [int(x) for x in offsets]
        self.w, self.K, self.alpha0, self.T, self.seed, self.exact = window, negative, alpha0, workers, seed, exact
        # local_corpus: every walk followed by -1
        self.corpus = []
        self.tidx = []
        for w in range(len(self.off) - 1):
            for i in range(self.off[w], self.off[w + 1]):
                self.corpus.append(self.tok[i])
                self.tidx.append(i)
            self.corpus.append(-1)
            self.tidx.append(-1)

    def sig(self, f):
        if self.exact:
            return 1.0 / (1.0 + math.exp(-f))
        if f > 6:
            return 1.0
        if f < -6:
            return 0.0
        return float(self.tr.expt[int((f + 6) * 83)])

    def run_worker(self, id_, syn0, syn1, epoch, actual, alpha):
        tr, K, W = self.tr, self.K, self.w
        key = epoch_key(self.seed, epoch)
        words_num = len(self.corpus)
        thread_words_num = words_num // self.T
        start_words = thread_words_num * id_
        if id_ == self.T - 1:
            thread_words_num = words_num - (self.T - 1) * thread_words_num
        end = start_words + thread_words_num
        cur = start_words
        word_count = last_word_count = 0
        sen = [[], []]
        pos = [0, 0]
        buf = [{}, {}]      # word -> [live, snapshot]
        neg_buffer = {}     # word -> [live, snapshot]
        neg_words = []
        change_neg, neg_count, train_sen_num, step, batch = True, 0, 0, 0, 0
        total = np.float32(1 * tr.train_words + 1)

        def read():
            nonlocal cur, word_count
            s = []
            while True:
                origin_id = self.corpus[cur]
                ti = self.tidx[cur]
                word_count += 1
                cur += 1
                if origin_id == -1 or cur >= end:
                    break
                w = int(tr.rank_of_vertex[origin_id])
                draw = int(dgoracle.philox((0, ti & 0xFFFFFFFF, ti >> 32, TAG_SAMPLE),
                                           (key & 0xFFFFFFFF, key >> 32))[0])
                if (draw & 0xFFFF) > int(tr.dsgl_keep[w]):
                    continue
                s.append(w)
                if len(s) >= MAX_SENT:
                    break
            return s

        while True:
            if word_count - last_word_count > 10000:
                actual += word_count - last_word_count
                last_word_count = word_count
                a = np.float32(self.alpha0) * (np.float32(1) - np.float32(actual * self.T) / total)
                alpha = max(a, np.float32(self.alpha0) * np.float32(0.0001))
            for k in (0, 1):
                while len(sen[k]) == 0:
                    if cur >= end:
                        break
                    s = read()
                    if len(s) <= 1:
                        continue
                    sen[k], pos[k] = s, 0
                    buf[k] = {}
                    for x in s:
                        if x not in buf[k]:
                            buf[k][x] = [syn0[x].copy(), syn0[x].copy()]
            if cur >= end:
                actual += word_count - last_word_count
                return actual, alpha
            if change_neg:
                neg_batch_num = max(len(sen[0]), len(sen[1]))
                neg_words = []
                neg_buffer = {}
                for i in range(neg_batch_num * K):
                    x = philox_word(key, id_, batch, TAG_NEG, i)
                    smp = int(tr.dsgl_table[x % len(tr.dsgl_table)])
                    if smp == 0:
                        smp = x % (tr.vocab - 1) + 1
                    neg_words.append(smp)
                    if smp not in neg_buffer:
                        neg_buffer[smp] = [syn1[smp].copy(), syn1[smp].copy()]
                batch += 1
                change_neg = False
            target, target2 = sen[0][pos[0]], sen[1][pos[1]]
            b = philox_word(key, id_, step, TAG_WIN, 0) % W
            step += 1
            inputs = []  # (sentence, word)
            first_input_num = 0
            for k in (0, 1):
                for i in range(b, 2 * W + 1 - b):
                    if i == W:
                        continue
                    c = pos[k] - W + i
                    if c < 0:
                        continue
                    if c >= len(sen[k]):
                        break
                    inputs.append((k, sen[k][c]))
                if k == 0:
                    first_input_num = len(inputs)
            neg_count += 1
            st = neg_count % neg_batch_num
            indices, meta = [target], [1]
            for k in range(K):
                smp = neg_words[st * K + k]
                if smp in indices:
                    meta[indices.index(smp)] += 1
                else:
                    indices.append(smp)
                    meta.append(1)
            meta[0] = 1
            indices.append(target2)
            meta.append(1)
            O = len(indices)
            outputM = [syn1[target].copy()] + [neg_buffer[x][0].copy() for x in indices[1:-1]] + \
                      [syn1[target2].copy()]
            inputM = [buf[k][x][0].copy() for k, x in inputs]
            corrM = np.zeros((O, len(inputs)))
            for i in range(O):
                for j in range(len(inputs)):
                    f = float(np.dot(outputM[i], inputM[j]))
                    label = (i == 0) if j < first_input_num else (i == O - 1)
                    corrM[i, j] = (label - self.sig(f)) * float(alpha) * meta[i]
            outputMd = [sum(corrM[i, j] * inputM[j] for j in range(len(inputs))) for i in range(O)]
            dX = [sum(corrM[i, j] * outputM[i] for i in range(O)) for j in range(len(inputs))]
            for j, (k, x) in enumerate(inputs):
                buf[k][x][0] += dX[j]
            syn1[target] += outputMd[0]
            syn1[target2] += outputMd[O - 1]
            for i in range(1, O - 1):
                neg_buffer[indices[i]][0] += outputMd[i]
            for k in (0, 1):
                pos[k] += 1
                if pos[k] >= len(sen[k]):
                    for x, (live, snap) in buf[k].items():
                        syn0[x] += live - snap
                    sen[k] = []
                    train_sen_num += 1
            if train_sen_num >= 2:
                for x, (live, snap) in neg_buffer.items():
                    syn1[x] += live - snap
                change_neg = True
                train_sen_num = 0

    def train(self, syn0, syn1, epoch=0):
        s0 = syn0.astype(np.float64)
        s1 = syn1.astype(np.float64)
        for t in range(self.T):
            self.run_worker(t, s0, s1, epoch, 0, np.float32(self.alpha0))
        return s0, s1


def _corpus(lens, V, seed):
    rng = np.random.default_rng(seed)
    toks = []
    for L in lens:
        toks.extend(rng.integers(0, V, L).tolist())
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    return np.array(toks, np.uint32), off


def _trainer(V, dim, sample, table_size=97, seed=3, counts=None):
    rng = np.random.default_rng(seed)
    cn = counts if counts is not None else rng.integers(1, 60, V)
    tr = dgsgns_oracle.Trainer(np.asarray(cn, np.uint64), dim, table_size=table_size, sample=sample,
                               float_init=True)
    tr.dsgl_setup(sample=sample, table_size=table_size)
    # rows big enough that the table sigmoid is exercised away from 1/2
    tr.syn0 = (rng.standard_normal((tr.vocab, dim)) * 0.4).astype(np.float32)
    tr.syn1 = (rng.standard_normal((tr.vocab, dim)) * 0.4).astype(np.float32)
    return tr


@pytest.mark.parametrize("lens,window,K,workers,sample,exact,seed", [
    ([5, 7, 1, 6, 4, 9, 3, 8], 3, 2, 1, 0.0, True, 1),      # one worker, single-token sentence
    ([5, 7, 2, 6, 4, 9, 3, 8, 5, 5], 2, 3, 3, 0.0, False, 2),  # ranges start mid-walk, table sigmoid
    ([12, 3, 11, 4, 10, 6, 2, 9], 4, 5, 2, 0.02, True, 3),    # subsampling on
    ([6] * 30, 2, 4, 4, 0.0, True, 4),                         # many negative batches
])
def test_dsgl_matches_independent_statement(lens, window, K, workers, sample, exact, seed):
    V, dim = 9, 8   # a small vocabulary: repeated negatives, negatives equal to targets
    tok, off = _corpus(lens, V, seed)
    tr = _trainer(V, dim, sample, seed=seed)
    r0, r1 = Ref(tr, tok, off, window, K, 0.05, workers, 11, exact).train(tr.syn0, tr.syn1)
    s0, s1 = tr.syn0.copy(), tr.syn1.copy()
    tr.dsgl(tok, off, window, K, 0.05, workers, 11, exact=exact)
    assert not np.array_equal(tr.syn1, s1)
    np.testing.assert_allclose(tr.syn0, r0, rtol=0, atol=1e-6)
    np.testing.assert_allclose(tr.syn1, r1, rtol=0, atol=1e-6)


def test_dsgl_rate_follows_counted_words():
    """Walks long enough that a worker counts past 10000 words: the rate steps
    down as alpha0 (1 - T * counted / (train_words + 1)) from then on."""
    V, dim = 40, 8
    lens = [25] * 900
    tok, off = _corpus(lens, V, 7)
    tr = _trainer(V, dim, 0.0, seed=7, counts=np.full(V, 600))
    r0, r1 = Ref(tr, tok, off, 2, 1, 0.05, 2, 5, True).train(tr.syn0, tr.syn1)
    tr.dsgl(tok, off, 2, 1, 0.05, 2, 5, exact=True)
    np.testing.assert_allclose(tr.syn0, r0, rtol=0, atol=1e-6)
    np.testing.assert_allclose(tr.syn1, r1, rtol=0, atol=1e-6)


def test_dsgl_first_step_closed_form():
    """Zero output rows, exact sigmoid, no negatives: the first step moves
    the target row by +alpha/2 times the sum of the first sentence's context
    rows and -alpha/2 times the second's (target is a negative of the second
    window), and target2 the other way round; the input rows do not move
    (their gradient is a sum of zero output rows).  Three walks of two
    tokens: the first two are trained for two steps, reading the third ends
    the range.  Step two sees only the first step's output moves."""
    V, dim = 6, 8
    tok = np.array([0, 1, 2, 3, 4, 5], np.uint32)
    off = np.array([0, 2, 4, 6], np.uint64)
    tr = _trainer(V, dim, 0.0, counts=np.array([9, 8, 7, 6, 5, 4]))
    tr.syn1[:] = 0
    x = {int(v): tr.syn0[tr.rank_of_vertex[v]].astype(np.float64) for v in range(V)}
    a = np.float32(0.05)
    tr.dsgl(tok, off, 1, 0, float(a), 1, 3, exact=True)
    r = lambda v: int(tr.rank_of_vertex[v])
    # step 1: targets 0 (contexts {1}) and 2 (contexts {3})
    d0 = 0.5 * a * x[1] - 0.5 * a * x[3]
    d2 = -0.5 * a * x[1] + 0.5 * a * x[3]
    # step 2: targets 1 (contexts {0}) and 3 (contexts {2}) from zero rows again
    d1 = 0.5 * a * x[0] - 0.5 * a * x[2]
    d3 = -0.5 * a * x[0] + 0.5 * a * x[2]
    np.testing.assert_allclose(tr.syn1[r(0)], d0, atol=1e-7)
    np.testing.assert_allclose(tr.syn1[r(2)], d2, atol=1e-7)
    np.testing.assert_allclose(tr.syn1[r(1)], d1, atol=1e-7)
    np.testing.assert_allclose(tr.syn1[r(3)], d3, atol=1e-7)
    assert not tr.syn1[r(4)].any() and not tr.syn1[r(5)].any()
    # the input rows' gradients are sums of the output rows each step reads:
    # rows 0 and 2 at step 1, rows 1 and 3 at step 2, all still zero then
    for v in range(V):
        np.testing.assert_array_equal(tr.syn0[r(v)], x[v].astype(np.float32))


def test_dsgl_f32_threads_match_f64():
    V, dim = 30, 16
    tok, off = _corpus([10] * 400, V, 8)
    tr = _trainer(V, dim, 0.001, table_size=1000, seed=8)
    init0, init1 = tr.syn0.copy(), tr.syn1.copy()
    tr.dsgl(tok, off, 3, 4, 0.025, 1, 9, exact=False)
    d64 = tr.syn0 - init0, tr.syn1 - init1
    tr.syn0, tr.syn1 = init0.copy(), init1.copy()
    tr.dsgl_f32(tok, off, 3, 4, 0.025, 1, 9, threads=1, exact=False)
    e0 = np.abs((tr.syn0 - init0) - d64[0]).max() / np.abs(d64[0]).max()
    e1 = np.abs((tr.syn1 - init1) - d64[1]).max() / np.abs(d64[1]).max()
    assert e0 < 1e-3 and e1 < 1e-3
    # several threads race on shared rows: same magnitude of training
    tr.syn0, tr.syn1 = init0.copy(), init1.copy()
    tr.dsgl_f32(tok, off, 3, 4, 0.025, 8, 9, threads=4, exact=False)
    assert np.isfinite(tr.syn0).all() and np.isfinite(tr.syn1).all()
    n = np.linalg.norm(tr.syn1 - init1) / np.linalg.norm(d64[1])
    assert 0.5 < n < 2.0


def test_dsgl_tables():
    """DSGL's unigram table follows pow(cn, 0.75) with a float cursor, and its
    keep limit is floor((sqrt(r) + r) * 65536) with r = s T / cn in float."""
    counts = np.array([400, 300, 100, 100, 50, 20, 10, 5, 2, 1], np.uint64)
    tr = dgsgns_oracle.Trainer(counts, 8, table_size=1_000_000)
    tr.dsgl_setup(sample=1e-2, table_size=1_000_000)
    p = counts.astype(np.float64) ** 0.75
    p /= p.sum()
    frac = np.bincount(tr.dsgl_table, minlength=10) / 1e6
    assert np.abs(frac - p).max() < 2e-4
    for w in range(10):
        r = np.float32(np.float32(1e-2) * np.float32(counts.sum())) / np.float32(counts[w])
        ran = np.float32(np.sqrt(r)) + r
        lim = ran * np.float32(65536)
        exp = 65536 if lim >= 65536 else int(lim)
        assert tr.dsgl_keep[w] == exp
