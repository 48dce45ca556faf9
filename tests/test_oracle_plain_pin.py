"""Independent derivation of the plain-form oracle (dgo_sg_plain_batch,
dgo_plain_batches), the restatement of skip_gram_kernel
(src/word2vec.cu:305-399) and of its host batching (:999-1041).

Every quirk is written out here from the reference's text, with Python
integers for the two LCGs and numpy float32 scalars for each rounding step:

* the window LCG multiplier 2514903917 (word2vec.cu:321, one digit short of
  the negative-sampling LCG's 25214903917, :363), seeded with the block index
  (the sentence's index in its batch);
* b = next_random % window, the shrunk window a in [b, 2w+1-b), a != w;
* a negative: next_random * 25214903917 + 11, table[(next_random >> 16) %
  table_size], an entry 0 replaced by next_random % (vocab - 1) + 1, a
  negative equal to the centre word skipped (:363-367);
* the sigmoid by the 1000-entry table at index (int)((f + 6) * 83): the
  integer division 1000 / 6 / 2 = 83 (:384-389), saturating outside [-6, 6];
* the dot product in the block-reduction order: per lane l the products at
  dims l + 64 j folded by halving (f[x] += f[x + FSIZE], :373-376), then the
  64 partials folded at distances 32, 16, ..., 1 (:377-381);
* the updates: neu1e += g y and y += g x per output (:391-395), x += neu1e
  after the context's outputs (:396-397), each a separately rounded float op;
* batches close at 15000 sentences or once they hold >= (15000 - 1) * 20
  tokens (:1002-1040).

A one-line change to any of these in the oracle makes a case here fail.
"""

import numpy as np

import dgsgns_oracle

F = np.float32
M64 = (1 << 64) - 1


def exp_table():
    """word2vec.cu:1569-1577 (float argument, exp in double, float ratio)."""
    t = np.zeros(1001, np.float32)
    for i in range(1000):
        arg = (F(i) / F(1000) * F(2) - F(1)) * F(6)
        e = F(np.exp(np.float64(arg)))
        t[i] = e / (e + F(1))
    t[1000] = t[999]
    return t


def tree_dot(x, y):
    dim = len(x)
    per = dim // 64
    s = []
    for lane in range(64):
        p = [F(x[lane + 64 * j]) * F(y[lane + 64 * j]) for j in range(per)]
        n = per
        while n > 1:
            h = n // 2
            for j in range(h):
                p[j] = F(p[j] + p[j + h])
            n = h
        s.append(p[0])
    off = 32
    while off >= 1:
        for lane in range(off):
            s[lane] = F(s[lane] + s[lane + off])
        off //= 2
    return s[0]


def grad(f, label, alpha, expt):
    if f > F(6):
        return F(label - 1) * F(alpha)
    if f < F(-6):
        return F(label) * F(alpha)
    idx = int((f + F(6)) * F(83))
    return (F(label) - expt[idx]) * F(alpha)


def plain_batch(sen, off, window, negative, table, vocab, alpha, expt, syn0, syn1):
    """Hand statement of one batch of skip_gram_kernel, sentences in order."""
    dim = syn0.shape[1]
    for s in range(len(off) - 1):
        nr = s
        for pos in range(int(off[s]), int(off[s + 1])):
            word = int(sen[pos])
            nr = (nr * 2514903917 + 11) & M64
            b = nr % window
            for a in range(b, 2 * window + 1 - b):
                if a == window:
                    continue
                c = pos - window + a
                if c < off[s] or c >= off[s + 1]:
                    continue
                x = syn0[int(sen[c])]
                neu1e = np.zeros(dim, np.float32)
                for d in range(negative + 1):
                    if d == 0:
                        target, label = word, 1
                    else:
                        nr = (nr * 25214903917 + 11) & M64
                        target = int(table[(nr >> 16) % len(table)])
                        if target == 0:
                            target = nr % (vocab - 1) + 1
                        if target == word:
                            continue
                        label = 0
                    y = syn1[target]
                    g = grad(tree_dot(x, y), label, alpha, expt)
                    for i in range(dim):
                        neu1e[i] = F(neu1e[i] + F(g * y[i]))
                        y[i] = F(y[i] + F(g * x[i]))
                for i in range(dim):
                    x[i] = F(x[i] + neu1e[i])


def _setup(vocab, dim, seed):
    rng = np.random.default_rng(seed)
    syn0 = (rng.standard_normal((vocab, dim)) * 0.5).astype(np.float32)
    syn1 = (rng.standard_normal((vocab, dim)) * 0.5).astype(np.float32)
    return syn0, syn1


def _oracle_batch(sen, off, window, negative, table, vocab, alpha, expt, syn0, syn1):
    L = dgsgns_oracle.lib()
    vp = dgsgns_oracle._vp
    sen = np.ascontiguousarray(sen, np.int32)
    off = np.ascontiguousarray(off, np.uint64)
    table = np.ascontiguousarray(table, np.int32)
    L.dgo_sg_plain_batch(vp(sen), vp(off), len(off) - 1, window, negative, vp(table), len(table),
                         vocab, alpha, vp(expt), syn0.shape[1], vp(syn0), vp(syn1))


def test_exp_table_matches_the_oracle():
    o = np.zeros(1001, np.float32)
    dgsgns_oracle.lib().dgo_exp_table(dgsgns_oracle._vp(o))
    assert np.array_equal(o, exp_table())


def test_plain_batch_by_hand():
    """Sentences of lengths 1..7 (one of a single token, repeated words, a
    window wider than the sentence), a table whose zero entries force the
    remap and whose draws hit the centre word, rows large enough that the
    dot products reach both saturated ends of the sigmoid."""
    vocab, dim, window, negative, alpha = 9, 64, 3, 4, F(0.05)
    table = np.array([0, 3, 3, 5, 0, 7, 1, 2, 2, 8, 0, 4, 6, 6, 1, 0, 5], np.int32)
    sen = np.array([4, 2, 6, 2, 8, 1, 3, 5, 5, 0, 7, 1, 2, 3, 4, 8, 8, 6, 2, 1, 3, 7, 0, 4,
                    5, 6, 2, 2, 1], np.int32)
    lens = [7, 1, 3, 6, 2, 5, 4, 1]
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    assert off[-1] == len(sen)
    expt = exp_table()
    s0, s1 = _setup(vocab, dim, 3)
    h0, h1 = s0.copy(), s1.copy()
    _oracle_batch(sen, off, window, negative, table, vocab, alpha, expt, s0, s1)
    plain_batch(sen, off, window, negative, table, vocab, alpha, expt, h0, h1)
    assert np.array_equal(s0, h0) and np.array_equal(s1, h1)
    assert not np.array_equal(s0, _setup(vocab, dim, 3)[0])


def test_plain_batch_by_hand_dim_128_small_rows():
    """Two dims per lane (the halving fold) and rows in the table's interior."""
    vocab, dim, window, negative, alpha = 7, 128, 2, 2, F(0.025)
    table = np.array([1, 0, 2, 3, 3, 4, 5, 6, 0, 2], np.int32)
    sen = np.array([1, 2, 3, 1, 4, 5, 6, 0, 2, 2, 3], np.int32)
    off = np.array([0, 5, 11], np.uint64)
    rng = np.random.default_rng(5)
    s0 = ((rng.random((vocab, dim)) - 0.5) / dim).astype(np.float32)
    s1 = (rng.standard_normal((vocab, dim)) * 0.05).astype(np.float32)
    h0, h1 = s0.copy(), s1.copy()
    expt = exp_table()
    _oracle_batch(sen, off, window, negative, table, vocab, F(alpha), expt, s0, s1)
    plain_batch(sen, off, window, negative, table, vocab, alpha, expt, h0, h1)
    assert np.array_equal(s0, h0) and np.array_equal(s1, h1)


def test_window_lcg_uses_the_short_multiplier():
    """The shrink b of position p of sentence s is the p-th value of
    r <- r * 2514903917 + 11 from r = s (mod window); with no negatives and a
    single context pair per position, the contexts trained reveal b: a wrong
    multiplier changes which pairs are trained."""
    vocab, dim, window = 40, 64, 5
    sen = np.arange(1, 31, dtype=np.int32) % vocab
    off = np.array([0, 30], np.uint64)
    expt = exp_table()
    s0, s1 = _setup(vocab, dim, 9)
    h0, h1 = s0.copy(), s1.copy()
    table = np.array([1], np.int32)
    _oracle_batch(sen, off, window, 0, table, vocab, F(0.1), expt, s0, s1)
    plain_batch(sen, off, window, 0, table, vocab, F(0.1), expt, h0, h1)
    assert np.array_equal(s0, h0) and np.array_equal(s1, h1)


def test_batch_boundaries():
    """15000 sentences per batch at most; a batch also closes once it holds
    >= 299,980 tokens (the sentence that crosses the mark is its last)."""
    L = dgsgns_oracle.lib()
    vp = dgsgns_oracle._vp

    def batches(lens):
        off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
        firsts = np.zeros(len(lens) + 2, np.uint64)
        nb = L.dgo_plain_batches(vp(off), len(lens), vp(firsts), len(lens) + 1)
        return [int(x) for x in firsts[: nb + 1]]

    assert batches([1] * 20000) == [0, 15000, 20000]
    assert batches([1] * 15000) == [0, 15000]
    assert batches([20] * 30000) == [0, 14999, 29998, 30000]       # 14999 * 20 = 299,980
    assert batches([30] * 20000) == [0, 10000, 20000]              # 9999 * 30 < 299,980 <= 10000 * 30
    assert batches([19] * 16000) == [0, 15000, 16000]              # 15000 * 19 < 299,980
    lens = [5] * 100 + [299_500] + [1] * 10
    assert batches(lens) == [0, 101, 111]                          # the long sentence closes the batch
    assert batches([]) == [0]


def test_sigmoid_boundaries_exactly_at_six():
    """f = 6 and f = -6 exactly read the table (entries 996 and 0); only
    |f| > 6 saturates (word2vec.cu:384-386)."""
    vocab, dim = 4, 64
    expt = exp_table()
    for sign in (1.0, -1.0):
        sen = np.array([1, 2], np.int32)
        off = np.array([0, 2], np.uint64)
        s0 = np.zeros((vocab, dim), np.float32)
        s1 = np.zeros((vocab, dim), np.float32)
        s0[2, 0] = 6.0 * sign  # context of position 0
        s1[1, 0] = 1.0         # its target
        h0, h1 = s0.copy(), s1.copy()
        _oracle_batch(sen, off, 1, 0, np.array([1], np.int32), vocab, F(0.5), expt, s0, s1)
        plain_batch(sen, off, 1, 0, np.array([1], np.int32), vocab, F(0.5), expt, h0, h1)
        assert np.array_equal(s0, h0) and np.array_equal(s1, h1)
        g = (F(1) - expt[996 if sign > 0 else 0]) * F(0.5)
        assert s1[1, 0] != F(1.0) or g == 0
        assert s1[1, 0] == F(F(1.0) + F(g * F(6.0 * sign)))


def reuse_batch(sen, off, window, negs, K, alpha, expt, syn0, syn1):
    """Hand statement of __sgNegReuse (word2vec.cu:231-303): the full window,
    the sentence's K shared negatives, a negative equal to the centre skipped;
    the outputs of a pair one after another from the same context row."""
    dim = syn0.shape[1]
    for s in range(len(off) - 1):
        for pos in range(int(off[s]), int(off[s + 1])):
            word = int(sen[pos])
            for a in range(2 * window + 1):
                if a == window:
                    continue
                c = pos - window + a
                if c < off[s] or c >= off[s + 1]:
                    continue
                x = syn0[int(sen[c])]
                neu1e = np.zeros(dim, np.float32)
                for d in range(K + 1):
                    target, label = (word, 1) if d == 0 else (int(negs[s * K + d - 1]), 0)
                    if d > 0 and target == word:
                        continue
                    y = syn1[target]
                    g = grad(tree_dot(x, y), label, alpha, expt)
                    for i in range(dim):
                        neu1e[i] = F(neu1e[i] + F(g * y[i]))
                        y[i] = F(y[i] + F(g * x[i]))
                for i in range(dim):
                    x[i] = F(x[i] + neu1e[i])


def test_reuse_batch_by_hand():
    vocab, dim, window, K, alpha = 9, 64, 2, 3, F(0.05)
    sen = np.array([4, 2, 6, 2, 8, 1, 3, 5, 5, 0, 7, 1, 2, 3], np.int32)
    off = np.array([0, 6, 7, 14], np.uint64)
    negs = np.array([2, 7, 7, 3, 3, 3, 1, 5, 2], np.int32)  # repeats; a centre word among them
    expt = exp_table()
    s0, s1 = _setup(vocab, dim, 4)
    h0, h1 = s0.copy(), s1.copy()
    L = dgsgns_oracle.lib()
    vp = dgsgns_oracle._vp
    L.dgo_sg_reuse_batch(vp(sen), vp(off), 3, window, K, vp(negs), alpha, vp(expt), dim, vp(s0), vp(s1))
    reuse_batch(sen, off, window, negs, K, alpha, expt, h0, h1)
    assert np.array_equal(s0, h0) and np.array_equal(s1, h1)


def test_reuse_negatives_by_hand():
    """x = word k of Philox(counter (k / 4, s, 0, REUSE), key seed ^ epoch << 32);
    table[x % table_size], an entry 0 replaced by x % (vocab - 1) + 1."""
    import dgoracle
    table = np.array([0, 3, 0, 5, 1, 2, 0], np.int32)
    vocab, K, seed, epoch = 8, 6, 77, 3
    out = np.zeros(4 * K, np.int32)
    dgsgns_oracle.lib().dgo_reuse_negs(4, K, seed, epoch, dgsgns_oracle._vp(table), len(table), vocab,
                                       dgsgns_oracle._vp(out))
    key = seed ^ (epoch << 32)
    for s in range(4):
        for k in range(K):
            x = int(dgoracle.philox((k >> 2, s, 0, 0x52555345), (key & 0xFFFFFFFF, key >> 32))[k & 3])
            v = int(table[x % len(table)])
            assert out[s * K + k] == (v if v else x % (vocab - 1) + 1)
