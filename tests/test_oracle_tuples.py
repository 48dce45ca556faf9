"""CPU pins of the tuple-interface oracle (oracle/sgns_oracle.c, tuple
section) and of the synthetic-input generators it is fed with.

The oracle restates the loss whose gradient skip_gram_kernel follows
(src/word2vec.cu:359-396, the negative skip at :366).  It is pinned by what
the mathematics fixes: FP64 central differences of the loss, the value
log 2 per term at zero rows, the reference's in-place order derived by hand
on one tuple (including a repeated negative, which must see its own update),
first-order agreement of that order with -lr * grad, and linearity in the
upstream gradient.
"""

import numpy as np
import pytest
import torch

import distger
import dgsgns_oracle as so


def tables(rows, dim, seed, sigma=0.1):
    rng = np.random.default_rng(seed)
    s0 = (rng.standard_normal((rows, dim)) * sigma).astype(np.float32)
    s1 = (rng.standard_normal((rows, dim)) * sigma).astype(np.float32)
    return s0, s1


def random_tuples(rows, n, K, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, rows, size=(n, K + 2)).astype(np.uint32)


def np_loss(t, K, s0, s1):
    """FP64 loss from numpy dots (live-term mask as in src/word2vec.cu:366)."""
    x = s0[t[:, 1]].astype(np.float64)
    tgt = np.concatenate([t[:, :1], t[:, 2:]], axis=1)
    y = s1[tgt].astype(np.float64)
    z = np.einsum("nd,njd->nj", x, y)
    live = np.ones_like(z, bool)
    live[:, 1:] = t[:, 2:] != t[:, :1]
    a = z.copy()
    a[:, 0] = -a[:, 0]
    return np.sum(np.logaddexp(0.0, a) * live), z, live


def test_forward_logits_and_loss_match_numpy():
    s0, s1 = tables(50, 64, 1)
    t = random_tuples(50, 300, 5, 2)
    t[::7, 3] = t[::7, 0]  # some negatives equal to the centre
    o = so.Tuples(t, 5, s0, s1)
    logits, loss = o.forward()
    ref_loss, z, live = np_loss(t, 5, s0, s1)
    assert np.allclose(logits, z, rtol=1e-12, atol=1e-14)
    assert abs(loss - ref_loss) < 1e-9 * abs(ref_loss)
    assert not live.all()


def test_bf16_operands_are_round_to_nearest_even():
    s0, s1 = tables(40, 128, 3)
    t = random_tuples(40, 100, 3, 4)
    b0 = torch.from_numpy(s0).to(torch.bfloat16).to(torch.float32).numpy()
    b1 = torch.from_numpy(s1).to(torch.bfloat16).to(torch.float32).numpy()
    lb, _ = so.Tuples(t, 3, s0, s1).forward(bf16=True)
    lr_, _ = so.Tuples(t, 3, b0, b1).forward(bf16=False)
    assert np.array_equal(lb, lr_)


def test_zero_rows_give_log2_per_live_term():
    s0 = np.zeros((10, 64), np.float32)
    s1 = np.zeros((10, 64), np.float32)
    t = np.array([[1, 2, 3, 4], [5, 6, 5, 7], [8, 9, 8, 8]], np.uint32)  # K = 2
    _, loss = so.Tuples(t, 2, s0, s1).forward()
    live_terms = 3 + 2 + 1
    assert abs(loss - live_terms * np.log(2.0)) < 1e-12


def test_fd_helper_equals_full_loss_difference():
    """dgo_tuple_fd evaluates only the tuples that read the row; the full
    numpy loss difference must agree."""
    s0, s1 = tables(20, 64, 5)
    t = random_tuples(20, 80, 4, 6)
    o = so.Tuples(t, 4, s0, s1)
    h = 1e-3
    for table, row, col in [(0, int(t[0, 1]), 3), (1, int(t[1, 0]), 17), (1, int(t[2, 3]), 63)]:
        a = [s0.astype(np.float64), s1.astype(np.float64)]
        plus = [a[0].copy(), a[1].copy()]
        minus = [a[0].copy(), a[1].copy()]
        plus[table][row, col] += h
        minus[table][row, col] -= h
        full = (np_loss(t, 4, *plus)[0] - np_loss(t, 4, *minus)[0]) / (2 * h)
        assert abs(o.fd(table, row, col, h) - full) < 1e-7 * max(1.0, abs(full))


@pytest.mark.parametrize("bf16", [False])
def test_gradient_matches_central_differences(bf16):
    s0, s1 = tables(30, 64, 7)
    t = random_tuples(30, 200, 5, 8)
    t[::5, 4] = t[::5, 0]
    o = so.Tuples(t, 5, s0, s1)
    rng = np.random.default_rng(9)
    for table in (0, 1):
        rows = np.unique(t[:, 1] if table == 0 else t[:, [0, 2, 3, 4, 5, 6]].ravel())
        g = o.grad_rows(table, rows, bf16=bf16)
        for _ in range(12):
            i = int(rng.integers(len(rows)))
            col = int(rng.integers(64))
            fd = o.fd(table, int(rows[i]), col, 1e-3)
            assert abs(g[i, col] - fd) < 1e-6 * max(1.0, abs(fd)), (table, rows[i], col)


def test_untouched_rows_have_zero_gradient():
    s0, s1 = tables(30, 64, 10)
    t = random_tuples(15, 50, 2, 11)  # rows 15..29 never used
    g = so.Tuples(t, 2, s0, s1).grad_rows(1, np.arange(15, 30))
    assert not g.any()


def test_skipped_negative_equals_tuple_without_it():
    s0, s1 = tables(12, 64, 12)
    with_skip = np.array([[3, 5, 3], [4, 6, 4]], np.uint32)  # K = 1, negative == centre
    without = np.array([[3, 5], [4, 6]], np.uint32)          # K = 0
    a = so.Tuples(with_skip, 1, s0, s1)
    b = so.Tuples(without, 0, s0, s1)
    assert abs(a.forward()[1] - b.forward()[1]) < 1e-15
    for table in (0, 1):
        rows = np.arange(12)
        assert np.array_equal(a.grad_rows(table, rows), b.grad_rows(table, rows))


def test_inplace_order_by_hand():
    """One tuple (c, x, n, n), K = 2: the repeated negative reads the row
    its first occurrence updated; the context row takes sum g_j y_j with
    each y_j as it was read (src/word2vec.cu:384-396)."""
    s0, s1 = tables(8, 64, 13)
    t = np.array([[2, 5, 6, 6]], np.uint32)
    lr = 0.3
    x = s0[5].astype(np.float64)
    yc = s1[2].astype(np.float64)
    yn = s1[6].astype(np.float64)
    sig = lambda v: 1.0 / (1.0 + np.exp(-v))
    g0 = (1.0 - sig(x @ yc)) * lr
    neu = g0 * yc
    yc2 = yc + g0 * x
    g1 = (0.0 - sig(x @ yn)) * lr
    neu = neu + g1 * yn
    yn2 = yn + g1 * x
    g2 = (0.0 - sig(x @ yn2)) * lr
    neu = neu + g2 * yn2
    yn3 = yn2 + g2 * x
    x2 = x + neu
    a, b = so.Tuples(t, 2, s0, s1).inplace(lr)
    assert np.allclose(a[5], x2, rtol=0, atol=1e-14)
    assert np.allclose(b[2], yc2, rtol=0, atol=1e-14)
    assert np.allclose(b[6], yn3, rtol=0, atol=1e-14)


def test_inplace_is_minus_lr_grad_to_first_order():
    s0, s1 = tables(200, 64, 14)
    t = random_tuples(200, 400, 5, 15)
    o = so.Tuples(t, 5, s0, s1)
    rows = np.arange(200)
    g0 = o.grad_rows(0, rows)
    g1 = o.grad_rows(1, rows)
    errs = []
    for lr in (1e-2, 1e-3):
        a, b = o.inplace(lr)
        d = np.concatenate([(a - s0).ravel(), (b - s1).ravel()])
        ref = -lr * np.concatenate([g0.ravel(), g1.ravel()])
        errs.append(np.abs(d - ref).max() / np.abs(ref).max())
    assert errs[0] < 0.05 and errs[1] < errs[0] / 5


def test_upstream_gradient_identity_and_linearity():
    s0, s1 = tables(25, 64, 16)
    t = random_tuples(25, 120, 3, 17)
    t[::4, 2] = t[::4, 0]
    o = so.Tuples(t, 3, s0, s1)
    z, _ = o.forward()
    live = np.ones_like(z, bool)
    live[:, 1:] = t[:, 2:] != t[:, :1]
    label = np.zeros_like(z)
    label[:, 0] = 1.0
    own = np.where(live, 1.0 / (1.0 + np.exp(-z)) - label, 0.0)
    rows = np.arange(25)
    for table in (0, 1):
        assert np.allclose(o.grad_rows(table, rows, dlogits=own), o.grad_rows(table, rows),
                           rtol=1e-12, atol=1e-15)
        rng = np.random.default_rng(18 + table)
        d1 = rng.standard_normal(z.shape)
        d2 = rng.standard_normal(z.shape)
        lhs = o.grad_rows(table, rows, dlogits=2.0 * d1 - 3.0 * d2)
        rhs = 2.0 * o.grad_rows(table, rows, dlogits=d1) - 3.0 * o.grad_rows(table, rows, dlogits=d2)
        assert np.allclose(lhs, rhs, rtol=1e-10, atol=1e-12)


def test_zipf_generator_frequencies():
    ids = distger.gen_zipf(100, 1.1, 400_000, 3)
    assert ids.max() < 100 and ids.dtype == np.uint32
    p = (np.arange(1, 101, dtype=np.float64)) ** -1.1
    p /= p.sum()
    f = np.bincount(ids, minlength=100) / len(ids)
    # binomial standard error per bin, 6 sigma
    assert np.all(np.abs(f - p) < 6 * np.sqrt(p * (1 - p) / len(ids)))
    assert np.array_equal(ids, distger.gen_zipf(100, 1.1, 400_000, 3))
    assert not np.array_equal(ids, distger.gen_zipf(100, 1.1, 400_000, 4))


def test_normal_generator_moments():
    x = distger.gen_normal(1_000_001, 0.1, 5).astype(np.float64)
    assert len(x) == 1_000_001
    assert abs(x.mean()) < 6 * 0.1 / 1000
    assert abs(x.std() - 0.1) < 6 * 0.1 / np.sqrt(2 * 1e6)
    # fraction inside one sigma: erf(1/sqrt 2) = 0.682689...
    assert abs(np.mean(np.abs(x) < 0.1) - 0.6826894921) < 6 * np.sqrt(0.22 / 1e6)
