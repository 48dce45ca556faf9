"""GPU parity of the differentiable tuple interface (dg_sgns_tuples_forward /
dg_sgns_tuples_backward / dg_sgns_grad_rows) against the FP64 oracle.

Tuples are (center, context, neg_1..neg_K) with ids from Zipf(1.1), so the
hot rows receive thousands of gradient entries and their segmented sums
cross many 64-entry blocks.  Tolerances (max abs error over the compared
values divided by their max abs value):
  FP32 form vs the FP64 oracle on the same FP32 rows      1e-5
  BF16 form vs the FP64 oracle on the same BF16-rounded rows 1e-5
  BF16 form vs the FP64 oracle on the FP32 rows (total)   2e-2
The reference's in-place update (src/word2vec.cu:359-396) agrees with the
device's W -= lr dL/dW to first order in lr (the gap shrinks in proportion
to lr).
"""

import numpy as np
import pytest

import distger
import dgsgns_oracle as so

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL_BF16_TOTAL = 2e-2


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def zipf_tuples(rows, n, K, seed):
    return distger.gen_zipf(rows, 1.1, n * (K + 2), seed).reshape(n, K + 2)


def normal_tables(rows, dim, seed):
    s0 = distger.gen_normal(rows * dim, 0.1, seed).reshape(rows, dim)
    s1 = distger.gen_normal(rows * dim, 0.1, seed + 1).reshape(rows, dim)
    return s0, s1


def touched(t):
    return np.unique(t[:, 1]), np.unique(np.concatenate([t[:, :1], t[:, 2:]], axis=1))


@pytest.fixture
def tctx():
    c = distger.Context(0)
    yield c
    c.close()


def check_case(c, t, K, s0, s1, bf16, label):
    print(f"[tuples-gpu] {label}: n={t.shape[0]} K={K} rows={s0.shape[0]} dim={s0.shape[1]} bf16={bf16}",
          flush=True)
    c.sgns_tables_create(s0, s1)
    o = so.Tuples(t, K, s0, s1)
    logits, loss = c.tuples_forward(t, K, bf16=bf16)
    ol, oloss = o.forward(bf16=bf16)
    assert rel(logits, ol) < TOL32, f"logits {rel(logits, ol):.2e}"
    assert abs(loss - oloss) < TOL32 * abs(oloss)
    c.tuples_backward(t, K, bf16=bf16)
    r0, r1 = touched(t)
    g0 = c.grad_rows(0, r0, s0.shape[1])
    g1 = c.grad_rows(1, r1, s0.shape[1])
    o0 = o.grad_rows(0, r0, bf16=bf16)
    o1 = o.grad_rows(1, r1, bf16=bf16)
    e0, e1 = rel(g0, o0), rel(g1, o1)
    assert e0 < TOL32 and e1 < TOL32, f"grads {e0:.2e} {e1:.2e}"
    if bf16:
        t0, t1 = rel(g0, o.grad_rows(0, r0)), rel(g1, o.grad_rows(1, r1))
        tl = rel(logits, o.forward()[0])
        print(f"[tuples-gpu]   BF16 vs FP64 on FP32 rows: logits {tl:.2e} grads {t0:.2e} {t1:.2e}",
              flush=True)
        assert tl < TOL_BF16_TOTAL and t0 < TOL_BF16_TOTAL and t1 < TOL_BF16_TOTAL
    return g0, g1


@pytest.mark.parametrize("dim,K,bf16", [(128, 5, False), (128, 5, True), (64, 0, False),
                                        (256, 15, False), (256, 7, True), (64, 10, True)])
def test_forward_backward_vs_oracle(tctx, dim, K, bf16):
    rows, n = 5000, 4096
    t = zipf_tuples(rows, n, K, 11 + K)
    if K:
        t[::9, 2] = t[::9, 0]  # negatives equal to the centre are skipped
    s0, s1 = normal_tables(rows, dim, 21)
    check_case(tctx, t, K, s0, s1, bf16, "zipf")


def test_one_hot_row_across_many_blocks(tctx):
    """Every tuple shares its context and centre: two rows with 6000
    entries each, summed over ~94 blocks by the carry path."""
    rows, n, K = 300, 6000, 5
    rng = np.random.default_rng(3)
    t = rng.integers(0, rows, size=(n, K + 2)).astype(np.uint32)
    t[:, 0] = 3
    t[:, 1] = 7
    s0, s1 = normal_tables(rows, 128, 5)
    check_case(tctx, t, K, s0, s1, False, "hot rows")


def test_backward_is_deterministic_and_untouched_rows_are_zero(tctx):
    rows, n, K = 20000, 8192, 5
    t = zipf_tuples(rows // 2, n, K, 31)  # rows >= 10000 untouched
    s0, s1 = normal_tables(rows, 128, 41)
    tctx.sgns_tables_create(s0, s1)
    probe = np.concatenate([np.arange(0, 200), np.arange(10000, 10050)]).astype(np.uint32)
    out = []
    for _ in range(2):
        tctx.tuples_backward(t, K)
        out.append((tctx.grad_rows(0, probe, 128), tctx.grad_rows(1, probe, 128)))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert not out[0][0][200:].any() and not out[0][1][200:].any()


def test_upstream_gradient(tctx):
    rows, n, K = 3000, 2048, 5
    t = zipf_tuples(rows, n, K, 51)
    s0, s1 = normal_tables(rows, 128, 61)
    tctx.sgns_tables_create(s0, s1)
    dl = np.random.default_rng(7).standard_normal((n, K + 1)).astype(np.float32)
    tctx.tuples_backward(t, K, dlogits=dl)
    o = so.Tuples(t, K, s0, s1)
    live = np.ones((n, K + 1), bool)
    live[:, 1:] = t[:, 2:] != t[:, :1]
    r0, r1 = touched(t)
    for table, r in ((0, r0), (1, r1)):
        g = tctx.grad_rows(table, r, 128)
        og = o.grad_rows(table, r, dlogits=np.where(live, dl, 0.0))
        assert rel(g, og) < TOL32


@pytest.mark.parametrize("bf16", [False, True])
def test_apply_matches_reference_inplace_to_first_order(tctx, bf16):
    """apply: W -= lr dL/dW (device) against the oracle gradient step, and
    against the reference's sequential in-place update, which differs from
    the gradient step by O(lr^2): on Zipf ids the hot rows take thousands of
    in-place steps, so the relative gap is ~16 lr here and falls tenfold per
    tenfold smaller lr."""
    rows, n, K = 4000, 4096, 5
    t = zipf_tuples(rows, n, K, 71)
    s0, s1 = normal_tables(rows, 128, 81)
    o = so.Tuples(t, K, s0, s1)
    r = np.arange(rows)
    og = np.concatenate([o.grad_rows(0, r, bf16=bf16).ravel(), o.grad_rows(1, r, bf16=bf16).ravel()])
    gaps = []
    for lr in (1e-4, 1e-5):
        tctx.sgns_tables_create(s0, s1)
        tctx.tuples_backward(t, K, bf16=bf16, lr=lr, apply=True)
        a, b = tctx.sgns_tables(128)
        d_dev = np.concatenate([(a - s0).ravel(), (b - s1).ravel()]).astype(np.float64)
        # the step itself: FP32 rounding of W + step is ~1e-8 absolute
        step_err = np.abs(d_dev + lr * og).max() / np.abs(lr * og).max()
        ia, ib = o.inplace(lr)
        gap = rel(d_dev, np.concatenate([(ia - s0).ravel(), (ib - s1).ravel()]))
        print(f"[tuples-gpu] lr {lr}: step vs oracle {step_err:.2e}, vs reference in-place {gap:.2e}",
              flush=True)
        assert step_err < 2e-3
        gaps.append(gap)
    assert gaps[0] < 3e-2
    if not bf16:  # the BF16 gradient's own rounding (~2e-3) does not shrink with lr
        assert gaps[1] < gaps[0] / 5
    # the BF16 copies follow the updated rows: a BF16 forward now equals the
    # oracle on the new FP32 rows rounded to BF16
    logits, _ = tctx.tuples_forward(t[:512], K, bf16=True)
    ol, _ = so.Tuples(t[:512], K, a, b).forward(bf16=True)
    assert rel(logits, ol) < TOL32


def test_errors(tctx):
    s0, s1 = normal_tables(100, 64, 1)
    tctx.sgns_tables_create(s0, s1)
    with pytest.raises(distger.DgError) as e:
        tctx.grad_rows(0, np.arange(3), 64)
    assert e.value.code == -5
    bad = np.array([[1, 2, 100]], np.uint32)
    with pytest.raises(distger.DgError) as e:
        tctx.tuples_forward(bad, 1)
    assert e.value.code == -1
    with pytest.raises(distger.DgError) as e:
        tctx.tuples_backward(np.zeros((4, 18), np.uint32), 16)
    assert e.value.code == -8


@pytest.mark.parametrize("bf16", [False, True])
def test_differentiation_workload_fd(tctx, bf16):
    """65,536 tuples, ids Zipf(1.1) over 2,238,731 rows, dim 128, 5
    negatives, rows N(0, 0.1^2): device gradients against FP64 central
    differences (step 1e-3) on 256 random coordinates, 128 in each table."""
    rows, n, K, dim = 2238731, 65536, 5, 128
    print(f"[tuples-gpu] differentiation workload, bf16={bf16}: generating", flush=True)
    t = zipf_tuples(rows, n, K, 2024)
    s0, s1 = normal_tables(rows, dim, 99)
    tctx.sgns_tables_create(s0, s1)
    logits, loss = tctx.tuples_forward(t, K, bf16=bf16)
    o = so.Tuples(t, K, s0, s1)
    ol, oloss = o.forward()
    print(f"[tuples-gpu]   logits err {rel(logits, ol):.2e}, loss {loss:.6f} vs {oloss:.6f}", flush=True)
    assert rel(logits, ol) < (TOL_BF16_TOTAL if bf16 else TOL32)
    tctx.tuples_backward(t, K, bf16=bf16)
    rng = np.random.default_rng(5 + bf16)
    r0, r1 = touched(t)
    errs, scale = [], []
    for table, pool in ((0, r0), (1, r1)):
        rows_s = pool[rng.integers(0, len(pool), 128)]
        cols = rng.integers(0, dim, 128)
        g = tctx.grad_rows(table, rows_s, dim)[np.arange(128), cols]
        fd = np.array([o.fd(table, int(r), int(col), 1e-3) for r, col in zip(rows_s, cols)])
        errs.append(np.abs(g - fd))
        scale.append(np.abs(fd))
    err = np.concatenate(errs).max() / np.concatenate(scale).max()
    print(f"[tuples-gpu]   gradient vs FP64 central differences on 256 coordinates: {err:.2e}",
          flush=True)
    assert err < (TOL_BF16_TOTAL if bf16 else 1e-4)


@pytest.mark.parametrize("rows,dim,n,label", [
    (3_000_000, 64, 200_000, "1.4M entries: more sort tiles than CUs, tiles ordered by ticket"),
    (8_500_000, 64, 40_000, "rows beyond 2^23: 25-bit keys, four radix passes"),
])
def test_sort_paths_at_scale(tctx, rows, dim, n, label):
    """The backward's sort takes tiles in ticket order once there are more
    tiles (4,096 entries each) than CUs, and a fourth 8-bit pass for keys
    wider than 24 bits: gradient rows against the oracle, and bit-equal over
    two backward calls.  A third of the tuples get uniform negatives, so the
    high digits vary."""
    K = 5
    print(f"[tuples-gpu] {label}: rows={rows} dim={dim} n={n}", flush=True)
    t = zipf_tuples(rows, n, K, 91)
    rng = np.random.default_rng(5)
    t[::3, 2:] = rng.integers(0, rows, size=t[::3, 2:].shape).astype(np.uint32)
    s0, s1 = normal_tables(rows, dim, 93)
    tctx.sgns_tables_create(s0, s1)
    o = so.Tuples(t, K, s0, s1)
    r0, r1 = touched(t)
    # the hottest rows and a sample of the rest
    probe0 = np.unique(np.concatenate([r0[:64], rng.choice(r0, min(len(r0), 1000), replace=False)]))
    probe1 = np.unique(np.concatenate([r1[:64], r1[-64:], rng.choice(r1, min(len(r1), 1000), replace=False)]))
    got = []
    for _ in range(2):
        tctx.tuples_backward(t, K)
        got.append((tctx.grad_rows(0, probe0, dim), tctx.grad_rows(1, probe1, dim)))
    assert np.array_equal(got[0][0], got[1][0]) and np.array_equal(got[0][1], got[1][1])
    e0 = rel(got[0][0], o.grad_rows(0, probe0))
    e1 = rel(got[0][1], o.grad_rows(1, probe1))
    print(f"[tuples-gpu]   gradient rows vs oracle: {e0:.2e} {e1:.2e}", flush=True)
    assert e0 < TOL32 and e1 < TOL32


@pytest.mark.parametrize("dim,K,bf16", [(128, 5, False), (128, 5, True), (256, 15, False), (64, 10, True)])
def test_fused_step_equals_forward_and_backward(tctx, dim, K, bf16):
    """dg_sgns_tuples_step: the logits and loss of the forward and the
    gradients of the backward (the loss's own gradient), bit for bit, and
    the same update as the backward's apply."""
    rows, n = 5000, 4096
    print(f"[tuples-gpu] fused step: dim={dim} K={K} bf16={bf16}", flush=True)
    t = zipf_tuples(rows, n, K, 61 + K)
    s0, s1 = normal_tables(rows, dim, 71)
    tctx.sgns_tables_create(s0, s1)
    r0, r1 = touched(t)
    logits, loss = tctx.tuples_forward(t, K, bf16=bf16)
    tctx.tuples_backward(t, K, bf16=bf16)
    g0, g1 = tctx.grad_rows(0, r0, dim), tctx.grad_rows(1, r1, dim)
    fl, floss = tctx.tuples_step(t, K, bf16=bf16)
    assert np.array_equal(fl, logits) and floss == loss
    assert np.array_equal(tctx.grad_rows(0, r0, dim), g0)
    assert np.array_equal(tctx.grad_rows(1, r1, dim), g1)
    tctx.tuples_backward(t, K, bf16=bf16, lr=0.01, apply=True)
    a0, a1 = tctx.sgns_tables(dim)
    tctx.sgns_tables_create(s0, s1)
    tctx.tuples_step(t, K, bf16=bf16, lr=0.01, apply=True)
    b0, b1 = tctx.sgns_tables(dim)
    assert np.array_equal(a0, b0) and np.array_equal(a1, b1)
    assert not np.array_equal(a1, s1)


@pytest.mark.parametrize("n,K", [(1, 0), (1, 5), (2, 1), (3, 5), (17, 2), (64, 0), (700, 15)])
def test_tiny_batches(tctx, n, K):
    """Batches far below one sort tile (4,096 entries) and one 64-entry block:
    gradient rows, logits and loss against the oracle, and the fused step
    against the separate calls."""
    rows, dim = 50, 64
    print(f"[tuples-gpu] tiny batch: n={n} K={K}", flush=True)
    rng = np.random.default_rng(n * 31 + K)
    t = rng.integers(0, rows, size=(n, K + 2)).astype(np.uint32)
    s0, s1 = normal_tables(rows, dim, 3)
    tctx.sgns_tables_create(s0, s1)
    o = so.Tuples(t, K, s0, s1)
    logits, loss = tctx.tuples_forward(t, K)
    ol, oloss = o.forward()
    assert rel(logits, ol) < TOL32 and abs(loss - oloss) < TOL32 * abs(oloss)
    tctx.tuples_backward(t, K)
    r = np.arange(rows, dtype=np.uint32)
    g0, g1 = tctx.grad_rows(0, r, dim), tctx.grad_rows(1, r, dim)
    assert rel(g0, o.grad_rows(0, r)) < TOL32 and rel(g1, o.grad_rows(1, r)) < TOL32
    fl, floss = tctx.tuples_step(t, K)
    assert np.array_equal(fl, logits) and floss == loss
    assert np.array_equal(tctx.grad_rows(0, r, dim), g0) and np.array_equal(tctx.grad_rows(1, r, dim), g1)
