"""CPU checks of the library's host utilities: AUC, the link-prediction split
and its files (src/tools/gtrain.cpp:226-386), the text -> binary edge
converter (src/tools/gconverter.cpp:45-76) and the learning-rate schedules
(src/lr_scheduler.hpp).  No GPU is touched."""

import math
import struct

import numpy as np
import pytest

import distger
from distger import ffi

F = np.float32


# ---- AUC ---------------------------------------------------------------------

def mann_whitney(pos, neg):
    """Brute force over all pairs: 1 for pos > neg, 1/2 for a tie."""
    s = 0.0
    for p in pos:
        s += float(np.sum(neg < p)) + 0.5 * float(np.sum(neg == p))
    return s / (len(pos) * len(neg))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_auc_matches_brute_force_with_ties(seed):
    rng = np.random.default_rng(seed)
    pos = rng.integers(0, 8, 300).astype(np.float32)  # few distinct values: many ties
    neg = rng.integers(-2, 6, 250).astype(np.float32)
    assert abs(distger.auc(pos, neg) - mann_whitney(pos, neg)) < 1e-12


def test_auc_matches_sklearn():
    from sklearn.metrics import roc_auc_score
    rng = np.random.default_rng(7)
    pos = (rng.standard_normal(2000) + 0.7).astype(np.float32)
    neg = rng.standard_normal(1500).astype(np.float32)
    pos[:100] = neg[:100]  # exact ties across the classes
    y = np.concatenate([np.ones(len(pos)), np.zeros(len(neg))])
    ref = roc_auc_score(y, np.concatenate([pos, neg]).astype(np.float64))
    assert abs(distger.auc(pos, neg) - ref) < 1e-12


def test_auc_special_cases():
    assert distger.auc([1, 2, 3], [0, -1]) == 1.0
    assert distger.auc([0, -1], [1, 2, 3]) == 0.0
    assert distger.auc([5, 5], [5, 5, 5]) == 0.5


# ---- link-prediction split ---------------------------------------------------

def _graph(v, e, seed):
    return distger.gen_chung_lu(v, e, 2.5, 2, seed)


def test_lp_split_invariants():
    V = 3000
    edges = _graph(V, 12000, 4)
    deg0 = np.bincount(np.concatenate([edges["src"], edges["dst"]]), minlength=V)
    train, pos, neg = distger.lp_split(V, edges, 0.1, 9)
    # gtrain.cpp:283: npos = sample * edges.size(), truncated; as many non-edges
    assert len(pos) == int(0.1 * len(edges)) and len(neg) == len(pos)
    # train + positives = the input edges (a permutation of them)
    key = lambda e: np.sort(e["src"].astype(np.uint64) << 32 | e["dst"].astype(np.uint64))
    assert np.array_equal(key(np.concatenate([train, pos])), key(edges))
    # gtrain.cpp:297-300: a positive is removed only while both endpoints keep
    # degree > 1, so no vertex of the input loses its last edge
    degt = np.bincount(np.concatenate([train["src"], train["dst"]]), minlength=V)
    assert np.all(degt[deg0 > 0] >= 1)
    # non-edges: not edges of the full graph in either direction, distinct as
    # unordered pairs, no self-pairs
    full = set(zip(edges["src"].tolist(), edges["dst"].tolist()))
    full |= {(b, a) for a, b in full}
    pairs = [(int(a), int(b)) for a, b in zip(neg["src"], neg["dst"])]
    assert all(a != b for a, b in pairs)
    assert not any(p in full for p in pairs)
    assert len({(min(a, b), max(a, b)) for a, b in pairs}) == len(pairs)
    # the same seed gives the same split
    t2, p2, n2 = distger.lp_split(V, edges, 0.1, 9)
    assert np.array_equal(t2, train) and np.array_equal(p2, pos) and np.array_equal(n2, neg)


def test_lp_split_degree_guard_on_a_star():
    """A star: every edge has a leaf of degree 1, so no edge may be held out and
    the split reports that it cannot draw its positives."""
    e = np.zeros(20, ffi.EDGE_DTYPE)
    e["src"] = 0
    e["dst"] = np.arange(1, 21)
    with pytest.raises(distger.DgError):
        distger.lp_split(21, e, 0.1, 1)


def test_lp_files_bytes(tmp_path):
    V = 500
    edges = _graph(V, 2000, 5)
    edges["data"] = np.arange(len(edges)) % 7
    train, pos, neg = distger.lp_split(V, edges, 0.2, 3)
    paths = [str(tmp_path / n) for n in ("train.txt", "test_edges.txt", "test_labels.txt")]
    distger.lp_write_files(*paths, train, pos, neg)
    exp_train = "".join(f"{a} {b} {int(d)}\n" for a, b, d in zip(train["src"], train["dst"], train["data"]))
    exp_edges = "".join(f"{a} {b}\n" for a, b in zip(pos["src"], pos["dst"]))
    exp_edges += "".join(f"{a} {b}\n" for a, b in zip(neg["src"], neg["dst"]))
    exp_labels = "1\n" * len(pos) + "0\n" * len(neg)
    for p, exp in zip(paths, (exp_train, exp_edges, exp_labels)):
        assert open(p).read() == exp


# ---- text -> binary converter ------------------------------------------------

def test_convert_weighted_and_unweighted(tmp_path):
    rng = np.random.default_rng(3)
    n = 1000
    src = rng.integers(0, 1 << 31, n, dtype=np.uint64).astype(np.uint32)
    dst = rng.integers(0, 1 << 31, n, dtype=np.uint64).astype(np.uint32)
    w = rng.random(n).astype(np.float32)
    txt = tmp_path / "w.txt"
    txt.write_text("".join(f"{a} {b} {x!r}\n" for a, b, x in zip(src, dst, w.tolist())))
    out = tmp_path / "w.bin"
    assert distger.convert_text_edges(str(txt), str(out), True) == n
    exp = b"".join(struct.pack("<IIf", int(a), int(b), float(x)) for a, b, x in zip(src, dst, w))
    assert out.read_bytes() == exp
    e = distger.read_edges_bin(str(out))
    assert np.array_equal(e["src"], src) and np.array_equal(e["dst"], dst) and np.array_equal(e["data"], w)
    # unweighted: 8-byte records; reading stops at the first line that does
    # not scan (here a comment)
    txt2 = tmp_path / "u.txt"
    txt2.write_text("".join(f"{a}\t{b}\n" for a, b in zip(src[:10], dst[:10])) + "# end\n1 2\n")
    out2 = tmp_path / "u.bin"
    assert distger.convert_text_edges(str(txt2), str(out2), False) == 10
    assert out2.read_bytes() == b"".join(struct.pack("<II", int(a), int(b)) for a, b in zip(src[:10], dst[:10]))


def test_convert_missing_input_is_an_error(tmp_path):
    with pytest.raises(distger.DgError):
        distger.convert_text_edges(str(tmp_path / "none.txt"), str(tmp_path / "o.bin"), True)


# ---- learning-rate schedules -------------------------------------------------

def test_lr_fixed_and_exponential():
    """ExponentialDecayLR(0.025, 0.1), the GPU trainer's schedule
    (word2vec.cu:1279): initial * exp(-decay * step) in float."""
    for step in range(40):
        # float argument -decay * step; expf is within half an ulp and a bit of
        # the exact value, so compare with the exactly rounded exp to 1 ulp
        arg = F(-F(0.1)) * F(step)
        exp = F(0.025) * F(math.exp(float(arg)))
        got = F(distger.lr_value(ffi.DG_LR_EXPONENTIAL, 0.025, 0.1, 0, step))
        assert abs(int(got.view(np.int32)) - int(exp.view(np.int32))) <= 1
        assert distger.lr_value(ffi.DG_LR_FIXED, 0.025, 0, 0, step) == F(0.025)
    v = [distger.lr_value(ffi.DG_LR_EXPONENTIAL, 0.025, 0.1, 0, k) for k in range(28)]
    assert all(a > b for a, b in zip(v, v[1:]))
    assert abs(v[10] - 0.025 * math.exp(-1.0)) < 1e-8


def test_lr_cosine_and_step():
    for step in range(25):
        cosine = F(math.cos(math.pi * step / 10))
        exp = F(float(F(0.001)) + 0.5 * (float(F(0.05)) - float(F(0.001))) * float(F(1) + cosine))
        assert distger.lr_value(ffi.DG_LR_COSINE, 0.05, 0.001, 10, step) == exp
        exp = F(float(F(0.05)) * float(F(0.5)) ** (step // 3))
        assert distger.lr_value(ffi.DG_LR_STEP, 0.05, 0.5, 3, step) == exp
    assert distger.lr_value(ffi.DG_LR_COSINE, 0.05, 0.001, 10, 0) == F(0.05)
    assert distger.lr_value(ffi.DG_LR_COSINE, 0.05, 0.001, 10, 10) == F(0.001)
    assert math.isnan(distger.lr_value(ffi.DG_LR_STEP, 0.05, 0.5, 0, 1))
    assert math.isnan(distger.lr_value(9, 0.05, 0, 0, 1))
