"""ctypes bindings for the SGNS half of the CPU oracle (oracle/sgns_oracle.c).

Test infrastructure: the reference results for the device trainer, never a product path.
"""

import ctypes

import numpy as np

import dgoracle


def _vp(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


_L = None


def lib():
    global _L
    if _L is not None:
        return _L
    L = dgoracle.load("dg")
    c = ctypes
    vp = c.c_void_p
    L.dgo_vocab.argtypes = [c.c_uint32, vp, c.c_uint32, vp, vp, vp, c.POINTER(c.c_uint32),
                            c.POINTER(c.c_uint64)]
    L.dgo_init_syn0.argtypes = [c.c_uint32, c.c_uint32, c.c_int, vp]
    L.dgo_unigram_table.argtypes = [vp, c.c_uint32, c.c_uint32, vp]
    L.dgo_exp_table.argtypes = [vp]
    L.dgo_keep_limits.argtypes = [vp, c.c_uint32, c.c_float, c.c_uint64, vp]
    L.dgo_subsample.restype = c.c_uint64
    L.dgo_subsample.argtypes = [vp, vp, c.c_uint64, vp, vp, c.c_uint64, c.c_uint32, vp, vp]
    L.dgo_sg_plain_batch.argtypes = [vp, vp, c.c_uint64, c.c_uint32, c.c_uint32, vp, c.c_uint32,
                                     c.c_uint32, c.c_float, vp, c.c_uint32, vp, vp]
    L.dgo_sg_reuse_batch.argtypes = [vp, vp, c.c_uint64, c.c_uint32, c.c_uint32, vp, c.c_float, vp,
                                     c.c_uint32, vp, vp]
    L.dgo_reuse_negs.argtypes = [c.c_uint64, c.c_uint32, c.c_uint64, c.c_uint32, vp, c.c_uint32,
                                 c.c_uint32, vp]
    L.dgo_plain_batches.restype = c.c_uint64
    L.dgo_plain_batches.argtypes = [vp, c.c_uint64, vp, c.c_uint64]
    L.dgo_position_draws.argtypes = [c.c_uint64, c.c_uint64, c.c_uint32, c.c_uint32, c.c_uint32,
                                     c.c_uint32, vp, c.c_uint32, c.c_uint32,
                                     c.POINTER(c.c_uint32), vp]
    L.dgo_sg_shared.argtypes = [vp, vp, c.c_uint64, c.c_uint64, c.c_uint32, c.c_uint32, c.c_uint32,
                                vp, c.c_uint32, c.c_uint32, c.c_float, vp, c.c_int, c.c_uint64,
                                c.c_uint32, c.c_uint32, vp, vp]
    L.dgo_sg_shared_first_order.argtypes = L.dgo_sg_shared.argtypes
    L.dgo_tuple_forward.restype = c.c_double
    L.dgo_tuple_forward.argtypes = [vp, c.c_uint64, c.c_uint32, vp, vp, c.c_uint32, c.c_int, vp]
    L.dgo_tuple_grad_rows.argtypes = [vp, c.c_uint64, c.c_uint32, vp, vp, c.c_uint32, c.c_uint32,
                                      c.c_int, vp, c.c_int, vp, c.c_uint64, vp]
    L.dgo_tuple_fd.restype = c.c_double
    L.dgo_tuple_fd.argtypes = [vp, c.c_uint64, c.c_uint32, vp, vp, c.c_uint32, c.c_int, c.c_uint32,
                               c.c_uint32, c.c_double]
    L.dgo_tuple_inplace.argtypes = [vp, c.c_uint64, c.c_uint32, c.c_double, c.c_uint32, vp, vp]
    L.dgo_dsgl_keep_limits.argtypes = [vp, c.c_uint32, c.c_float, c.c_uint64, vp]
    L.dgo_dsgl_unigram_table.argtypes = [vp, c.c_uint32, c.c_uint32, vp]
    L.dgo_dsgl_train_f64.argtypes = [c.POINTER(DsglArgs), vp, vp]
    L.dgo_dsgl_train_f32.argtypes = [c.POINTER(DsglArgs), vp, vp, c.c_int]
    _L = L
    return L


class DsglArgs(ctypes.Structure):
    _fields_ = [("tokens", ctypes.c_void_p), ("offsets", ctypes.c_void_p), ("n_walks", ctypes.c_uint64),
                ("rank_of_vertex", ctypes.c_void_p), ("keep", ctypes.c_void_p),
                ("table", ctypes.c_void_p), ("table_size", ctypes.c_uint32), ("vocab", ctypes.c_uint32),
                ("expt", ctypes.c_void_p), ("exact_sigmoid", ctypes.c_int),
                ("dim", ctypes.c_uint32), ("window", ctypes.c_uint32), ("negative", ctypes.c_uint32),
                ("iter", ctypes.c_uint32), ("workers", ctypes.c_uint32), ("alpha0", ctypes.c_float),
                ("train_words", ctypes.c_uint64), ("seed", ctypes.c_uint64), ("epoch0", ctypes.c_uint32)]


_BASE = None


def baseline_lib():
    """The optimised FP32 DSGL restatement (CPU baseline build)."""
    global _BASE
    if _BASE is None:
        import os
        path = os.path.join(dgoracle.ORACLE_DIR, "libdgbaseline.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: build the oracle (make -C oracle)")
        B = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        B.dgo_dsgl_keep_limits.argtypes = [vp, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint64, vp]
        B.dgo_dsgl_unigram_table.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, vp]
        B.dgo_dsgl_train_f32.argtypes = [ctypes.POINTER(DsglArgs), vp, vp, ctypes.c_int]
        _BASE = B
    return _BASE


class Trainer:
    """Oracle state mirroring dg_sgns_init for a given count vector."""

    def __init__(self, counts, dim, min_count=0, table_size=1_000_000, sample=1e-3,
                 float_init=False):
        L = lib()
        counts = np.ascontiguousarray(counts, dtype=np.uint64)
        V = len(counts)
        self.vertex_of_rank = np.zeros(V, np.uint32)
        self.rank_of_vertex = np.zeros(V, np.uint32)
        self.cn_of_rank = np.zeros(V, np.uint64)
        n = ctypes.c_uint32()
        tw = ctypes.c_uint64()
        L.dgo_vocab(V, _vp(counts), min_count, _vp(self.vertex_of_rank), _vp(self.rank_of_vertex),
                    _vp(self.cn_of_rank), ctypes.byref(n), ctypes.byref(tw))
        self.vocab = n.value
        self.train_words = tw.value
        self.vertex_of_rank = self.vertex_of_rank[: self.vocab]
        self.cn_of_rank = self.cn_of_rank[: self.vocab]
        self.dim = dim
        self.syn0 = np.zeros((self.vocab, dim), np.float32)
        L.dgo_init_syn0(self.vocab, dim, int(float_init), _vp(self.syn0))
        self.syn1 = np.zeros((self.vocab, dim), np.float32)
        self.table_size = table_size
        self.table = np.zeros(table_size, np.int32)
        L.dgo_unigram_table(_vp(self.cn_of_rank), self.vocab, table_size, _vp(self.table))
        self.expt = np.zeros(1001, np.float32)
        L.dgo_exp_table(_vp(self.expt))
        self.keep = np.zeros(self.vocab, np.uint32)
        L.dgo_keep_limits(_vp(self.cn_of_rank), self.vocab, sample, self.train_words, _vp(self.keep))

    def subsample(self, tokens, offsets, seed, epoch):
        L = lib()
        tokens = np.ascontiguousarray(tokens, dtype=np.uint32)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        nw = len(offsets) - 1
        sen = np.zeros(max(len(tokens), 1), np.int32)
        off = np.zeros(nw + 1, np.uint64)
        n = L.dgo_subsample(_vp(tokens), _vp(offsets), nw, _vp(self.rank_of_vertex), _vp(self.keep),
                            seed, epoch, _vp(sen), _vp(off))
        return sen[:n].copy(), off

    def plain(self, sen, off, window, negative, alpha):
        """All reference batches in order, one alpha."""
        L = lib()
        nsen = len(off) - 1
        firsts = np.zeros(nsen + 2, np.uint64)
        nb = L.dgo_plain_batches(_vp(off), nsen, _vp(firsts), nsen + 1)
        bounds = list(firsts[: nb + 1])
        for b in range(nb):
            s0, s1 = int(bounds[b]), int(bounds[b + 1])
            sub_off = np.ascontiguousarray(off[s0: s1 + 1])
            L.dgo_sg_plain_batch(_vp(sen), _vp(sub_off), s1 - s0, window, negative, _vp(self.table),
                                 self.table_size, self.vocab, alpha, _vp(self.expt), self.dim,
                                 _vp(self.syn0), _vp(self.syn1))
        return nb

    def reuse(self, sen, off, window, negative, alpha, seed, epoch):
        """The negative-reuse form: per-sentence shared negatives, the full
        window, in the reference's batches, one alpha.  Returns the negatives."""
        L = lib()
        nsen = len(off) - 1
        negs = np.zeros(max(nsen * negative, 1), np.int32)
        L.dgo_reuse_negs(nsen, negative, seed, epoch, _vp(self.table), self.table_size, self.vocab,
                         _vp(negs))
        firsts = np.zeros(nsen + 2, np.uint64)
        nb = L.dgo_plain_batches(_vp(off), nsen, _vp(firsts), nsen + 1)
        bounds = list(firsts[: nb + 1])
        for b in range(nb):
            s0, s1 = int(bounds[b]), int(bounds[b + 1])
            sub_off = np.ascontiguousarray(off[s0: s1 + 1])
            sub_negs = np.ascontiguousarray(negs[s0 * negative: s1 * negative + 1])
            L.dgo_sg_reuse_batch(_vp(sen), _vp(sub_off), s1 - s0, window, negative, _vp(sub_negs),
                                 alpha, _vp(self.expt), self.dim, _vp(self.syn0), _vp(self.syn1))
        return negs[: nsen * negative]

    def shared(self, sen, off, window, negative, alpha, exact, seed, epoch, first_sen=0, draws=0):
        """Sequential SGNS with per-position shared negatives (the batched forms' order)."""
        L = lib()
        nsen = len(off) - 1
        L.dgo_sg_shared(_vp(sen), _vp(off), first_sen, nsen, window, negative, draws,
                        _vp(self.table), self.table_size, self.vocab, alpha, _vp(self.expt),
                        int(exact), seed, epoch, self.dim, _vp(self.syn0), _vp(self.syn1))

    def shared_first_order(self, sen, off, window, negative, alpha, exact, seed, epoch,
                           first_sen=0, draws=0):
        """The batched forms' first-order evaluation of the shared() order."""
        L = lib()
        nsen = len(off) - 1
        L.dgo_sg_shared_first_order(_vp(sen), _vp(off), first_sen, nsen, window, negative, draws,
                                    _vp(self.table), self.table_size, self.vocab, alpha,
                                    _vp(self.expt), int(exact), seed, epoch, self.dim,
                                    _vp(self.syn0), _vp(self.syn1))

    def dsgl_setup(self, sample=1e-3, table_size=None):
        """DSGL's unigram table (float d1) and keep limits (sqrt(r) + r)."""
        L = lib()
        ts = table_size or self.table_size
        self.dsgl_table = np.zeros(ts, np.int32)
        L.dgo_dsgl_unigram_table(_vp(self.cn_of_rank), self.vocab, ts, _vp(self.dsgl_table))
        self.dsgl_keep = np.zeros(self.vocab, np.uint32)
        L.dgo_dsgl_keep_limits(_vp(self.cn_of_rank), self.vocab, sample, self.train_words,
                               _vp(self.dsgl_keep))

    def dsgl_args(self, tokens, offsets, window, negative, alpha, workers, seed, epoch=0, iter=1,
                  exact=False):
        self._keep_alive = (np.ascontiguousarray(tokens, dtype=np.uint32),
                            np.ascontiguousarray(offsets, dtype=np.uint64))
        t, o = self._keep_alive
        return DsglArgs(t.ctypes.data, o.ctypes.data, len(o) - 1, self.rank_of_vertex.ctypes.data,
                        self.dsgl_keep.ctypes.data, self.dsgl_table.ctypes.data, len(self.dsgl_table),
                        self.vocab, self.expt.ctypes.data, int(exact), self.dim, window, negative,
                        iter, workers, alpha, self.train_words, seed, epoch)

    def dsgl(self, tokens, offsets, window, negative, alpha, workers, seed, epoch=0, iter=1,
             exact=False):
        """DSGL in FP64, the workers one after another."""
        a = self.dsgl_args(tokens, offsets, window, negative, alpha, workers, seed, epoch, iter, exact)
        if lib().dgo_dsgl_train_f64(ctypes.byref(a), _vp(self.syn0), _vp(self.syn1)) != 0:
            raise RuntimeError("dgo_dsgl_train_f64 failed")

    def dsgl_f32(self, tokens, offsets, window, negative, alpha, workers, seed, threads, epoch=0,
                 iter=1, exact=False, baseline=False):
        """DSGL in FP32 on `threads` host threads (Hogwild across workers)."""
        a = self.dsgl_args(tokens, offsets, window, negative, alpha, workers, seed, epoch, iter, exact)
        L = baseline_lib() if baseline else lib()
        if L.dgo_dsgl_train_f32(ctypes.byref(a), _vp(self.syn0), _vp(self.syn1), threads) != 0:
            raise RuntimeError("dgo_dsgl_train_f32 failed")

    def position_draws(self, seed, gs, q, epoch, window, draws):
        L = lib()
        b = ctypes.c_uint32()
        negs = np.zeros(max(draws, 1), np.int32)
        L.dgo_position_draws(seed, gs, q, epoch, window, draws, _vp(self.table), self.table_size,
                             self.vocab, ctypes.byref(b), _vp(negs))
        return b.value, negs[:draws].copy()


class Tuples:
    """FP64 oracle of the tuple interface over float tables (rows x dim)."""

    def __init__(self, tuples, K, syn0, syn1):
        self.t = np.ascontiguousarray(tuples, dtype=np.uint32)
        self.K = int(K)
        assert self.t.ndim == 2 and self.t.shape[1] == 2 + self.K
        self.s0 = np.ascontiguousarray(syn0, dtype=np.float32)
        self.s1 = np.ascontiguousarray(syn1, dtype=np.float32)
        self.dim = self.s0.shape[1]

    def forward(self, bf16=False):
        """(logits[n, 1+K], loss) in FP64."""
        logits = np.zeros((self.t.shape[0], 1 + self.K), np.float64)
        loss = lib().dgo_tuple_forward(_vp(self.t), self.t.shape[0], self.K, _vp(self.s0),
                                       _vp(self.s1), self.dim, int(bf16), _vp(logits))
        return logits, loss

    def grad_rows(self, table, rows, bf16=False, dlogits=None):
        r = np.ascontiguousarray(rows, dtype=np.uint32)
        out = np.zeros((len(r), self.dim), np.float64)
        dl = None if dlogits is None else np.ascontiguousarray(dlogits, dtype=np.float64)
        lib().dgo_tuple_grad_rows(_vp(self.t), self.t.shape[0], self.K, _vp(self.s0), _vp(self.s1),
                                  self.dim, self.s0.shape[0], int(bf16), _vp(dl), int(table), _vp(r),
                                  len(r), _vp(out))
        return out

    def fd(self, table, row, col, h=1e-3):
        return lib().dgo_tuple_fd(_vp(self.t), self.t.shape[0], self.K, _vp(self.s0), _vp(self.s1),
                                  self.dim, int(table), int(row), int(col), float(h))

    def inplace(self, lr):
        """Tables after the reference's sequential in-place pass, FP64."""
        x = self.s0.astype(np.float64)
        y = self.s1.astype(np.float64)
        lib().dgo_tuple_inplace(_vp(self.t), self.t.shape[0], self.K, float(lr), self.dim, _vp(x), _vp(y))
        return x, y
