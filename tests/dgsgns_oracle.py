"""ctypes bindings for the SGNS half of the CPU oracle (oracle/sgns_oracle.c).

Test infrastructure: the checker for the device trainer, never a product path.
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
    L.dgo_plain_batches.restype = c.c_uint64
    L.dgo_plain_batches.argtypes = [vp, c.c_uint64, vp, c.c_uint64]
    L.dgo_position_draws.argtypes = [c.c_uint64, c.c_uint64, c.c_uint32, c.c_uint32, c.c_uint32,
                                     c.c_uint32, vp, c.c_uint32, c.c_uint32,
                                     c.POINTER(c.c_uint32), vp]
    L.dgo_sg_shared.argtypes = [vp, vp, c.c_uint64, c.c_uint64, c.c_uint32, c.c_uint32, c.c_uint32,
                                vp, c.c_uint32, c.c_uint32, c.c_float, vp, c.c_int, c.c_uint64,
                                c.c_uint32, c.c_uint32, vp, vp]
    L.dgo_sg_shared_first_order.argtypes = L.dgo_sg_shared.argtypes
    L.dgo_tuple_loss.restype = c.c_double
    L.dgo_tuple_loss.argtypes = [vp, c.c_uint64, c.c_uint32, vp, vp, c.c_uint32]
    _L = L
    return L


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

    def position_draws(self, seed, gs, q, epoch, window, draws):
        L = lib()
        b = ctypes.c_uint32()
        negs = np.zeros(max(draws, 1), np.int32)
        L.dgo_position_draws(seed, gs, q, epoch, window, draws, _vp(self.table), self.table_size,
                             self.vocab, ctypes.byref(b), _vp(negs))
        return b.value, negs[:draws].copy()


def tuple_loss(tuples, K, syn0, syn1):
    L = lib()
    t = np.ascontiguousarray(tuples, dtype=np.uint32)
    s0 = np.ascontiguousarray(syn0, dtype=np.float64)
    s1 = np.ascontiguousarray(syn1, dtype=np.float64)
    return L.dgo_tuple_loss(_vp(t), t.shape[0], K, _vp(s0), _vp(s1), s0.shape[1])
