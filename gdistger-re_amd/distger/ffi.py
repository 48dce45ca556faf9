"""ctypes bindings for include/distger_hip.h.

Every wrapper checks the returned status and raises DgError with the
library's message (dg_last_error) on failure.  Host buffers are numpy arrays
owned by the caller, as the C ABI requires.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libdistger_hip.so"

DG_OK = 0
DG_ERR_ARG, DG_ERR_HIP, DG_ERR_NOMEM, DG_ERR_IO, DG_ERR_STATE = -1, -2, -3, -4, -5
DG_ERR_WALK_TOO_LONG, DG_ERR_RE_UNDEFINED, DG_ERR_UNSUPPORTED = -6, -7, -8
DG_CN_FROM_DATA = 0
DG_CN_COMPUTE = 1
DG_CN_COMPUTE_WEIGHTED = 2
DG_ROUNDS_REFERENCE = 0
DG_ROUNDS_CONVERGE = 1
DG_MATH_SQRT, DG_MATH_LOG2, DG_MATH_LOG, DG_MATH_TANH, DG_MATH_POW = 0, 1, 2, 3, 4

# 12-byte binary edge record (include/type.hpp Edge<float>)
EDGE_DTYPE = np.dtype([("src", "<u4"), ("dst", "<u4"), ("data", "<f4")])

u8p = ctypes.POINTER(ctypes.c_uint8)


class DgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"distger error {code}: {msg}")
        self.code = code


class WalkParams(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("walkers_per_round", ctypes.c_uint32),
        ("init_round", ctypes.c_uint32),
        ("min_length", ctypes.c_uint32),
        ("max_length", ctypes.c_uint32),
        ("mu", ctypes.c_double),
        ("delta", ctypes.c_double),
        ("max_rounds", ctypes.c_uint32),
        ("rounds_mode", ctypes.c_uint32),
        ("length_limit", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class RoundStats(ctypes.Structure):
    _fields_ = [
        ("round", ctypes.c_uint32),
        ("walkers", ctypes.c_uint32),
        ("tokens", ctypes.c_uint64),
        ("steps", ctypes.c_uint64),
        ("re", ctypes.c_double),
        ("delta_re", ctypes.c_double),
        ("walk_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("max_walk_len", ctypes.c_uint32),
        ("stop", ctypes.c_uint32),
        ("draws", ctypes.c_uint64),
        ("moves", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


DG_SGNS_PLAIN_FP32 = 0
DG_SGNS_BATCHED_FP32 = 1
DG_SGNS_BATCHED_BF16 = 2
DG_SGNS_DSGL_FP32 = 3
DG_SGNS_NEGREUSE_FP32 = 4
DG_VOCAB_DEGREE = 0
DG_VOCAB_VISITS = 1
DG_ALPHA_LINEAR = 0
DG_ALPHA_FIXED = 1
DG_CORPUS_ROUND = 0
DG_CORPUS_TRAIN = 1
DG_LR_FIXED, DG_LR_EXPONENTIAL, DG_LR_COSINE, DG_LR_STEP = 0, 1, 2, 3


class SgnsParams(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_uint32),
        ("window", ctypes.c_uint32),
        ("negative", ctypes.c_uint32),
        ("min_count", ctypes.c_uint32),
        ("iter", ctypes.c_uint32),
        ("batch_size", ctypes.c_uint32),
        ("sample", ctypes.c_float),
        ("alpha", ctypes.c_float),
        ("form", ctypes.c_uint32),
        ("vocab_source", ctypes.c_uint32),
        ("alpha_schedule", ctypes.c_uint32),
        ("deterministic", ctypes.c_uint32),
        ("exact_sigmoid", ctypes.c_uint32),
        ("minibatch", ctypes.c_uint32),
        ("table_size", ctypes.c_uint32),
        ("init_float_ops", ctypes.c_uint32),
        ("negative_draws", ctypes.c_uint32),
        ("block_waves", ctypes.c_uint32),
        ("pool_split", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("workers", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class SgnsStats(ctypes.Structure):
    _fields_ = [
        ("words", ctypes.c_uint64),
        ("kept", ctypes.c_uint64),
        ("sentences", ctypes.c_uint64),
        ("launches", ctypes.c_uint64),
        ("ms", ctypes.c_double),
        ("alpha_end", ctypes.c_float),
        ("reserved", ctypes.c_float),
        ("pool_rows", ctypes.c_uint64),
    ]


_lib = None


def lib_path():
    return os.path.normpath(os.path.join(_HERE, "..", "lib", _LIB_NAME))


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype)) if a is not None else None


def load_library():
    """Load libdistger_hip.so from the package tree; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise DgError(-1, f"{path} is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(path)
    c = ctypes
    vp = c.c_void_p
    sig = {
        "dg_open": (c.c_int, [c.c_int, c.POINTER(vp)]),
        "dg_close": (None, [vp]),
        "dg_last_error": (c.c_char_p, [vp]),
        "dg_sync": (c.c_int, [vp]),
        "dg_graph_from_edges": (c.c_int, [vp, c.c_uint32, vp, c.c_uint64, c.c_int, c.c_int]),
        "dg_graph_load_bin": (c.c_int, [vp, c.c_char_p, c.c_uint32, c.c_int, c.c_int]),
        "dg_graph_info": (c.c_int, [vp, c.POINTER(c.c_uint32), c.POINTER(c.c_uint64)]),
        "dg_graph_get_csr": (c.c_int, [vp, vp, vp, vp]),
        "dg_graph_get_thresholds": (c.c_int, [vp, vp]),
        "dg_walk_params_default": (None, [c.POINTER(WalkParams), c.c_uint32]),
        "dg_walk_round": (c.c_int, [vp, c.POINTER(WalkParams), c.c_uint32, c.c_uint32,
                                    c.c_uint32, c.POINTER(RoundStats)]),
        "dg_walk_run": (c.c_int, [vp, c.POINTER(WalkParams), c.POINTER(RoundStats), c.c_uint32,
                                  c.POINTER(c.c_uint32), c.c_int]),
        "dg_walk_reset": (c.c_int, [vp]),
        "dg_corpus_info": (c.c_int, [vp, c.POINTER(c.c_uint64), c.POINTER(c.c_uint64)]),
        "dg_corpus_copy": (c.c_int, [vp, vp, vp]),
        "dg_corpus_write_text": (c.c_int, [vp, c.c_char_p, c.c_uint32]),
        "dg_vertex_freq": (c.c_int, [vp, vp]),
        "dg_corpus_append": (c.c_int, [vp]),
        "dg_train_corpus_info": (c.c_int, [vp, c.POINTER(c.c_uint64), c.POINTER(c.c_uint64)]),
        "dg_train_corpus_clear": (c.c_int, [vp]),
        "dg_timing_get": (c.c_int, [vp, c.c_char_p, c.POINTER(c.c_uint64), c.POINTER(c.c_double)]),
        "dg_timing_reset": (c.c_int, [vp]),
        "dg_gen_chung_lu": (c.c_int, [c.c_uint32, c.c_uint64, c.c_double, c.c_uint32, c.c_uint64,
                                      vp, c.c_uint64, c.POINTER(c.c_uint64)]),
        "dg_gen_rmat": (c.c_int, [c.c_uint32, c.c_uint32, c.c_double, c.c_double, c.c_double,
                                  c.c_uint64, vp, c.c_uint64, c.POINTER(c.c_uint64)]),
        "dg_lp_split": (c.c_int, [c.c_uint32, vp, c.c_uint64, c.c_double, c.c_uint64,
                                  c.POINTER(c.c_uint64), vp, c.c_uint64, c.POINTER(c.c_uint64)]),
        "dg_auc": (c.c_double, [vp, c.c_uint64, vp, c.c_uint64]),
        "dg_copy_bench": (c.c_int, [vp, c.c_uint64, c.c_uint32, c.POINTER(c.c_double)]),
        "dg_math_eval": (c.c_int, [vp, c.c_int, vp, vp, c.c_uint64, vp]),
        "dg_sgns_params_default": (None, [c.POINTER(SgnsParams)]),
        "dg_sgns_init": (c.c_int, [vp, c.POINTER(SgnsParams)]),
        "dg_sgns_train": (c.c_int, [vp, c.c_int, c.POINTER(SgnsParams), c.c_float,
                                    c.POINTER(SgnsStats)]),
        "dg_sgns_autotune": (c.c_int, [vp, c.c_int, c.POINTER(SgnsParams), c.c_uint64,
                                       c.POINTER(SgnsParams), vp]),
        "dg_pipeline_step": (c.c_int, [vp, c.POINTER(WalkParams), c.c_uint32, c.c_uint32,
                                       c.c_uint32, c.POINTER(SgnsParams), c.c_float,
                                       c.POINTER(RoundStats), c.POINTER(SgnsStats)]),
        "dg_pipeline_info": (c.c_int, [vp, c.POINTER(c.c_uint64), c.POINTER(c.c_uint64),
                                       c.POINTER(c.c_uint64)]),
        "dg_sgns_vocab": (c.c_int, [vp, vp, c.POINTER(c.c_uint32)]),
        "dg_sgns_get_tables": (c.c_int, [vp, vp, vp]),
        "dg_sgns_set_tables": (c.c_int, [vp, vp, vp]),
        "dg_sgns_score_pairs": (c.c_int, [vp, vp, vp, c.c_uint64, vp]),
        "dg_emb_write": (c.c_int, [vp, c.c_char_p, c.c_int]),
        "dg_sgns_tuples_forward": (c.c_int, [vp, vp, c.c_uint64, c.c_uint32, c.c_int, vp,
                                             c.POINTER(c.c_double)]),
        "dg_sgns_tuples_backward": (c.c_int, [vp, vp, c.c_uint64, c.c_uint32, c.c_int, vp,
                                              c.c_float, c.c_int]),
        "dg_sgns_tuples_step": (c.c_int, [vp, vp, c.c_uint64, c.c_uint32, c.c_int, c.c_float, c.c_int,
                                          vp, c.POINTER(c.c_double)]),
        "dg_sgns_tables_create": (c.c_int, [vp, c.c_uint32, c.c_uint32, vp, vp]),
        "dg_gen_zipf": (c.c_int, [c.c_uint32, c.c_double, c.c_uint64, c.c_uint64, vp]),
        "dg_gen_normal": (c.c_int, [c.c_uint64, c.c_double, c.c_uint64, vp]),
        "dg_sgns_grad_rows": (c.c_int, [vp, c.c_int, vp, c.c_uint64, vp]),
        "dg_write_edges_bin": (c.c_int, [c.c_char_p, vp, c.c_uint64]),
        "dg_read_edges_bin": (c.c_int, [c.c_char_p, vp, c.c_uint64, c.POINTER(c.c_uint64)]),
        "dg_convert_text_edges": (c.c_int, [c.c_char_p, c.c_char_p, c.c_int, c.POINTER(c.c_uint64)]),
        "dg_lp_write_files": (c.c_int, [c.c_char_p, c.c_char_p, c.c_char_p, vp, c.c_uint64, vp,
                                        c.c_uint64, vp, c.c_uint64]),
        "dg_lr_value": (c.c_float, [c.c_int, c.c_float, c.c_float, c.c_float, c.c_int]),
        "dg_sgns_last_alphas": (c.c_int, [vp, vp, c.c_uint32, c.POINTER(c.c_uint32)]),
        "dg_sgns_tuples_forward_noalloc": (c.c_int, [vp, c.c_void_p, c.c_uint64, c.c_uint32,
                                                     c.c_void_p, c.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _vp(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _check_host(rc, what):
    if rc != DG_OK:
        raise DgError(rc, what)


class Context:
    """One device context (dg_ctx).  Use as a context manager or call close()."""

    def __init__(self, device=0):
        self.L = load_library()
        h = ctypes.c_void_p()
        rc = self.L.dg_open(device, ctypes.byref(h))
        if rc != DG_OK:
            raise DgError(rc, f"dg_open(device={device}) failed")
        self.h = h

    def close(self):
        if self.h:
            self.L.dg_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != DG_OK:
            msg = self.L.dg_last_error(self.h)
            raise DgError(rc, msg.decode() if msg else "")

    def sync(self):
        self._check(self.L.dg_sync(self.h))

    # ---- graph
    def graph_from_edges(self, v_num, edges, undirected=True, cn_mode=DG_CN_COMPUTE):
        edges = np.ascontiguousarray(edges, dtype=EDGE_DTYPE)
        self._check(self.L.dg_graph_from_edges(self.h, v_num, _vp(edges), len(edges),
                                               int(undirected), cn_mode))

    def graph_load_bin(self, path, v_num, undirected=True, cn_mode=DG_CN_FROM_DATA):
        self._check(self.L.dg_graph_load_bin(self.h, path.encode(), v_num, int(undirected), cn_mode))

    def graph_info(self):
        v = ctypes.c_uint32()
        m = ctypes.c_uint64()
        self._check(self.L.dg_graph_info(self.h, ctypes.byref(v), ctypes.byref(m)))
        return v.value, m.value

    def graph_csr(self):
        v, m = self.graph_info()
        rp = np.zeros(v + 1, np.uint64)
        nbr = np.zeros(m, np.uint32)
        cn = np.zeros(m, np.uint32)
        self._check(self.L.dg_graph_get_csr(self.h, _vp(rp), _vp(nbr), _vp(cn)))
        return rp, nbr, cn

    def graph_thresholds(self):
        _, m = self.graph_info()
        thr = np.zeros(m, np.uint32)
        self._check(self.L.dg_graph_get_thresholds(self.h, _vp(thr)))
        return thr

    # ---- walker
    def default_walk_params(self, v_num=None, **kw):
        if v_num is None:
            v_num = self.graph_info()[0]
        p = WalkParams()
        self.L.dg_walk_params_default(ctypes.byref(p), v_num)
        for k, val in kw.items():
            setattr(p, k, val)
        return p

    def walk_reset(self):
        self._check(self.L.dg_walk_reset(self.h))

    def walk_round(self, params, round_no, walker_begin, walker_num):
        st = RoundStats()
        self._check(self.L.dg_walk_round(self.h, ctypes.byref(params), round_no, walker_begin,
                                         walker_num, ctypes.byref(st)))
        return st

    def walk_run(self, params, keep_corpus=False, max_stats=256):
        arr = (RoundStats * max_stats)()
        n = ctypes.c_uint32()
        self._check(self.L.dg_walk_run(self.h, ctypes.byref(params), arr, max_stats,
                                       ctypes.byref(n), int(keep_corpus)))
        return [arr[i] for i in range(min(n.value, max_stats))]

    def corpus(self):
        nw = ctypes.c_uint64()
        nt = ctypes.c_uint64()
        self._check(self.L.dg_corpus_info(self.h, ctypes.byref(nw), ctypes.byref(nt)))
        tok = np.zeros(nt.value, np.uint32)
        off = np.zeros(nw.value + 1, np.uint64)
        self._check(self.L.dg_corpus_copy(self.h, _vp(tok), _vp(off)))
        return tok, off

    def corpus_write_text(self, path, threads=1):
        self._check(self.L.dg_corpus_write_text(self.h, path.encode(), threads))

    def vertex_freq(self):
        v, _ = self.graph_info()
        f = np.zeros(v, np.uint32)
        self._check(self.L.dg_vertex_freq(self.h, _vp(f)))
        return f

    def corpus_append(self):
        self._check(self.L.dg_corpus_append(self.h))

    def train_corpus_info(self):
        nw = ctypes.c_uint64()
        nt = ctypes.c_uint64()
        self._check(self.L.dg_train_corpus_info(self.h, ctypes.byref(nw), ctypes.byref(nt)))
        return nw.value, nt.value

    def train_corpus_clear(self):
        self._check(self.L.dg_train_corpus_clear(self.h))

    # ---- SGNS trainer
    def default_sgns_params(self, **kw):
        p = SgnsParams()
        self.L.dg_sgns_params_default(ctypes.byref(p))
        for k, val in kw.items():
            setattr(p, k, val)
        return p

    def sgns_init(self, params):
        self._check(self.L.dg_sgns_init(self.h, ctypes.byref(params)))

    def sgns_train(self, params, corpus=DG_CORPUS_TRAIN, alpha=0.0):
        st = SgnsStats()
        self._check(self.L.dg_sgns_train(self.h, corpus, ctypes.byref(params), float(alpha),
                                         ctypes.byref(st)))
        return st

    def sgns_autotune(self, params, corpus=DG_CORPUS_TRAIN, sample_walks=0):
        """Fastest launch shape for `params` on the corpus: (tuned params,
        default words/s, tuned words/s).  Leaves the tables as they were."""
        tuned = SgnsParams()
        rates = np.zeros(2, np.float64)
        self._check(self.L.dg_sgns_autotune(self.h, corpus, ctypes.byref(params), sample_walks,
                                            ctypes.byref(tuned), _vp(rates)))
        return tuned, float(rates[0]), float(rates[1])

    def pipeline_step(self, wparams, round_no, walker_begin, walker_num, sparams, alpha=0.0,
                      allow_undefined_re=False):
        """One walk round and one trainer pass over its corpus as one HIP graph
        (the same results as walk_round followed by sgns_train on the round).
        allow_undefined_re: a round whose relative entropy is undefined (a
        vertex never visited, e.g. an isolated one) returns its statistics
        instead of raising; the round and its training are complete either way."""
        rs = RoundStats()
        ss = SgnsStats()
        rc = self.L.dg_pipeline_step(self.h, ctypes.byref(wparams), round_no, walker_begin,
                                     walker_num, ctypes.byref(sparams), float(alpha),
                                     ctypes.byref(rs), ctypes.byref(ss))
        if not (allow_undefined_re and rc == DG_ERR_RE_UNDEFINED):
            self._check(rc)
        return rs, ss

    def pipeline_info(self):
        """(instantiations, in-place updates, launches) of the step's graph."""
        a, b, n = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self.L.dg_pipeline_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)))
        return a.value, b.value, n.value

    def sgns_last_alphas(self):
        n = ctypes.c_uint32()
        self._check(self.L.dg_sgns_last_alphas(self.h, None, 0, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1), np.float32)
        self._check(self.L.dg_sgns_last_alphas(self.h, _vp(out), n.value, ctypes.byref(n)))
        return out[: n.value]

    def sgns_vocab(self):
        n = ctypes.c_uint32()
        self._check(self.L.dg_sgns_vocab(self.h, None, ctypes.byref(n)))
        v = np.zeros(n.value, np.uint32)
        self._check(self.L.dg_sgns_vocab(self.h, _vp(v), ctypes.byref(n)))
        return v

    def sgns_tables(self, dim):
        n = len(self.sgns_vocab())
        s0 = np.zeros((n, dim), np.float32)
        s1 = np.zeros((n, dim), np.float32)
        self._check(self.L.dg_sgns_get_tables(self.h, _vp(s0), _vp(s1)))
        return s0, s1

    def sgns_set_tables(self, syn0, syn1):
        s0 = np.ascontiguousarray(syn0, dtype=np.float32)
        s1 = np.ascontiguousarray(syn1, dtype=np.float32)
        self._check(self.L.dg_sgns_set_tables(self.h, _vp(s0), _vp(s1)))

    def sgns_score_pairs(self, src, dst):
        src = np.ascontiguousarray(src, dtype=np.uint32)
        dst = np.ascontiguousarray(dst, dtype=np.uint32)
        out = np.zeros(len(src), np.float32)
        self._check(self.L.dg_sgns_score_pairs(self.h, _vp(src), _vp(dst), len(src), _vp(out)))
        return out

    def emb_write(self, path, binary=False):
        self._check(self.L.dg_emb_write(self.h, path.encode(), int(binary)))

    def tuples_forward(self, tuples, K, bf16=False):
        t = np.ascontiguousarray(tuples, dtype=np.uint32)
        n = t.shape[0]
        logits = np.zeros(n * (1 + K), np.float32)
        loss = ctypes.c_double()
        self._check(self.L.dg_sgns_tuples_forward(self.h, _vp(t), n, K, int(bf16), _vp(logits),
                                                  ctypes.byref(loss)))
        return logits.reshape(n, 1 + K), loss.value

    def tuples_backward(self, tuples, K, bf16=False, dlogits=None, lr=0.0, apply=False):
        t = np.ascontiguousarray(tuples, dtype=np.uint32)
        dl = None if dlogits is None else np.ascontiguousarray(dlogits, dtype=np.float32)
        if dl is not None and dl.size != t.shape[0] * (K + 1):
            raise ValueError("dlogits needs n * (K + 1) values")
        self._check(self.L.dg_sgns_tuples_backward(self.h, _vp(t), t.shape[0], K, int(bf16), _vp(dl),
                                                   float(lr), int(apply)))

    def tuples_step(self, tuples, K, bf16=False, lr=0.0, apply=False):
        """Forward and own-gradient backward in one call: (logits, loss), the
        gradients readable through grad_rows, the update applied if asked."""
        t = np.ascontiguousarray(tuples, dtype=np.uint32)
        n = t.shape[0]
        logits = np.zeros(n * (1 + K), np.float32)
        loss = ctypes.c_double()
        self._check(self.L.dg_sgns_tuples_step(self.h, _vp(t), n, K, int(bf16), float(lr), int(apply),
                                               _vp(logits), ctypes.byref(loss)))
        return logits.reshape(n, 1 + K), loss.value

    def sgns_tables_create(self, syn0, syn1):
        s0 = np.ascontiguousarray(syn0, dtype=np.float32)
        s1 = np.ascontiguousarray(syn1, dtype=np.float32)
        if s0.shape != s1.shape or s0.ndim != 2:
            raise ValueError("syn0 and syn1 must be equal rows x dim arrays")
        self._check(self.L.dg_sgns_tables_create(self.h, s0.shape[0], s0.shape[1], _vp(s0), _vp(s1)))

    def grad_rows(self, table, rows, dim):
        r = np.ascontiguousarray(rows, dtype=np.uint32)
        out = np.zeros((len(r), dim), np.float32)
        self._check(self.L.dg_sgns_grad_rows(self.h, table, _vp(r), len(r), _vp(out)))
        return out

    def math_eval(self, fn, x, y=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        yy = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros_like(x)
        self._check(self.L.dg_math_eval(self.h, int(fn), _vp(x), _vp(yy), len(x), _vp(out)))
        return out

    def copy_bench(self, nbytes, iters=3):
        g = ctypes.c_double()
        self._check(self.L.dg_copy_bench(self.h, int(nbytes), iters, ctypes.byref(g)))
        return g.value

    # ---- timing
    def timing(self, name):
        n = ctypes.c_uint64()
        ms = ctypes.c_double()
        self._check(self.L.dg_timing_get(self.h, name.encode(), ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def timing_reset(self):
        self._check(self.L.dg_timing_reset(self.h))


# ---- host utilities (no device needed) -------------------------------------

def gen_chung_lu(v_num, e_target, gamma, min_degree, seed):
    L = load_library()
    cap = int(e_target) + int(v_num) * max(int(min_degree), 1) + 16
    buf = np.zeros(cap, EDGE_DTYPE)
    n = ctypes.c_uint64()
    _check_host(L.dg_gen_chung_lu(v_num, int(e_target), float(gamma), min_degree, seed, _vp(buf),
                                  cap, ctypes.byref(n)), "dg_gen_chung_lu")
    return buf[: n.value].copy()


def gen_rmat(scale, edge_factor, a, b, c, seed):
    L = load_library()
    cap = (int(edge_factor) << int(scale)) + 16
    buf = np.zeros(cap, EDGE_DTYPE)
    n = ctypes.c_uint64()
    _check_host(L.dg_gen_rmat(scale, edge_factor, a, b, c, seed, _vp(buf), cap, ctypes.byref(n)),
                "dg_gen_rmat")
    return buf[: n.value].copy()


def gen_zipf(v_num, s, n, seed):
    """n ids in [0, v_num), P(k) proportional to (k + 1)^-s."""
    L = load_library()
    out = np.zeros(int(n), np.uint32)
    _check_host(L.dg_gen_zipf(v_num, float(s), int(n), seed, _vp(out)), "dg_gen_zipf")
    return out


def gen_normal(n, sigma, seed):
    L = load_library()
    out = np.zeros(int(n), np.float32)
    _check_host(L.dg_gen_normal(int(n), float(sigma), seed, _vp(out)), "dg_gen_normal")
    return out


def lp_split(v_num, edges, frac, seed):
    """Returns (train_edges, pos_edges, neg_edges)."""
    L = load_library()
    e = np.ascontiguousarray(edges, dtype=EDGE_DTYPE).copy()
    npos = int(frac * len(e))
    neg = np.zeros(max(npos, 1), EDGE_DTYPE)
    ntr = ctypes.c_uint64()
    nn = ctypes.c_uint64()
    _check_host(L.dg_lp_split(v_num, _vp(e), len(e), float(frac), seed, ctypes.byref(ntr),
                              _vp(neg), len(neg), ctypes.byref(nn)), "dg_lp_split")
    return e[: ntr.value].copy(), e[ntr.value:].copy(), neg[: nn.value].copy()


def auc(pos_scores, neg_scores):
    L = load_library()
    p = np.ascontiguousarray(pos_scores, dtype=np.float32)
    n = np.ascontiguousarray(neg_scores, dtype=np.float32)
    return L.dg_auc(_vp(p), len(p), _vp(n), len(n))


def write_edges_bin(path, edges):
    L = load_library()
    e = np.ascontiguousarray(edges, dtype=EDGE_DTYPE)
    _check_host(L.dg_write_edges_bin(path.encode(), _vp(e), len(e)), "dg_write_edges_bin")


def read_edges_bin(path):
    L = load_library()
    n = ctypes.c_uint64()
    _check_host(L.dg_read_edges_bin(path.encode(), None, 0, ctypes.byref(n)), "dg_read_edges_bin")
    buf = np.zeros(n.value, EDGE_DTYPE)
    _check_host(L.dg_read_edges_bin(path.encode(), _vp(buf), n.value, ctypes.byref(n)),
                "dg_read_edges_bin")
    return buf


def convert_text_edges(in_path, out_path, weighted):
    """Text edge list -> binary edge file (12-byte records if weighted, else 8)."""
    L = load_library()
    n = ctypes.c_uint64()
    _check_host(L.dg_convert_text_edges(in_path.encode(), out_path.encode(), int(weighted),
                                        ctypes.byref(n)), "dg_convert_text_edges")
    return n.value


def lp_write_files(train_path, test_edges_path, test_labels_path, train, pos, neg):
    L = load_library()
    arrs = [np.ascontiguousarray(x, dtype=EDGE_DTYPE) for x in (train, pos, neg)]
    enc = [p.encode() if p else None for p in (train_path, test_edges_path, test_labels_path)]
    _check_host(L.dg_lp_write_files(*enc, _vp(arrs[0]), len(arrs[0]), _vp(arrs[1]), len(arrs[1]),
                                    _vp(arrs[2]), len(arrs[2])), "dg_lp_write_files")


def lr_value(kind, initial, a=0.0, b=0.0, step=0):
    return load_library().dg_lr_value(int(kind), float(initial), float(a), float(b), int(step))
