"""ctypes bindings for the CPU oracle (oracle/_build/libdgoracle*.so).

Test infrastructure: used only by the tests, the smoke check and the
benchmark's CPU-baseline leg, to check results.  Never on a product path.
"""

import ctypes
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle", "_build")

EDGE_DTYPE = np.dtype([("src", "<u4"), ("dst", "<u4"), ("data", "<f4")])


class _Graph(ctypes.Structure):
    _fields_ = [
        ("v_num", ctypes.c_uint32),
        ("m", ctypes.c_uint64),
        ("row_ptr", ctypes.c_void_p),
        ("nbr", ctypes.c_void_p),
        ("cn", ctypes.c_void_p),
        ("w", ctypes.c_void_p),
    ]


class _WalkCfg(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("min_length", ctypes.c_uint32),
        ("max_length", ctypes.c_uint32),
        ("mu", ctypes.c_double),
    ]


def _vp(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


_libs = {}


def load(variant="dg"):
    """variant 'dg' = dg_math.h functions (bitwise the GPU's), 'libm' = glibc."""
    if variant in _libs:
        return _libs[variant]
    name = "libdgoracle.so" if variant == "dg" else "libdgoracle_libm.so"
    path = os.path.join(ORACLE_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: build the oracle (make -C oracle)")
    L = ctypes.CDLL(path)
    c = ctypes
    vp = c.c_void_p
    L.dgo_build_csr.argtypes = [c.c_uint32, vp, c.c_uint64, c.c_int, vp, vp, vp]
    L.dgo_common_neighbors.argtypes = [c.c_uint32, vp, vp, vp]
    L.dgo_walk_one.restype = c.c_int64
    L.dgo_walk_one.argtypes = [c.POINTER(_Graph), c.POINTER(_WalkCfg), c.c_uint32, c.c_uint32, vp,
                               c.c_uint64, vp]
    L.dgo_walk_round.restype = c.c_int64
    L.dgo_walk_round.argtypes = [c.POINTER(_Graph), c.POINTER(_WalkCfg), c.c_uint32, c.c_uint32,
                                 c.c_uint32, vp, c.c_uint64, vp, c.c_int]
    L.dgo_relative_entropy.restype = c.c_double
    L.dgo_relative_entropy.argtypes = [c.c_uint32, vp, vp, c.POINTER(c.c_int)]
    L.dgo_write_walks.argtypes = [c.c_char_p, vp, vp, c.c_uint64, c.c_uint32]
    for fn in ("dgo_pow",):
        getattr(L, fn).restype = c.c_double
        getattr(L, fn).argtypes = [c.c_double, c.c_double]
    for fn in ("dgo_log2", "dgo_log", "dgo_sqrt", "dgo_tanh"):
        getattr(L, fn).restype = c.c_double
        getattr(L, fn).argtypes = [c.c_double]
    L.dgo_philox.argtypes = [c.c_uint32] * 6 + [vp]
    L.dgo_stream_draws.argtypes = [c.c_uint64, c.c_uint32, c.c_uint32, c.c_uint32, c.c_uint32,
                                   c.c_uint32, vp]
    _libs[variant] = L
    return L


class OracleGraph:
    """CSR built by the oracle from 12-byte edges (reference neighbour order)."""

    def __init__(self, v_num, edges, undirected=True, cn=None, variant="dg", weighted=False):
        L = load(variant)
        self.L = L
        edges = np.ascontiguousarray(edges, dtype=EDGE_DTYPE)
        m = len(edges) * (2 if undirected else 1)
        self.v_num = v_num
        self.row_ptr = np.zeros(v_num + 1, np.uint64)
        self.nbr = np.zeros(max(m, 1), np.uint32)
        self.data = np.zeros(max(m, 1), np.float32)
        rc = L.dgo_build_csr(v_num, _vp(edges), len(edges), int(undirected), _vp(self.row_ptr),
                             _vp(self.nbr), _vp(self.data))
        if rc != 0:
            raise ValueError("edge out of range")
        self.m = m
        if cn is None:
            self.cn = np.zeros(max(m, 1), np.uint32)
            L.dgo_common_neighbors(v_num, _vp(self.row_ptr), _vp(self.nbr), _vp(self.cn))
        elif isinstance(cn, str) and cn == "data":
            self.cn = self.data.astype(np.int64).astype(np.uint32)
        else:
            self.cn = np.ascontiguousarray(cn, dtype=np.uint32)
        # weighted: the edge data is w(u, v) and the acceptance tanh(pa * w)
        self.w = self.data if weighted else None
        self._g = _Graph(v_num, m, self.row_ptr.ctypes.data, self.nbr.ctypes.data,
                         self.cn.ctypes.data, self.w.ctypes.data if weighted else None)

    @classmethod
    def from_csr(cls, row_ptr, nbr, cn, w=None, variant="dg"):
        """Oracle view of a CSR built elsewhere (the device's, as dg_graph_get_csr
        returns it): the same neighbour order and common-neighbour counts;
        `w` the per-entry weights for the weighted acceptance."""
        g = cls.__new__(cls)
        g.L = load(variant)
        g.v_num = len(row_ptr) - 1
        g.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint64)
        g.nbr = np.ascontiguousarray(nbr, dtype=np.uint32)
        g.cn = np.ascontiguousarray(cn, dtype=np.uint32)
        g.m = len(g.nbr)
        g.w = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
        g.data = g.w
        g._g = _Graph(g.v_num, g.m, g.row_ptr.ctypes.data, g.nbr.ctypes.data, g.cn.ctypes.data,
                      g.w.ctypes.data if g.w is not None else None)
        return g

    @property
    def degrees(self):
        return np.diff(self.row_ptr).astype(np.uint32)

    def cfg(self, seed, min_length=20, max_length=0, mu=0.999):
        return _WalkCfg(seed, min_length, max_length, mu)

    def walk_one(self, cfg, round_no, walk_id, cap=1 << 16):
        path = np.zeros(cap, np.uint32)
        state = np.zeros(9, np.float64)
        n = self.L.dgo_walk_one(ctypes.byref(self._g), ctypes.byref(cfg), round_no, walk_id,
                                _vp(path), cap, _vp(state))
        if n < 0:
            raise RuntimeError("walk capacity exceeded")
        return path[:n].copy(), state

    def walk_round(self, cfg, round_no, walker_begin, walker_num, threads=1, tok_cap=None):
        if tok_cap is None:
            tok_cap = walker_num * 256 + 1024
        tok = np.zeros(tok_cap, np.uint32)
        off = np.zeros(walker_num + 1, np.uint64)
        n = self.L.dgo_walk_round(ctypes.byref(self._g), ctypes.byref(cfg), round_no, walker_begin,
                                  walker_num, _vp(tok), tok_cap, _vp(off), threads)
        if n < 0:
            raise RuntimeError("oracle walk round failed")
        return tok[:n].copy(), off

    def relative_entropy(self, freq):
        st = ctypes.c_int()
        freq = np.ascontiguousarray(freq, dtype=np.uint32)
        deg = np.ascontiguousarray(self.degrees)
        h = self.L.dgo_relative_entropy(self.v_num, _vp(deg), _vp(freq), ctypes.byref(st))
        return h, st.value


def walk_run(g, seed, walkers_per_round, init_round=5, min_length=20, max_length=0, mu=0.999,
             max_rounds=30, delta=0.001, converge=False, threads=1):
    """Oracle restatement of the round loop of include/walk.hpp:608-782 for one
    partition: round 1 runs walkers_per_round*init_round walks, later rounds
    walkers_per_round; after each round the relative entropy against the
    cumulative visit counts; stop after max_rounds (or on |dRE| <= delta)."""
    cfg = g.cfg(seed, min_length, max_length, mu)
    freq = np.zeros(g.v_num, np.uint64)
    out = []
    prev = None
    it = 1
    while True:
        nw = walkers_per_round * init_round if it == 1 else walkers_per_round
        tok, off = g.walk_round(cfg, it, 0, nw, threads=threads)
        freq += np.bincount(tok, minlength=g.v_num).astype(np.uint64)
        h, st = g.relative_entropy(freq.astype(np.uint32))
        d = h if prev is None else h - prev
        prev = h
        stop = it + 1 > max_rounds or (converge and abs(d) <= delta)
        out.append(dict(round=it, tokens=tok, offsets=off, re=h, delta_re=d, status=st, stop=stop))
        it += 1
        if stop:
            break
    return out


def csr_edge_data(v_num, edges, undirected=True):
    """The edges' data fields in the reference CSR order (doubled list, stable
    by source): the per-entry weights of the CSR built from `edges`."""
    L = load()
    edges = np.ascontiguousarray(edges, dtype=EDGE_DTYPE)
    m = len(edges) * (2 if undirected else 1)
    row_ptr = np.zeros(v_num + 1, np.uint64)
    data = np.zeros(max(m, 1), np.float32)
    if L.dgo_build_csr(v_num, _vp(edges), len(edges), int(undirected), _vp(row_ptr), None,
                       _vp(data)) != 0:
        raise ValueError("edge out of range")
    return row_ptr, data[:m]


def write_walks(path, tokens, offsets, threads=1, variant="dg"):
    L = load(variant)
    tokens = np.ascontiguousarray(tokens, dtype=np.uint32)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    rc = L.dgo_write_walks(path.encode(), _vp(tokens), _vp(offsets), len(offsets) - 1, threads)
    if rc != 0:
        raise OSError(path)


def philox(c, k):
    L = load()
    out = np.zeros(4, np.uint32)
    L.dgo_philox(*(list(c) + list(k)), _vp(out))
    return out


def stream_draws(seed, c1, c2, tag, n, bound=0):
    L = load()
    out = np.zeros(n, np.uint32)
    L.dgo_stream_draws(seed, c1, c2, tag, n, bound, _vp(out))
    return out
