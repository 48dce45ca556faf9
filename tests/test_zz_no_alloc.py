"""The allocation-free tuple forward (dg_sgns_tuples_forward_noalloc) against
the allocating one (dg_sgns_tuples_forward, bf16 = 0): the same logits and
the same loss, bit for bit, over two calls in a row (the second smaller, so
the scratch array still holds loss partials of the first call's extra
wavefronts past those the second one writes); tuple counts beyond the
documented maximum and out-of-range rows are refused.  The test owns its
device buffers (hipMalloc through the HIP runtime), as such a caller would."""

import ctypes

import numpy as np
import pytest

import distger
from distger import ffi

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), max(nbytes, 8)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, n):
        out = np.zeros(n, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc(ctx, d_t, n, K, d_l, d_loss):
    return ctx.L.dg_sgns_tuples_forward_noalloc(ctx.h, d_t.p, n, K, d_l.p, d_loss.p)


def test_noalloc_forward_matches_the_allocating_one():
    rows, dim, K = 60_000, 128, 5
    hip = _hip()
    with distger.Context(0) as ctx:
        s0 = distger.gen_normal(rows * dim, 0.1, 7).reshape(rows, dim)
        s1 = distger.gen_normal(rows * dim, 0.1, 8).reshape(rows, dim)
        ctx.sgns_tables_create(s0, s1)
        n_max = 50_000
        d_t = DevBuf(hip, n_max * (K + 2) * 4)
        d_l = DevBuf(hip, n_max * (K + 1) * 4)
        d_loss = DevBuf(hip, 8)
        try:
            for seed, n in ((3, 50_000), (4, 7_001)):  # the second call smaller
                t = distger.gen_zipf(rows, 1.1, n * (K + 2), seed).reshape(n, K + 2)
                ref_logits, ref_loss = ctx.tuples_forward(t, K, bf16=False)
                d_t.put(t)
                assert _noalloc(ctx, d_t, n, K, d_l, d_loss) == ffi.DG_OK
                ctx.sync()
                logits = d_l.get(np.float32, n * (K + 1)).reshape(n, K + 1)
                loss = float(d_loss.get(np.float64, 1)[0])
                print(f"[no-alloc] {n} tuples: loss {loss!r} vs {ref_loss!r}", flush=True)
                assert np.array_equal(logits, ref_logits.reshape(n, K + 1))
                assert loss == ref_loss
            # beyond the fixed array, or a row out of range: refused
            assert _noalloc(ctx, d_t, (1 << 20) + 1, K, d_l, d_loss) == ffi.DG_ERR_ARG
            bad = t.copy()
            bad[17, 3] = rows
            d_t.put(bad)
            assert _noalloc(ctx, d_t, 7_001, K, d_l, d_loss) == ffi.DG_ERR_ARG
            assert b"out of range" in ctx.L.dg_last_error(ctx.h)
        finally:
            d_t.free()
            d_l.free()
            d_loss.free()
