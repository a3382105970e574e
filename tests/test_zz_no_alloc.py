"""fe_step_delta_noalloc, the step that allocates nothing, against
fe_step_delta (through the C ABI).

The entry point runs the same force kernel on the same plan as fe_step_delta;
only the heavy-row partials and split counters move into the library's
static scratch array. Its delta must equal fe_step_delta's bit for bit, on
two calls in a row where the second has the smaller input (fewer rows and
fewer partials), so stale scratch contents from the first call would show.
Device memory for the output is taken with the HIP runtime directly.
"""
import ctypes as C

import numpy as np
import pytest

from distembed_amd import _lib, api

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        _lib.load()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipStreamDestroy.argtypes = [C.c_void_p]
        self.h.hipStreamSynchronize.argtypes = [C.c_void_p]

    def ok(self, r, what):
        assert r == 0, f"{what}: hip error {r}"

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.ok(self.h.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        self.ok(self.h.hipMemset(p, 0xFF, nbytes), "hipMemset")  # NaN patterns: unwritten rows show
        return p.value

    def to_host(self, p, shape):
        out = np.empty(shape, dtype=np.float32)
        self.ok(self.h.hipMemcpy(out.ctypes.data, C.c_void_p(p), out.nbytes, 2), "hipMemcpy")
        return out


def _ctx(g, d, B, lists, tuning=None, seed=0):
    ctx = api.ForceEmbed(g.n, d)
    ctx.upload_csr(g)
    ctx.upload_x(api.init_uniform(g.n, d, seed=seed))
    ctx.set_negatives(lists)
    if tuning:
        ctx.set_tuning(**tuning)
    ctx.configure(B, 0.02, 5.0)
    return ctx


def test_noalloc_step_equals_step_delta_twice_second_smaller():
    hip = _Hip()
    # call 1: R-MAT hub rows cut into many block chunks (partials, counters)
    g1 = api.Graph.rmat(13, 16, seed=5)
    l1 = api.reference_negatives(g1.n, 40, 5)[0]
    c1 = _ctx(g1, 128, 1024, l1, tuning=dict(chunk=8, max_chunks=512))
    # call 2: a smaller graph and batch, fewer partials
    g2 = api.Graph.rmat(10, 8, seed=6)
    l2 = api.reference_negatives(g2.n, 20, 5)[0]
    c2 = _ctx(g2, 64, 100, l2, tuning=dict(chunk=4, max_chunks=8), seed=1)
    stream = C.c_void_p()
    hip.ok(hip.h.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    bufs = []
    try:
        cases = [(c1, 1024, 128, 2, 0, 0), (c2, 100, 64, 1, 10, stream.value)]
        for (ctx, B, d, epoch, batch, st) in cases:
            rows = min(B, ctx.n - batch * B)
            want = ctx.step_delta(epoch, batch, rows)  # prepares the context, too
            p = hip.malloc(rows * d * 4)
            bufs.append(p)
            ctx.step_delta_noalloc(epoch, batch, p, st)
            hip.ok(hip.h.hipStreamSynchronize(C.c_void_p(st) if st else None), "sync")
            ctx.synchronize()
            got = hip.to_host(p, (rows, d))
            assert np.isfinite(want).all()
            np.testing.assert_array_equal(got, want)
            # X is left unchanged, as by fe_step_delta
            np.testing.assert_array_equal(ctx.step_delta(epoch, batch, rows), want)
        # the first case again after the second: the scratch array is shared
        p = hip.malloc(1024 * 128 * 4)
        bufs.append(p)
        c1.step_delta_noalloc(0, 0, p)
        c1.synchronize()
        np.testing.assert_array_equal(hip.to_host(p, (1024, 128)), c1.step_delta(0, 0, 1024))
    finally:
        for p in bufs:
            hip.h.hipFree(C.c_void_p(p))
        hip.h.hipStreamDestroy(stream)
        c1.close()
        c2.close()


def test_noalloc_step_rejects_unprepared_context_and_oversized_batch():
    hip = _Hip()
    g = api.Graph.rmat(10, 8, seed=6)
    lists = api.reference_negatives(g.n, 20, 5)[0]
    ctx = _ctx(g, 64, 128, lists)
    p = hip.malloc(128 * 64 * 4)
    try:
        with pytest.raises(api.FeError) as e:
            ctx.step_delta_noalloc(0, 0, p)  # nothing prepared the plan yet
        assert e.value.status == _lib.FE_ERR_STATE
        ctx.step_delta(0, 0, 128)
        with pytest.raises(api.FeError) as e:
            ctx.step_delta_noalloc(0, 0, 0)  # no output
        assert e.value.status == _lib.FE_ERR_INVALID_ARG
    finally:
        hip.h.hipFree(C.c_void_p(p))
        ctx.close()
    # a batch whose partials exceed the array: 9 hubs of degree 4096 at
    # d = 512, cut into chunks of 4 entries (1024 partials of 2 KiB each)
    n = 4096 + 9
    u = np.repeat(np.arange(9), 4096)
    v = 9 + np.tile(np.arange(4096), 9)
    g = api.Graph.from_edges(n, u, v)
    lists = api.reference_negatives(n, 40, 3)[0]
    ctx = _ctx(g, 512, 16, lists, tuning=dict(chunk=1, max_chunks=1024))
    p = hip.malloc(16 * 512 * 4)
    try:
        ctx.step_delta(0, 0, 16)
        with pytest.raises(api.FeError) as e:
            ctx.step_delta_noalloc(0, 0, p)
        assert e.value.status == _lib.FE_ERR_INVALID_ARG
        assert "scratch" in str(e.value)
    finally:
        hip.h.hipFree(C.c_void_p(p))
        ctx.close()
    # a width held padded (d = 12 on rows of 16): refused before any launch
    # (the buffer has room for padded rows all the same)
    g = api.Graph.rmat(10, 8, seed=6)
    lists = api.reference_negatives(g.n, 20, 5)[0]
    ctx = _ctx(g, 12, 128, lists)
    p = hip.malloc(128 * 16 * 4)
    try:
        ctx.step_delta(0, 0, 128)
        with pytest.raises(api.FeError) as e:
            ctx.step_delta_noalloc(0, 0, p)
        assert e.value.status == _lib.FE_ERR_INVALID_ARG
        assert "padded" in str(e.value)
    finally:
        hip.h.hipFree(C.c_void_p(p))
        ctx.close()
