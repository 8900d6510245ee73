"""mgl_modularity_dev_noalloc, the allocation-free variant of
mgl_modularity_dev, through the C ABI: the same result as the allocating
entry point (bit-equal for unit weights, whose sums are exact integers;
within 1e-12 for float weights, whose FP64 sums follow the device's order) on
two calls in a row, the second with a smaller graph (the working memory keeps
the last call's contents and must be cleared); the exact result also equals
the oracle's. Graphs above the vertex limit are refused before anything runs,
and repeated calls leave the device's free memory unchanged."""
import ctypes as C

import numpy as np
import pytest

from hipmem import DeviceArray, hip

pytestmark = pytest.mark.gpu


def _dev(g, c):
    return (DeviceArray(g.offsets), DeviceArray(g.cols), DeviceArray(np.asarray(c, np.int32)),
            DeviceArray(g.w.astype(np.float32)) if g.w is not None else None)


def _free_bytes():
    free, total = C.c_size_t(), C.c_size_t()
    h = hip()
    h.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    assert h.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("weighted", [False, True])
def test_noalloc_equals_allocating_entry_point(ctx, mgl, orc, weighted):
    wt = 1 if weighted else 0
    big = mgl.gen_rmat(17, 16, seed=61, weights=wt)
    small = mgl.gen_rmat(11, 8, seed=62, weights=wt)
    rng = np.random.default_rng(7)
    for g in (big, small):  # the second call has the smaller input
        assert g.n <= mgl.NOALLOC_MAX_VERTICES
        c = (rng.integers(0, g.n, g.n) // 16 * 16).astype(np.int32)
        off, col, dc, dw = _dev(g, c)
        pw = dw.ptr if dw else 0
        q_alloc = ctx.modularity_dev(g.n, g.nnz, off.ptr, col.ptr, pw, dc.ptr)
        q_free = ctx.modularity_dev_noalloc(g.n, g.nnz, off.ptr, col.ptr, pw, dc.ptr)
        print(f"n={g.n} weighted={weighted}: {q_alloc!r} {q_free!r}", flush=True)
        if not weighted:
            assert q_free == q_alloc
            k = orc.vertex_weights(g.offsets)
            a = orc.community_totals(k, c, g.n)
            assert q_free == orc.modularity(g.offsets, g.cols, None, c, a, g.nnz)
        else:
            assert abs(q_free - q_alloc) <= 1e-12 * max(1.0, abs(q_alloc))
            assert abs(q_free - orc.modularity_fp64(g.offsets, g.cols, g.w, c)) <= 1e-12


def test_noalloc_limits_and_memory(ctx, mgl):
    g = mgl.gen_rmat(12, 8, seed=63)
    c = np.arange(g.n, dtype=np.int32)
    off, col, dc, _ = _dev(g, c)
    q0 = ctx.modularity_dev_noalloc(g.n, g.nnz, off.ptr, col.ptr, 0, dc.ptr)
    before = _free_bytes()
    for _ in range(20):
        assert ctx.modularity_dev_noalloc(g.n, g.nnz, off.ptr, col.ptr, 0, dc.ptr) == q0
    assert _free_bytes() == before
    # above the limit: refused with a message, nothing launched (the arrays are
    # far smaller than the size claimed and are never read)
    with pytest.raises(mgl.MglError) as e:
        ctx.modularity_dev_noalloc(mgl.NOALLOC_MAX_VERTICES + 1, g.nnz, off.ptr, col.ptr, 0,
                                   dc.ptr)
    assert e.value.code == mgl.INVALID_ARG and "allocation-free" in str(e.value)
    # malformed input is reported as with the allocating entry point
    bad = c.copy()
    bad[3] = -1
    _, _, dbad, _ = _dev(g, bad)
    with pytest.raises(mgl.MglError) as e:
        ctx.modularity_dev_noalloc(g.n, g.nnz, off.ptr, col.ptr, 0, dbad.ptr)
    assert e.value.code == mgl.INVALID_GRAPH
    assert ctx.modularity_dev_noalloc(g.n, g.nnz, off.ptr, col.ptr, 0, dc.ptr) == q0
