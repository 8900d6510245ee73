"""mgl_modularity_dev: the modularity of any partition of a graph in device
memory (compute_modularity, louvain.cu:105-238) against the oracle: bit-equal
for unit weights (exact integer sums), within 1e-12 for float weights (FP64
sums whose order is the device's); malformed inputs are refused with
INVALID_GRAPH and nothing outside the arrays is touched."""
import numpy as np
import pytest

from hipmem import DeviceArray

pytestmark = pytest.mark.gpu


def dev_graph(g, c):
    bufs = [DeviceArray(g.offsets), DeviceArray(g.cols), DeviceArray(np.asarray(c, np.int32))]
    bufs.append(DeviceArray(g.w.astype(np.float32)) if g.w is not None else None)
    return bufs


@pytest.mark.parametrize("weighted", [False, True])
def test_modularity_dev_matches_oracle(ctx, mgl, orc, weighted):
    g = mgl.gen_rmat(15, 16, seed=41, weights=1 if weighted else 0)
    rng = np.random.default_rng(2)
    assigns = [np.arange(g.n, dtype=np.int32), np.zeros(g.n, np.int32),
               (rng.integers(0, g.n, g.n) // 64 * 64).astype(np.int32),
               orc.louvain(g.offsets, g.cols, None if g.w is None else g.w.astype(np.float64))
               .levels[0]["assign"]]
    for c in assigns:
        off, col, dc, dw = dev_graph(g, c)
        q = ctx.modularity_dev(g.n, g.nnz, off.ptr, col.ptr, dw.ptr if dw else 0, dc.ptr)
        if not weighted:
            k = orc.vertex_weights(g.offsets)
            a = orc.community_totals(k, c, g.n)
            assert q == orc.modularity(g.offsets, g.cols, None, c, a, g.nnz)
        else:
            q64 = orc.modularity_fp64(g.offsets, g.cols, g.w, c)
            assert abs(q - q64) <= 1e-12, (q, q64)
    # the singleton start and a one-community partition: closed forms
    off, col, dc, dw = dev_graph(g, np.zeros(g.n, np.int32))
    assert ctx.modularity_dev(g.n, g.nnz, off.ptr, col.ptr, dw.ptr if dw else 0, dc.ptr) == \
        pytest.approx(0.0, abs=1e-15)


def test_modularity_dev_empty_and_errors(ctx, mgl):
    g = mgl.gen_rmat(10, 8, seed=3)
    off, col, dc, _ = dev_graph(g, np.arange(g.n, dtype=np.int32))
    assert ctx.modularity_dev(0, 0, off.ptr, col.ptr, 0, dc.ptr) == 0.0
    bad = np.arange(g.n, dtype=np.int32)
    bad[5] = g.n  # community id out of range
    _, _, dbad, _ = dev_graph(g, bad)
    with pytest.raises(mgl.MglError) as e:
        ctx.modularity_dev(g.n, g.nnz, off.ptr, col.ptr, 0, dbad.ptr)
    assert e.value.code == mgl.INVALID_GRAPH and "community id" in str(e.value)
    cols = g.cols.copy()
    cols[7] = -3
    dcol = DeviceArray(cols)
    with pytest.raises(mgl.MglError) as e:
        ctx.modularity_dev(g.n, g.nnz, off.ptr, dcol.ptr, 0, dc.ptr)
    assert e.value.code == mgl.INVALID_GRAPH and "column" in str(e.value)
    o = g.offsets.copy()
    o[3] = g.nnz + 5  # a row past the entries
    doff = DeviceArray(o)
    with pytest.raises(mgl.MglError) as e:
        ctx.modularity_dev(g.n, g.nnz, doff.ptr, col.ptr, 0, dc.ptr)
    assert e.value.code == mgl.INVALID_GRAPH and "row" in str(e.value)
    with pytest.raises(mgl.MglError) as e:
        ctx.modularity_dev(-1, 0, off.ptr, col.ptr, 0, dc.ptr)
    assert e.value.code == mgl.INVALID_ARG
    # still usable
    assert ctx.modularity_dev(g.n, g.nnz, off.ptr, col.ptr, 0, dc.ptr) < 0
