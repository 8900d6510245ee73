"""Bit parity of the headline workload: R-MAT scale 24, edge factor 32,
(0.57, 0.19, 0.19, 0.05), seed 3, unit weights (1.02e9 directed entries),
version 2 with fused modularity, against the CPU oracle: the level count,
every level's vertex count, iteration count, modularity trajectory and
assignment, the level-1 coarse CSR and the final labels. The oracle takes a
minute or two on the host's cores."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(1500)
def test_rmat24_bit_parity(ctx, mgl, orc):
    t0 = time.time()
    g = mgl.gen_rmat(24, 32, seed=3)
    print(f"R-MAT 24: n={g.n} nnz={g.nnz} ({time.time() - t0:.0f} s)", flush=True)
    res = ctx.louvain(g, keep_levels=True, keep_graphs=True)
    print(f"GPU: {len(res.levels)} levels {[l.iters for l in res.levels]} Q={res.final_q!r} "
          f"{res.seconds:.2f} s", flush=True)
    t0 = time.time()
    ref = orc.louvain(g.offsets, g.cols, None, mode=0, version=2)
    print(f"oracle: {len(ref.levels)} levels {[l['iters'] for l in ref.levels]} "
          f"Q={ref.final_q!r} ({time.time() - t0:.0f} s)", flush=True)
    assert len(res.levels) == len(ref.levels)
    for l, (lg, lr) in enumerate(zip(res.levels, ref.levels)):
        assert lg.n == lr["n"] and lg.iters == lr["iters"], l
        assert lg.q == lr["q"], l
        np.testing.assert_array_equal(lg.qtraj, lr["qtraj"])
        np.testing.assert_array_equal(lg.assign, lr["assign"])
        print(f"level {l}: n={lg.n} iters={lg.iters} Q={lg.q!r} equal", flush=True)
    go, gc, gw = ref.graphs[1]
    g1 = res.levels[1].graph
    np.testing.assert_array_equal(g1.offsets, go)
    np.testing.assert_array_equal(g1.cols, gc)
    np.testing.assert_array_equal(res.levels[1].graph_w64, gw)
    print(f"level-1 coarse CSR equal: n={len(go) - 1} nnz={len(gc)}", flush=True)
    np.testing.assert_array_equal(res.final_assign, ref.final)
    assert res.final_q == ref.final_q
