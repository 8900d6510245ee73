"""GPU parity of the batched form (mgl_louvain_batched): every graph of a
batch against the CPU oracle run on that graph alone (levels, iteration
counts, modularity trajectories, per-level assignments, final labels), and
against the plain HIP path. The reference has no batched form; it runs one
graph per process (main.cu:118-131, louvain.cu:1473-1693), which is what
each graph's oracle run restates.

Batches exercise the union renumbering, the per-graph control state, the
shared fixed-point weight scale of a weighted batch (unit-weight graphs carry
2^S per entry), compaction of finished graphs, per-graph coarsening and the
hipGraph replays; the last test runs 256 graphs of the batched workload's own
generator in one call.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def compare(gpu, ref, what=""):
    assert len(gpu.levels) == len(ref.levels), (what, len(gpu.levels), len(ref.levels))
    for l, (g, r) in enumerate(zip(gpu.levels, ref.levels)):
        assert g.n == r["n"] and g.iters == r["iters"], (what, l, g.n, r["n"], g.iters,
                                                        r["iters"])
        assert g.q == r["q"], (what, l, g.q, r["q"])
        np.testing.assert_array_equal(g.qtraj, r["qtraj"], err_msg=what)
        if g.assign is not None:
            np.testing.assert_array_equal(g.assign, r["assign"], err_msg=what)
    np.testing.assert_array_equal(gpu.final_assign, ref.final, err_msg=what)


def oracle_run(orc, g, version=2):
    w = None if g.w is None else g.w.astype(np.float64)
    return orc.louvain(g.offsets, g.cols, w, mode=0, version=version)


def two_triangles(mgl):
    return mgl.CSR.from_edges(6, [(0, 1), (0, 2), (1, 2), (3, 4), (3, 5), (4, 5), (2, 3)])


@pytest.mark.parametrize("graph", [True, False])
def test_batched_unit_graphs(ctx, mgl, orc, graph):
    graphs = [mgl.gen_rmat(10 + (i % 3), 8 + 4 * (i % 3), seed=100 + i) for i in range(5)]
    graphs.append(two_triangles(mgl))
    graphs.append(mgl.gen_sbm(3000, 12, min_size=64, seed=11, weights=0, deg_in=16, deg_out=3))
    batch = ctx.louvain_batched(graphs, keep_levels=True, use_graph=graph)
    for i, (g, b) in enumerate(zip(graphs, batch)):
        ref = oracle_run(orc, g)
        compare(b, ref, f"graph {i}")
        compare(ctx.louvain(g, keep_levels=True, use_graph=False), ref, f"plain graph {i}")


def test_batched_weighted_and_unit_graphs(ctx, mgl, orc):
    # unit-weight graphs inside a batch with FP32 weights run in the batch's
    # fixed-point scale (each unit weight is 2^S); their 2m must be scaled too
    graphs = [mgl.gen_rmat(11, 24, seed=900, weights=1), mgl.gen_rmat(12, 25, seed=901),
              mgl.gen_sbm(3000, 12, min_size=64, seed=902, weights=0, deg_in=16, deg_out=3),
              mgl.gen_sbm(2000, 8, min_size=64, seed=903, weights=1, deg_in=16, deg_out=3)]
    batch = ctx.louvain_batched(graphs, keep_levels=True)
    for i, (g, b) in enumerate(zip(graphs, batch)):
        compare(b, oracle_run(orc, g), f"graph {i}")


@pytest.mark.parametrize("version", [0, 1, 2])
@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("fused", [True, False])
def test_batched_weighted_mixed_sizes(ctx, mgl, orc, graph, fused, version):
    # graphs of very different sizes finish at different iterations and levels,
    # which exercises the batch compaction and per-graph coarsening
    graphs = []
    for i in range(24):
        if i % 2 == 0:
            graphs.append(mgl.gen_sbm(500 + 300 * i, 4 + i, min_size=8, deg_in=16, deg_out=3,
                                      seed=200 + i, weights=1))
        else:
            graphs.append(mgl.gen_rmat(9 + (i % 5), 8 + i, seed=300 + i, weights=1,
                                       n_limit=400 + 250 * i))
    batch = ctx.louvain_batched(graphs, keep_levels=True, use_graph=graph,
                                fuse_modularity=fused, version=version)
    for i, (g, b) in enumerate(zip(graphs, batch)):
        compare(b, oracle_run(orc, g, version), f"graph {i}")


@pytest.mark.parametrize("fused", [True, False])
def test_binned_layout_batched(ctx, mgl, orc, fused):
    # a batch with compaction (graphs finish at different iterations), records
    # in degree-bin order
    graphs = []
    for i in range(12):
        if i % 2 == 0:
            graphs.append(mgl.gen_sbm(600 + 400 * i, 4 + i, min_size=8, deg_in=16, deg_out=3,
                                      seed=500 + i, weights=1))
        else:
            graphs.append(mgl.gen_rmat(9 + (i % 5), 8 + i, seed=600 + i, weights=1,
                                       n_limit=500 + 300 * i))
    batch = ctx.louvain_batched(graphs, keep_levels=True, fuse_modularity=fused,
                                record_layout=mgl.LAYOUT_BINNED)
    for i, (g, b) in enumerate(zip(graphs, batch)):
        compare(b, oracle_run(orc, g), f"graph {i}")


def test_batched_tie_smallest_and_max_iter(ctx, mgl, orc):
    graphs = [mgl.gen_batch_graph(i) for i in range(8)]
    batch = ctx.louvain_batched(graphs, keep_levels=True, tie_rule=mgl.TIE_SMALLEST)
    for i, (g, b) in enumerate(zip(graphs, batch)):
        ref = orc.louvain(g.offsets, g.cols, g.w.astype(np.float64), mode=0, tie=1)
        compare(b, ref, f"graph {i} smallest")
    batch = ctx.louvain_batched(graphs, keep_levels=True, max_iter=3)
    for i, (g, b) in enumerate(zip(graphs, batch)):
        ref = orc.louvain(g.offsets, g.cols, g.w.astype(np.float64), mode=0, max_iter=3)
        compare(b, ref, f"graph {i} max_iter")


@pytest.mark.timeout(900)
def test_batched_workload_256_graphs(ctx, mgl, orc):
    # the batched workload's first 256 graphs (1k-200k vertices, planted
    # partitions and R-MAT graphs, weights (0.5, 1.5]) in one call
    graphs = [mgl.gen_batch_graph(i) for i in range(256)]
    nv = sum(g.n for g in graphs)
    ne = sum(g.nnz for g in graphs)
    print(f"256 graphs: {nv} vertices, {ne} entries", flush=True)
    batch = ctx.louvain_batched(graphs, keep_levels=True)
    print(f"batched call: {batch[0].seconds:.3f} s", flush=True)
    for i, (g, b) in enumerate(zip(graphs, batch)):
        compare(b, oracle_run(orc, g), f"graph {i}")
        if i % 32 == 31:
            print(f"graphs 0..{i} equal to the oracle", flush=True)


@pytest.mark.parametrize("group", [1, 40_000, 300_000])
def test_batched_coarsening_in_groups(ctx, mgl, orc, tmp_path, group):
    # a batch is coarsened in groups of consecutive graphs whose kept entries
    # fit a bound (a graph larger than the bound is a group of its own): the
    # coarse graphs of every level and the runs equal the oracle's whatever
    # the grouping (group = 1: every graph alone)
    tf = tmp_path / "group.txt"
    tf.write_text(f"coarsen_group_entries {group}\n")
    graphs = []
    for i in range(16):
        if i % 2 == 0:
            graphs.append(mgl.gen_sbm(800 + 500 * i, 4 + i, min_size=8, deg_in=16, deg_out=3,
                                      seed=700 + i, weights=1))
        else:
            graphs.append(mgl.gen_rmat(10 + (i % 4), 8 + i, seed=800 + i,
                                       weights=1 if i % 3 else 0, n_limit=700 + 400 * i))
    graphs.append(mgl.CSR(np.zeros(41, np.int64), np.zeros(0, np.int32)))  # isolated only
    print("entries", [g.nnz for g in graphs], flush=True)
    batch = ctx.louvain_batched(graphs, keep_levels=True, keep_graphs=True,
                                tuning_file=str(tf))
    for i, (g, b) in enumerate(zip(graphs, batch)):
        w = None if g.w is None else g.w.astype(np.float64)
        ref = orc.louvain(g.offsets, g.cols, w, mode=0)
        compare(b, ref, f"graph {i}")
        # batch weights carry the batch's scale; the levels' graphs are in graph units
        S = 0 if w is None else orc.scale_of(w)
        for l in range(1, len(b.levels)):
            go, gc, gw = ref.graphs[l]
            np.testing.assert_array_equal(b.levels[l].graph.offsets, go, err_msg=f"{i}/{l}")
            np.testing.assert_array_equal(b.levels[l].graph.cols, gc, err_msg=f"{i}/{l}")
            np.testing.assert_array_equal(b.levels[l].graph_w64, np.ldexp(gw, -S),
                                          err_msg=f"{i}/{l}")
