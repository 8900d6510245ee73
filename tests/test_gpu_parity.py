"""GPU parity: the HIP path (through libmgl.so's C ABI) against the CPU oracle.

Exact modes must be bit-exact: community assignments, coarse CSRs, iteration
counts and every modularity value of the trajectory. Inputs are seeded
synthetic graphs at sizes the oracle finishes in seconds, plus hand-built edge
cases (empty graph, isolated vertices, self-loops, stars whose hub takes the
HBM-hash bin).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def two_triangles(mgl):
    return mgl.CSR.from_edges(6, [(0, 1), (0, 2), (1, 2), (3, 4), (3, 5), (4, 5), (2, 3)])


def compare(gpu, ref, check_graphs=True, scale=0):
    assert len(gpu.levels) == len(ref.levels), (len(gpu.levels), len(ref.levels))
    for l, (g, r) in enumerate(zip(gpu.levels, ref.levels)):
        assert g.n == r["n"], (l, g.n, r["n"])
        assert g.iters == r["iters"], (l, g.iters, r["iters"])
        assert g.q == r["q"], (l, g.q, r["q"])
        np.testing.assert_array_equal(g.qtraj, r["qtraj"])
        if g.assign is not None:
            np.testing.assert_array_equal(g.assign, r["assign"])
        if check_graphs and l >= 1 and g.graph is not None:
            go, gc, gw = ref.graphs[l]
            np.testing.assert_array_equal(g.graph.offsets, go)
            np.testing.assert_array_equal(g.graph.cols, gc)
            np.testing.assert_array_equal(g.graph_w64, np.ldexp(gw, -scale))
    np.testing.assert_array_equal(gpu.final_assign, ref.final)


def compare_fp64(gpu, ref, tol=1e-12):
    """FP64 mode on dyadic weights: every sum of weights is exact, so the gains
    and hence the assignments, iteration counts and coarse graphs are the
    oracle's bit for bit; sum a^2 is not exact in FP64 and its rounding depends
    on the summation order, so modularity values agree to tol."""
    assert len(gpu.levels) == len(ref.levels), (len(gpu.levels), len(ref.levels))
    for l, (g, r) in enumerate(zip(gpu.levels, ref.levels)):
        assert g.n == r["n"] and g.iters == r["iters"], (l, g.n, r["n"], g.iters, r["iters"])
        assert abs(g.q - r["q"]) <= tol
        np.testing.assert_allclose(g.qtraj, r["qtraj"], rtol=0, atol=tol)
        if g.assign is not None:
            np.testing.assert_array_equal(g.assign, r["assign"])
        if l >= 1 and g.graph is not None:
            go, gc, gw = ref.graphs[l]
            np.testing.assert_array_equal(g.graph.offsets, go)
            np.testing.assert_array_equal(g.graph.cols, gc)
            np.testing.assert_array_equal(g.graph_w64, gw)
    np.testing.assert_array_equal(gpu.final_assign, ref.final)


@pytest.mark.parametrize("version", [0, 1, 2])
def test_two_triangles_known_answer(ctx, mgl, version):
    # hand-derived known answer (two triangles joined by edge 2-3)
    r = ctx.louvain(two_triangles(mgl), version=version, keep_levels=True, keep_graphs=True)
    assert len(r.levels) == 2
    l0 = r.levels[0]
    # exact rationals; the FP64 expression Q = sum_in/2m - sum a^2/(2m)^2 may
    # differ from the correctly rounded fraction in the last place
    if version == 1:
        assert l0.qtraj[0] == pytest.approx(18 / 196, abs=1e-15)
    else:
        assert l0.qtraj[0] == pytest.approx(38 / 196, abs=1e-15)
    assert l0.qtraj[1] == pytest.approx(5 / 14, abs=1e-15)
    assert l0.q == pytest.approx(5 / 14, abs=1e-15)
    np.testing.assert_array_equal(l0.assign, [0, 0, 0, 4, 4, 4])
    g1 = r.levels[1].graph
    np.testing.assert_array_equal(g1.offsets, [0, 2, 4])
    np.testing.assert_array_equal(g1.cols, [0, 1, 0, 1])
    np.testing.assert_array_equal(r.levels[1].graph_w64, [6, 1, 1, 6])
    assert r.levels[1].q == pytest.approx(5 / 14, abs=1e-15)
    np.testing.assert_array_equal(r.final_assign, [0, 0, 0, 1, 1, 1])


@pytest.mark.parametrize("version", [0, 1, 2])
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("graph", [True, False])
def test_rmat16_full_parity(ctx, mgl, orc, version, fused, graph):
    g = mgl.gen_rmat(16, 16, seed=1)
    print(f"rmat16 v{version} fused={fused} graph={graph} n={g.n} nnz={g.nnz}", flush=True)
    res = ctx.louvain(g, version=version, fuse_modularity=fused, use_graph=graph,
                      keep_levels=True, keep_graphs=True)
    ref = orc.louvain(g.offsets, g.cols, None, mode=0, version=version)
    compare(res, ref)


def test_move_stage_random_assignments(ctx, mgl, orc):
    g = mgl.gen_rmat(14, 16, seed=7)
    s = ctx.session(g)
    rng = np.random.default_rng(0)
    k = orc.vertex_weights(g.offsets)
    for trial in range(4):
        # random coarse-grained assignment: community = id of a random member
        c = rng.integers(0, g.n, g.n).astype(np.int32)
        c = c[c]  # ids that are themselves used ids more often
        s.set_assign(c)
        a = orc.community_totals(k, c, g.n)
        ga, gk = s.weights()
        np.testing.assert_array_equal(gk, k)
        np.testing.assert_array_equal(ga, a)
        for up in (True, False):
            got = s.move(up)
            want = orc.move(g.offsets, g.cols, None, k, c, a, g.nnz, up)
            np.testing.assert_array_equal(got, want)
        assert s.modularity() == orc.modularity(g.offsets, g.cols, None, c, a, g.nnz)
    s.close()


def test_coarsen_stage(ctx, mgl, orc):
    g = mgl.gen_rmat(14, 16, seed=9)
    s = ctx.session(g)
    rng = np.random.default_rng(3)
    c = rng.integers(0, g.n // 8, g.n).astype(np.int32) * 8
    s.set_assign(c)
    off2, col2, w2, rank = s.coarsen()
    ro, rc, rw, rr = orc.coarsen(g.offsets, g.cols, None, c)
    np.testing.assert_array_equal(off2, ro)
    np.testing.assert_array_equal(col2, rc)
    np.testing.assert_array_equal(w2, rw.astype(np.float64))
    np.testing.assert_array_equal(rank, rr)
    s.close()


def test_hub_path_star(ctx, mgl, orc):
    # a star with 40000 leaves plus a ring: the centre's degree exceeds every
    # LDS-table bin, so its row goes through the bucketed hub path
    n = 40001
    edges = [(0, i) for i in range(1, n)] + [(i, i + 1) for i in range(1, n - 1)]
    g = mgl.CSR.from_edges(n, edges)
    s = ctx.session(g)
    assert s.bins()[-1] == 1
    s.close()
    res = ctx.louvain(g, keep_levels=True, keep_graphs=True)
    ref = orc.louvain(g.offsets, g.cols, None)
    compare(res, ref)


def test_hub_path_giant_star(ctx, mgl, orc):
    # a star with 4.4 M leaves plus a ring: the centre's row has more chunks
    # than the segmented hub form takes (so the scatter form runs), its 1,024
    # buckets hold about 4,300 distinct singletons each in the first pass,
    # beyond the owner table and the scanned LDS table, so the aggregate falls
    # back to its HBM table
    n = 4_400_001
    leaves = np.arange(1, n, dtype=np.int64)
    u = np.concatenate([np.zeros(n - 1, dtype=np.int64), leaves[:-1]])
    v = np.concatenate([leaves, leaves[1:]])
    g = mgl.CSR.from_edges(n, np.stack([u, v], axis=1))
    s = ctx.session(g)
    assert s.bins()[-1] == 1
    # the centre's proposal under singletons (all candidates distinct), random
    # communities and coarse groups (many duplicates per bucket)
    k = orc.vertex_weights(g.offsets)
    rng = np.random.default_rng(11)
    for c in (np.arange(n, dtype=np.int32), rng.integers(0, n, n).astype(np.int32),
              (np.arange(n, dtype=np.int32) // 1000) * 1000):
        s.set_assign(c)
        a = orc.community_totals(k, c, n)
        for up in (True, False):
            np.testing.assert_array_equal(
                s.move(up), orc.move(g.offsets, g.cols, None, k, c, a, g.nnz, up))
    s.close()
    res = ctx.louvain(g, keep_levels=True, keep_graphs=True)
    ref = orc.louvain(g.offsets, g.cols, None)
    compare(res, ref)


def hub_graph(mgl, weights=0):
    # R-MAT with a heavy tail: many rows in the hub bin
    return mgl.gen_rmat(17, 48, seed=13, weights=weights)


HUB_TUNE_BINS = "".join(
    f"{u} {sh} {s_} {sz}\n" for u, sh, s_, sz in
    [(0, 0, 0, 0), (4, 1, 4, 3), (8, 1, 8, 4), (16, 1, 16, 5), (32, 1, 32, 6),
     (64, 1, 64, 7), (128, 1, 64, 8), (256, 1, 64, 9), (1024, 2, 256, 2048),
     (-1, 3, 0, 0)])


class HubOracle:
    """The oracle's move pass for the hub graph in one of three arithmetic
    forms: unit weights, dyadic weights scaled to int64 (the exact mode), or
    the same weights as doubles (the FP64 mode; every sum of these weights is
    exact in FP64, so the two float orders agree bit for bit)."""

    def __init__(self, orc, g, form):
        self.orc, self.g, self.form = orc, g, form
        if form == "unit":
            self.w, self.mass = None, g.nnz
            self.k = orc.vertex_weights(g.offsets)
        elif form == "exact":
            S = orc.scale_of(g.w)
            self.w = orc.int_weights(g.w, S)
            self.mass = int(self.w.sum())
            self.k = orc.vertex_weights(g.offsets, self.w)
        else:
            self.w = g.w.astype(np.float64)
            rows = np.repeat(np.arange(g.n), np.diff(g.offsets))
            self.k = np.bincount(rows, weights=self.w, minlength=g.n)
            self.mass = float(self.w.sum())

    def move(self, c, up, version):
        if self.form == "fp64":
            a = np.zeros(self.g.n)
            np.add.at(a, c, self.k)
            return self.orc.move_f64(self.g.offsets, self.g.cols, self.w, self.k, c, a,
                                     self.mass, up, version)
        a = self.orc.community_totals(self.k, c, self.g.n)
        return self.orc.move(self.g.offsets, self.g.cols, self.w, self.k, c, a, self.mass, up,
                             version)


@pytest.mark.parametrize("version", [0, 1, 2])
@pytest.mark.parametrize("segmented", [1, 0])
@pytest.mark.parametrize("bucket", [4096, 1 << 20])
@pytest.mark.parametrize("form", ["unit", "exact", "fp64"])
def test_hub_path_random_assignments(ctx, mgl, orc, tmp_path, form, bucket, segmented, version):
    # bucket = 2^20 makes each hub a single bucket larger than the LDS table,
    # which exercises the HBM-table fallback of the aggregate step;
    # segmented = 0 forces the scatter form of the hub path; the weighted forms
    # take the 8-byte-sum tables (ModeInt, ModeF64)
    g = hub_graph(mgl, weights=0 if form == "unit" else 1)
    tf = tmp_path / "tune.txt"
    tf.write_text(f"hub_bucket_unclamped {bucket}\nhub_segmented {segmented}\n" + HUB_TUNE_BINS)
    prec = {"unit": mgl.PREC_EXACT, "exact": mgl.PREC_EXACT, "fp64": mgl.PREC_FP64}[form]
    s = ctx.session(g, tuning_file=str(tf), version=version, precision=prec)
    bins = s.bins()
    print(f"{form} v{version} bins {bins}", flush=True)
    assert bins[-1] > 10
    ho = HubOracle(orc, g, form)
    rng = np.random.default_rng(5)
    for trial in range(3):
        c = rng.integers(0, g.n, g.n).astype(np.int32)
        if trial > 0:
            c = (c // (7 * trial)) * (7 * trial)
        s.set_assign(c)
        for up in (True, False):
            np.testing.assert_array_equal(s.move(up), ho.move(c, up, version))
    s.close()


@pytest.mark.parametrize("threads", [512, 1024])
@pytest.mark.parametrize("form", ["unit", "exact"])
def test_hub_count_workgroup_sizes(ctx, mgl, orc, tmp_path, threads, form):
    # hub_count_threads: the count step's workgroups of 512 and 1024 threads
    # (8,192 entries each, so 16 and 8 entries per thread) give the default's
    # proposals; many hubs (rows above 1,024 entries)
    g = hub_graph(mgl, weights=0 if form == "unit" else 1)
    tf = tmp_path / "tune.txt"
    tf.write_text(f"hub_count_threads {threads}\n" + "".join(
        f"{u} {sh} {s_} {sz}\n" for u, sh, s_, sz in
        [(0, 0, 0, 0), (4, 1, 4, 3), (8, 1, 8, 4), (16, 1, 16, 5), (32, 1, 32, 6),
         (64, 1, 64, 7), (128, 1, 64, 8), (256, 1, 64, 9), (1024, 2, 256, 2048),
         (-1, 3, 0, 0)]))
    s = ctx.session(g, tuning_file=str(tf))
    ho = HubOracle(orc, g, form)
    rng = np.random.default_rng(31)
    for trial in range(2):
        c = rng.integers(0, g.n, g.n).astype(np.int32)
        if trial > 0:
            c = (c // 11) * 11
        s.set_assign(c)
        for up in (True, False):
            np.testing.assert_array_equal(s.move(up), ho.move(c, up, 2))
    s.close()
    res = ctx.louvain(g, tuning_file=str(tf), keep_levels=True)
    w = None if g.w is None else g.w.astype(np.float64)
    ref = orc.louvain(g.offsets, g.cols, w, mode=0)
    assert len(res.levels) == len(ref.levels)
    for lg_, lr in zip(res.levels, ref.levels):
        assert lg_.q == lr["q"] and lg_.iters == lr["iters"]
    np.testing.assert_array_equal(res.final_assign, ref.final)


@pytest.mark.parametrize("threads,per_cu", [(256, 2), (512, 4), (1024, 2), (1024, 1)])
def test_hub_aggregate_workgroup_shapes(ctx, mgl, orc, tmp_path, threads, per_cu):
    g = hub_graph(mgl)
    tf = tmp_path / "tune.txt"
    tf.write_text(f"hub_agg_threads {threads}\nhub_agg_per_cu {per_cu}\n" + "".join(
        f"{u} {sh} {s_} {sz}\n" for u, sh, s_, sz in
        [(0, 0, 0, 0), (4, 1, 4, 3), (8, 1, 8, 4), (16, 1, 16, 5), (32, 1, 32, 6),
         (64, 1, 64, 7), (128, 1, 64, 8), (256, 1, 64, 9), (1024, 2, 256, 2048),
         (-1, 3, 0, 0)]))
    s = ctx.session(g, tuning_file=str(tf))
    k = orc.vertex_weights(g.offsets)
    rng = np.random.default_rng(9)
    for trial in range(2):
        c = rng.integers(0, g.n, g.n).astype(np.int32)
        c = (c // (3 + 5 * trial)) * (3 + 5 * trial)
        s.set_assign(c)
        a = orc.community_totals(k, c, g.n)
        for up in (True, False):
            np.testing.assert_array_equal(
                s.move(up), orc.move(g.offsets, g.cols, None, k, c, a, g.nnz, up))
    s.close()


def test_tuning_file_applies_to_its_call_only(ctx, mgl, tmp_path):
    g = mgl.gen_rmat(12, 16, seed=3)
    s = ctx.session(g)
    before = s.bins()
    s.close()
    tf = tmp_path / "tune.txt"
    tf.write_text("hub_bucket 64\n0 0 0 0\n32 1 32 6\n-1 3 0 0\n")
    s = ctx.session(g, tuning_file=str(tf))
    assert len(s.bins()) == 3
    s.close()
    s = ctx.session(g)
    assert s.bins() == before
    s.close()


_GOOD_BINS = "0 0 0 0\n32 1 32 6\n-1 3 0 0\n"


@pytest.mark.parametrize("text, reason", [
    ("hub_agg_threads 300\n" + _GOOD_BINS, "hub_agg_threads"),
    ("hub_count_threads 128\n" + _GOOD_BINS, "hub_count_threads"),
    ("hub_agg_per_cu 9\n" + _GOOD_BINS, "hub_agg_per_cu"),
    ("hub_bucket 0\n" + _GOOD_BINS, "hub_bucket"),
    ("hub_fork_upper -1\n" + _GOOD_BINS, "hub_fork_upper"),
    ("hub_buckets 3072\n" + _GOOD_BINS, "unknown line"),
    ("hub_bucket 3072 4096\n" + _GOOD_BINS, "unknown line"),
    ("0 0 0 0\n32 1 32 6 7\n-1 3 0 0\n", "unknown line"),
    ("".join(f"{2 * i} 1 4 5\n" for i in range(1, 20)), "more bins"),
    ("0 0 0 0\n32 1 32 6\n16 1 16 5\n-1 3 0 0\n", "bounds"),
    ("0 0 0 0\n32 1 32 6\n", "hub path"),
    ("4 4 1 0\n-1 3 0 0\n", "degree 0"),
    ("0 0 0 0\n64 1 32 5\n-1 3 0 0\n", "table too small"),
    ("0 0 0 0\n16 1 0 5\n-1 3 0 0\n", "lanes per vertex"),
    ("0 0 0 0\n16 1 512 5\n-1 3 0 0\n", "lanes per vertex"),
    ("0 0 0 0\n16 1 4 6\n-1 3 0 0\n", "no compiled kernel"),
    ("0 0 0 0\n512 2 128 600\n-1 3 0 0\n", "threads"),
    ("0 0 0 0\n16384 5 512 1021\n-1 3 0 0\n", "shape"),
])
def test_tuning_file_errors(ctx, mgl, tmp_path, text, reason):
    # every malformed tuning file is refused with INVALID_ARG naming the
    # problem; the context stays usable
    g = mgl.gen_rmat(10, 8, seed=1)
    tf = tmp_path / "tune.txt"
    tf.write_text(text)
    with pytest.raises(mgl.MglError) as e:
        ctx.session(g, tuning_file=str(tf))
    assert e.value.code == 1 and reason in str(e.value), str(e.value)
    with pytest.raises(mgl.MglError) as e:
        ctx.louvain(g, tuning_file=str(tf))
    assert e.value.code == 1 and reason in str(e.value), str(e.value)
    with pytest.raises(mgl.MglError) as e:
        ctx.session(g, tuning_file=str(tmp_path / "missing.txt"))
    assert e.value.code == 1
    s = ctx.session(g)
    s.pass_()
    s.close()


@pytest.mark.parametrize("version", [0, 1, 2])
@pytest.mark.parametrize("form", ["unit", "exact", "fp64"])
def test_hub_graph_full_parity(ctx, mgl, orc, form, version):
    g = hub_graph(mgl, weights=0 if form == "unit" else 1)
    print(f"hub graph {form} v{version} n={g.n} nnz={g.nnz}", flush=True)
    if form == "unit":
        res = ctx.louvain(g, version=version, keep_levels=True, keep_graphs=True)
        compare(res, orc.louvain(g.offsets, g.cols, None, version=version))
    elif form == "exact":
        S = orc.scale_of(g.w)
        res = ctx.louvain(g, version=version, precision=mgl.PREC_EXACT, keep_levels=True,
                          keep_graphs=True)
        compare(res, orc.louvain(g.offsets, g.cols, g.w.astype(np.float64), mode=0,
                                 version=version), scale=S)
    else:
        # dyadic weights whose sums are exact in FP64: the FP64 mode is bit-equal
        # to the oracle's FP64 mode
        res = ctx.louvain(g, version=version, precision=mgl.PREC_FP64, keep_levels=True,
                          keep_graphs=True)
        assert res.precision == mgl.PREC_FP64
        compare_fp64(res, orc.louvain(g.offsets, g.cols, g.w.astype(np.float64), mode=1,
                                      version=version))


@pytest.mark.parametrize("tuning", ["hub_stream 0", "hub_fork_upper 0", "hub_fork_upper 128",
                                    "hub_fork_upper 2048", "hub_fork_upper 1000000"])
def test_hub_stream_placement_parity(ctx, mgl, orc, tmp_path, tuning):
    # the hub path on the main stream, or forked onto the second stream after
    # all bins, before the short-row bins, before rows of up to 2,048 entries,
    # or before every bin: the same results as the oracle in each case
    tf = tmp_path / "tune.txt"
    tf.write_text(tuning + "\n")
    g = hub_graph(mgl)
    res = ctx.louvain(g, tuning_file=str(tf), keep_levels=True, keep_graphs=True)
    compare(res, orc.louvain(g.offsets, g.cols, None))
    s = ctx.session(g, tuning_file=str(tf))
    k = orc.vertex_weights(g.offsets)
    c = (np.random.default_rng(5).integers(0, g.n, g.n) // 7 * 7).astype(np.int32)
    s.set_assign(c)
    a = orc.community_totals(k, c, g.n)
    for up in (True, False):
        np.testing.assert_array_equal(
            s.move(up), orc.move(g.offsets, g.cols, None, k, c, a, g.nnz, up))
    s.close()


def test_edge_cases(ctx, mgl, orc):
    cases = {
        "isolated_only": mgl.CSR(np.zeros(6, np.int64), np.zeros(0, np.int32)),
        "single_edge": mgl.CSR.from_edges(2, [(0, 1)]),
        "self_loops": mgl.CSR.from_edges(5, [(0, 0), (0, 1), (1, 2), (2, 2), (3, 4), (4, 4)]),
        "with_isolated": mgl.CSR.from_edges(7, [(0, 1), (1, 2), (2, 0), (4, 5)]),
    }
    for name, g in cases.items():
        print("edge case", name, flush=True)
        res = ctx.louvain(g, keep_levels=True, keep_graphs=True)
        ref = orc.louvain(g.offsets, g.cols, None)
        compare(res, ref)


def test_empty_graph(ctx, mgl):
    g = mgl.CSR(np.zeros(1, np.int64), np.zeros(0, np.int32))
    res = ctx.louvain(g)
    assert len(res.levels) == 1 and res.levels[0].n == 0 and res.final_q == 0.0


def test_sbm_weighted_exact(ctx, mgl, orc):
    g = mgl.gen_sbm(1 << 14, 64, min_size=64, seed=5, weights=1)
    S = orc.scale_of(g.w)
    assert S == 24
    res = ctx.louvain(g, precision=mgl.PREC_EXACT, keep_levels=True, keep_graphs=True)
    assert res.precision == mgl.PREC_EXACT
    ref = orc.louvain(g.offsets, g.cols, g.w.astype(np.float64), mode=0)
    compare(res, ref, scale=S)


def test_repeat_runs_identical(ctx, mgl):
    g = mgl.gen_rmat(15, 16, seed=21)
    a = ctx.louvain(g, keep_levels=True)
    b = ctx.louvain(g, keep_levels=True)
    assert a.final_q == b.final_q
    np.testing.assert_array_equal(a.final_assign, b.final_assign)


# ---- records in degree-bin order (forced on graphs below the automatic size) ----
@pytest.mark.parametrize("version", [0, 1, 2])
def test_binned_record_layout_parity(ctx, mgl, orc, version):
    g = mgl.gen_rmat(16, 16, seed=1)
    res = ctx.louvain(g, version=version, record_layout=mgl.LAYOUT_BINNED, keep_levels=True,
                      keep_graphs=True)
    ref = orc.louvain(g.offsets, g.cols, None, mode=0, version=version)
    compare(res, ref)


def test_binned_layout_hub_weighted_selfloops(ctx, mgl, orc):
    g = hub_graph(mgl)
    print(f"hub graph n={g.n} nnz={g.nnz}", flush=True)
    res = ctx.louvain(g, record_layout=mgl.LAYOUT_BINNED, keep_levels=True, keep_graphs=True)
    compare(res, orc.louvain(g.offsets, g.cols, None))
    g = mgl.gen_sbm(1 << 14, 64, min_size=64, seed=5, weights=1)
    S = orc.scale_of(g.w)
    res = ctx.louvain(g, precision=mgl.PREC_EXACT, record_layout=mgl.LAYOUT_BINNED,
                      keep_levels=True, keep_graphs=True)
    compare(res, orc.louvain(g.offsets, g.cols, g.w.astype(np.float64), mode=0), scale=S)
    # self-loops: the own-community weight excludes the loop, found by record position
    g = mgl.CSR.from_edges(5, [(0, 0), (0, 1), (1, 2), (2, 2), (3, 4), (4, 4)])
    res = ctx.louvain(g, record_layout=mgl.LAYOUT_BINNED, keep_levels=True, keep_graphs=True)
    compare(res, orc.louvain(g.offsets, g.cols, None))


def test_binned_layout_move_stage(ctx, mgl, orc):
    g = mgl.gen_rmat(14, 16, seed=7)
    s = ctx.session(g, record_layout=mgl.LAYOUT_BINNED)
    rng = np.random.default_rng(1)
    k = orc.vertex_weights(g.offsets)
    for trial in range(3):
        c = rng.integers(0, g.n, g.n).astype(np.int32)
        c = c[c]
        s.set_assign(c)
        a = orc.community_totals(k, c, g.n)
        for up in (True, False):
            np.testing.assert_array_equal(
                s.move(up), orc.move(g.offsets, g.cols, None, k, c, a, g.nnz, up))
        assert s.modularity() == orc.modularity(g.offsets, g.cols, None, c, a, g.nnz)
    s.close()


@pytest.mark.parametrize("version", [0, 1, 2])
@pytest.mark.parametrize("form", ["unit", "exact", "fp64"])
def test_binned_layout_weighted_move_stage(ctx, mgl, orc, form, version):
    # records and rows (with their int64 or FP64 weights) placed in degree-bin
    # order on a graph below the automatic size: every bin shape, the hub path
    # included, reads the placed rows; random assignments against the oracle
    g = hub_graph(mgl, weights=0 if form == "unit" else 1)
    prec = {"unit": mgl.PREC_EXACT, "exact": mgl.PREC_EXACT, "fp64": mgl.PREC_FP64}[form]
    s = ctx.session(g, record_layout=mgl.LAYOUT_BINNED, version=version, precision=prec)
    bins = s.bins()
    assert bins[-1] > 0, bins  # hub rows present
    ho = HubOracle(orc, g, form)
    rng = np.random.default_rng(41)
    for trial in range(3):
        c = rng.integers(0, g.n, g.n).astype(np.int32)
        if trial > 0:
            c = (c // (5 * trial)) * (5 * trial)
        s.set_assign(c)
        for up in (True, False):
            np.testing.assert_array_equal(s.move(up), ho.move(c, up, version))
    s.close()


def test_autotune_layout_parity(ctx, mgl, orc, tmp_path):
    # the autotuner times candidate bin layouts and hub-aggregate shapes on the
    # graph and writes the fastest as a tuning file; runs under that file match
    # the oracle
    g = hub_graph(mgl)
    tf = str(tmp_path / "tuned.txt")
    ctx.autotune(g, tf)
    text = open(tf).read()
    print(text, flush=True)
    assert "hub_agg_threads" in text and "hub_bucket" in text
    assert text.rstrip().endswith("-1 3 0 0")
    s = ctx.session(g, tuning_file=tf)
    k = orc.vertex_weights(g.offsets)
    rng = np.random.default_rng(17)
    c = (rng.integers(0, g.n, g.n) // 3 * 3).astype(np.int32)
    s.set_assign(c)
    a = orc.community_totals(k, c, g.n)
    for up in (True, False):
        np.testing.assert_array_equal(s.move(up),
                                      orc.move(g.offsets, g.cols, None, k, c, a, g.nnz, up))
    s.close()
    compare(ctx.louvain(g, tuning_file=tf, keep_levels=True, keep_graphs=True),
            orc.louvain(g.offsets, g.cols, None))
