"""GPU parity of the move rule's corners and of the control options, through
libmgl.so's C ABI against the CPU oracle:

* the hand-derived known answer with a down-pass tie and a zero-gain move
  (tests/golden/kat_path_ties.json), under the reference tie rule and the
  smallest-id rule, versions 0/1/2, fused and unfused modularity
* the smallest-id tie rule on R-MAT 16 (v0/v1/v2) against the oracle's
* honoured max_iter and max_phases cut-offs (main.cu:20-26, louvain.cu:1473)
* malformed CSRs refused with MGL_INVALID_GRAPH before any kernel runs
* the scan variants of the move kernels (k_move_group, k_move_block,
  k_move_lane<..., 16>) forced by a tuning file
* an LDS table smaller than its bin's rows reported as MGL_DEVICE_ERROR
* FP64 weights from a general .mtx (0.1, 0.3, 1/3, not float-representable)
  in FP64 mode against the oracle's FP64 mode
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def compare(gpu, ref, scale=0, check_graphs=True):
    assert len(gpu.levels) == len(ref.levels), (len(gpu.levels), len(ref.levels))
    for l, (g, r) in enumerate(zip(gpu.levels, ref.levels)):
        assert g.n == r["n"] and g.iters == r["iters"], (l, g.n, r["n"], g.iters, r["iters"])
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


@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("version", [0, 1, 2])
@pytest.mark.parametrize("rule", ["reference", "smallest"])
def test_known_answer_path_ties(ctx, mgl, rule, version, fused):
    with open(os.path.join(GOLDEN, "kat_path_ties.json")) as f:
        kat = json.load(f)
    R = kat["rules"][rule]
    g = mgl.CSR.from_edges(kat["n"], kat["edges"])
    res = ctx.louvain(g, version=version, tie_rule=R["tie"], fuse_modularity=fused,
                      keep_levels=True, keep_graphs=True)
    assert len(res.levels) == len(R["levels"])
    for li, (lv, le) in enumerate(zip(res.levels, R["levels"])):
        assert lv.n == le["n"] and lv.iters == le["iters"]
        # 2m = 8: every value is a dyadic rational, exact in FP64
        assert lv.q0 == float(Fraction(le["q0"]))
        assert list(lv.qtraj) == [float(Fraction(x)) for x in le["qtraj"]]
        np.testing.assert_array_equal(lv.assign, le["assign"])
        if li >= 1:
            np.testing.assert_array_equal(lv.graph.offsets, le["graph"]["offsets"])
            np.testing.assert_array_equal(lv.graph.cols, le["graph"]["cols"])
            np.testing.assert_array_equal(lv.graph_w64, le["graph"]["w"])
    np.testing.assert_array_equal(res.final_assign, R["final"])
    assert res.final_q == float(Fraction(R["final_q"]))
    # pass by pass through the move stage
    s = ctx.session(g, version=version, tie_rule=R["tie"])
    for ps in R["levels"][0]["passes"]:
        s.set_assign(np.asarray(ps["c"], np.int32))
        np.testing.assert_array_equal(s.move(ps["up"]), ps["proposal"])
    s.close()


@pytest.mark.parametrize("version", [0, 1, 2])
def test_tie_smallest_rmat16(ctx, mgl, orc, version):
    g = mgl.gen_rmat(16, 16, seed=1)
    res = ctx.louvain(g, version=version, tie_rule=mgl.TIE_SMALLEST, keep_levels=True,
                      keep_graphs=True)
    ref = orc.louvain(g.offsets, g.cols, None, version=version, tie=1)
    compare(res, ref)
    # the rule changes the answer on this graph
    base = orc.louvain(g.offsets, g.cols, None, version=version, tie=0)
    assert not np.array_equal(ref.final, base.final) or \
        [l["iters"] for l in ref.levels] != [l["iters"] for l in base.levels]


def test_tie_smallest_hub_and_weighted(ctx, mgl, orc):
    # hub rows and the 8-byte (weighted) tables under the smallest-id rule
    g = mgl.gen_rmat(17, 48, seed=13)
    compare(ctx.louvain(g, tie_rule=mgl.TIE_SMALLEST, keep_levels=True, keep_graphs=True),
            orc.louvain(g.offsets, g.cols, None, tie=1))
    g = mgl.gen_sbm(1 << 14, 64, min_size=64, seed=5, weights=1)
    S = orc.scale_of(g.w)
    compare(ctx.louvain(g, tie_rule=mgl.TIE_SMALLEST, precision=mgl.PREC_EXACT, keep_levels=True,
                        keep_graphs=True),
            orc.louvain(g.offsets, g.cols, g.w.astype(np.float64), mode=0, tie=1), scale=S)


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("max_iter", [1, 2, 5])
def test_max_iter_honoured(ctx, mgl, orc, max_iter, graph):
    g = mgl.gen_rmat(14, 16, seed=3)
    res = ctx.louvain(g, max_iter=max_iter, use_graph=graph, keep_levels=True, keep_graphs=True)
    ref = orc.louvain(g.offsets, g.cols, None, max_iter=max_iter)
    assert all(l["iters"] <= max_iter for l in ref.levels)
    assert any(l["iters"] == max_iter for l in ref.levels)
    compare(res, ref)


@pytest.mark.parametrize("max_phases", [1, 2])
def test_max_phases_honoured(ctx, mgl, orc, max_phases):
    g = mgl.gen_rmat(14, 16, seed=3)
    res = ctx.louvain(g, max_phases=max_phases, keep_levels=True, keep_graphs=True)
    ref = orc.louvain(g.offsets, g.cols, None, max_phases=max_phases)
    assert len(ref.levels) == max_phases
    compare(res, ref)
    full = orc.louvain(g.offsets, g.cols, None)
    assert len(full.levels) > max_phases


def test_max_phases_zero_runs_nothing(ctx, mgl, orc):
    g = mgl.gen_rmat(10, 8, seed=3)
    res = ctx.louvain(g, max_phases=0)
    ref = orc.louvain(g.offsets, g.cols, None, max_phases=0)
    assert len(res.levels) == len(ref.levels) == 0
    # no level ran: every vertex is its own community
    np.testing.assert_array_equal(ref.final, np.arange(g.n, dtype=np.int32))
    np.testing.assert_array_equal(res.final_assign, ref.final)


def _bad_graphs(mgl):
    good = mgl.CSR.from_edges(5, [(0, 1), (1, 2), (2, 3), (3, 4)], [1.0, 2.0, 1.0, 2.0])
    off, col, w = good.offsets.copy(), good.cols.copy(), good.w.copy()
    cases = {}
    o = off.copy()
    o[2], o[3] = o[3], o[2]  # non-monotone offsets (the total is unchanged)
    cases["offsets are not non-decreasing"] = (mgl.CSR(o, col.copy(), w.copy()), 1)
    c = col.copy()
    c[3] = 5
    cases["column index out of range"] = (mgl.CSR(off.copy(), c, w.copy()), 1)
    c = col.copy()
    c[0] = -1
    cases["column index out of range "] = (mgl.CSR(off.copy(), c, w.copy()), 1)
    ww = w.copy()
    ww[2] = -1.0
    cases["edge weight negative or not finite"] = (mgl.CSR(off.copy(), col.copy(), ww), 1)
    ww = w.astype(np.float64)
    ww[1] = np.nan
    cases["edge weight negative or not finite "] = (mgl.CSR(off.copy(), col.copy(), ww), 1)
    ww = w.copy()
    ww[0] = 3.0  # (0,1) weight 3 but (1,0) weight 1
    cases["graph is not symmetric"] = (mgl.CSR(off.copy(), col.copy(), ww), 2)
    asym = mgl.CSR(np.array([0, 1, 1, 1, 1, 1], np.int64), np.array([1], np.int32))
    cases["graph is not symmetric "] = (asym, 2)
    o = off.copy()
    o[0] = 1
    o[1] = max(o[1], 1)
    cases["offsets[0] must be 0"] = (mgl.CSR(o, col.copy(), w.copy()), 1)
    return cases


def test_malformed_csr_rejected(ctx, mgl):
    # host-side validation: each case fails with INVALID_GRAPH naming the
    # problem (asymmetry only with validate=2), before anything runs; the
    # context stays usable
    for reason, (g, level) in _bad_graphs(mgl).items():
        with pytest.raises(mgl.MglError) as e:
            ctx.louvain(g, validate=level)
        assert e.value.code == mgl.INVALID_GRAPH, (reason, str(e.value))
        assert reason.strip() in str(e.value), (reason, str(e.value))
        with pytest.raises(mgl.MglError) as e:
            ctx.louvain_batched([mgl.gen_rmat(8, 4, seed=1), g], validate=level)
        assert e.value.code == mgl.INVALID_GRAPH
        with pytest.raises(mgl.MglError) as e:
            ctx.session(g, validate=level)
        assert e.value.code == mgl.INVALID_GRAPH
    # an asymmetric graph passes the default validation level (as the reference,
    # which does not check) and runs
    asym = _bad_graphs(mgl)["graph is not symmetric "][0]
    ctx.louvain(asym)
    ctx.louvain(mgl.gen_rmat(10, 8, seed=1))


# bins whose (lanes, table) pairs have no owner-based instance, so the scan
# kernels run: k_move_lane<16>, k_move_group<16, 5>, k_move_group<64, 10>,
# k_move_block<256>
SCAN_LAYOUT = "0 0 0 0\n16 4 1 0\n32 1 16 5\n1024 1 64 10\n8192 2 256 8192\n-1 3 0 0\n"


@pytest.mark.parametrize("weighted", [False, True])
def test_scan_variants_forced_by_tuning_file(ctx, mgl, orc, tmp_path, weighted):
    tf = tmp_path / "scan.txt"
    tf.write_text(SCAN_LAYOUT)
    g = mgl.gen_rmat(16, 16, seed=1, weights=1 if weighted else 0)
    prec = mgl.PREC_EXACT
    s = ctx.session(g, tuning_file=str(tf), precision=prec)
    bins = s.bins()
    print("bins", bins, flush=True)
    assert len(bins) == 6 and all(b > 0 for b in bins[1:5])
    if weighted:
        S = orc.scale_of(g.w)
        wi = orc.int_weights(g.w, S)
        mass = int(wi.sum())
    else:
        S, wi, mass = 0, None, g.nnz
    k = orc.vertex_weights(g.offsets, wi)
    rng = np.random.default_rng(23)
    for trial in range(3):
        c = rng.integers(0, g.n, g.n).astype(np.int32)
        if trial:
            c = (c // (5 * trial)) * (5 * trial)
        s.set_assign(c)
        a = orc.community_totals(k, c, g.n)
        for version in (2,):
            for up in (True, False):
                np.testing.assert_array_equal(
                    s.move(up), orc.move(g.offsets, g.cols, wi, k, c, a, mass, up, version))
    s.close()
    for version in (0, 1, 2):
        res = ctx.louvain(g, tuning_file=str(tf), precision=prec, version=version,
                          keep_levels=True, keep_graphs=True)
        ref = orc.louvain(g.offsets, g.cols, None if not weighted else g.w.astype(np.float64),
                          mode=0, version=version)
        compare(res, ref, scale=S)


def test_lds_table_overflow_is_a_device_error(ctx, mgl, tmp_path):
    # "table_check 0" lets a 16-slot table serve rows of up to 64 entries: the
    # first pass from singletons overflows it, which the call reports as
    # DEVICE_ERROR instead of a wrong answer (the insert loops are bounded by
    # the table size, so nothing reads or writes outside the table)
    tf = tmp_path / "small.txt"
    tf.write_text("table_check 0\n0 0 0 0\n64 1 8 4\n-1 3 0 0\n")
    g = mgl.gen_rmat(12, 16, seed=2)
    with pytest.raises(mgl.MglError) as e:
        ctx.louvain(g, tuning_file=str(tf))
    assert e.value.code == mgl.DEVICE_ERROR and "overflow" in str(e.value)
    s = ctx.session(g, tuning_file=str(tf))
    with pytest.raises(mgl.MglError) as e:
        s.move(True)
    assert e.value.code == mgl.DEVICE_ERROR
    s.close()
    # a timed pass reads the error word with the graph states
    s = ctx.session(g, tuning_file=str(tf))
    with pytest.raises(mgl.MglError) as e:
        s.pass_()
    assert e.value.code == mgl.DEVICE_ERROR and "overflow" in str(e.value)
    s.close()
    # without the switch the same layout is refused up front
    tf.write_text("0 0 0 0\n64 1 8 4\n-1 3 0 0\n")
    with pytest.raises(mgl.MglError) as e:
        ctx.louvain(g, tuning_file=str(tf))
    assert e.value.code == mgl.INVALID_ARG
    ctx.louvain(g)


def _write_general_mtx(path, n, edges, weights):
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{n} {n} {len(edges)}\n")
        for (u, v), w in zip(edges, weights):
            f.write(f"{u + 1} {v + 1} {w}\n")


@pytest.mark.parametrize("version", [0, 1, 2])
def test_fp64_weights_from_mtx(ctx, mgl, orc, tmp_path, version):
    # weights 0.1, 0.3 and 1/3 (and random reals scaled from them) are not
    # floats: the loader keeps the parsed doubles (the reference's weight_t)
    # and FP64 mode runs on them. Most weights are random reals so that no two
    # candidates tie exactly (a tie between sums of equal weights added in
    # different orders could round either way in either implementation).
    rng = np.random.default_rng(31)
    n = 3000
    pairs = rng.integers(0, n, (14000, 2))
    edges = sorted({(int(max(u, v)), int(min(u, v))) for u, v in pairs if u != v})
    base = [0.1, 0.3, 1.0 / 3.0]
    w = []
    for i in range(len(edges)):
        b = base[i % 3]
        w.append(repr(b) if i % 10 == 0 else repr(b * (1.0 + float(rng.random()))))
    p = tmp_path / "w64.mtx"
    _write_general_mtx(p, n, edges, w)
    g = mgl.read_mtx(str(p))
    assert g.w.dtype == np.float64
    assert {0.1, 0.3, 1.0 / 3.0} <= set(np.unique(g.w).tolist())
    # the doubles arrive unchanged; a float would have rounded them
    want = sorted(float(x) for x in w for _ in range(2))
    np.testing.assert_array_equal(np.sort(g.w), want)
    assert not np.array_equal(g.w.astype(np.float32).astype(np.float64), g.w)
    assert orc.scale_of(g.w) < 0  # not dyadic with a usable scale
    res = ctx.louvain(g, version=version, keep_levels=True)
    assert res.precision == mgl.PREC_FP64  # AUTO falls back to FP64
    ref = orc.louvain(g.offsets, g.cols, g.w, mode=1, version=version)
    assert len(res.levels) == len(ref.levels)
    for lg, lr in zip(res.levels, ref.levels):
        assert lg.n == lr["n"] and lg.iters == lr["iters"]
        np.testing.assert_array_equal(lg.assign, lr["assign"])
        assert abs(lg.q - lr["q"]) <= 1e-12
    np.testing.assert_array_equal(res.final_assign, ref.final)
    assert abs(res.final_q - ref.final_q) <= 1e-12
    # the literal values only (ties are common): the final modularity is the
    # FP64 recomputation of the partition found
    lit = [base[i % 3] for i in range(len(edges))]
    g2 = mgl.CSR.from_edges(n, edges, lit, w64=True)
    assert g2.w.dtype == np.float64
    r2 = ctx.louvain(g2, version=version, precision=mgl.PREC_FP64)
    q2 = orc.modularity_fp64(g2.offsets, g2.cols, g2.w, r2.final_assign)
    assert abs(r2.final_q - q2) <= 1e-12
