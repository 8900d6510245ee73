"""GPU checks of the precision modes and of the full-size workloads.

* FP64 mode against the oracle's FP64 mode (dyadic weights: identical sums)
* FP32-hash mode against an all-FP64 GPU run (lognormal weights): modularity
  difference and the fraction of vertices whose final community agrees
* final modularity of every mode within 1e-6 relative of an independent FP64
  recomputation on the CPU
* planted partition 2^20 vertices / 1,024 communities / weights in (0.5, 1.5]:
  bit-exact against the oracle in the exact mode
* R-MAT scale 24 edge factor 32: size-independent properties of the full run
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def agreement(a, b):
    """Fraction of vertices whose community agrees after matching each
    community of a with its most frequent counterpart in b."""
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    key = a * (int(b.max()) + 1) + b
    uk, cnt = np.unique(key, return_counts=True)
    ua = uk // (int(b.max()) + 1)
    best = {}
    for x, c in zip(ua.tolist(), cnt.tolist()):
        best[x] = max(best.get(x, 0), c)
    return sum(best.values()) / len(a)


@pytest.mark.parametrize("layout", ["AUTO", "BINNED"])
def test_fp64_mode_matches_oracle(ctx, mgl, orc, layout):
    # BINNED: records, rows and their FP64 weights placed in degree-bin order
    g = mgl.gen_sbm(1 << 13, 32, min_size=64, seed=21, weights=1)
    res = ctx.louvain(g, precision=mgl.PREC_FP64, keep_levels=True,
                      record_layout=getattr(mgl, "LAYOUT_" + layout))
    assert res.precision == mgl.PREC_FP64
    ref = orc.louvain(g.offsets, g.cols, g.w.astype(np.float64), mode=1)
    assert len(res.levels) == len(ref.levels)
    for lg, lr in zip(res.levels, ref.levels):
        np.testing.assert_array_equal(lg.assign, lr["assign"])
        assert lg.q == pytest.approx(lr["q"], abs=1e-12)
    np.testing.assert_array_equal(res.final_assign, ref.final)


@pytest.mark.parametrize("prec", ["EXACT", "FP64", "FP32_HASH"])
def test_final_q_matches_fp64_recompute(ctx, mgl, orc, prec):
    g = mgl.gen_rmat(18, 16, seed=5, weights=1)
    p = getattr(mgl, "PREC_" + prec)
    res = ctx.louvain(g, precision=p)
    q64 = orc.modularity_fp64(g.offsets, g.cols, g.w, res.final_assign)
    print(f"{prec}: Q={res.final_q:.12f} fp64 recompute={q64:.12f}", flush=True)
    assert abs(res.final_q - q64) <= 1e-6 * abs(q64)


@pytest.mark.parametrize("layout", ["AUTO", "BINNED"])
def test_fp32_hash_vs_fp64_lognormal(ctx, mgl, orc, layout):
    # lognormal(0,1) weights: neither mode is exact; report the difference.
    # BINNED: the FP32-hash run reads its rows and FP32 weights placed in
    # degree-bin order
    g = mgl.gen_rmat(18, 16, seed=5, weights=2)
    r32 = ctx.louvain(g, precision=mgl.PREC_FP32_HASH,
                      record_layout=getattr(mgl, "LAYOUT_" + layout))
    r64 = ctx.louvain(g, precision=mgl.PREC_FP64)
    q32 = orc.modularity_fp64(g.offsets, g.cols, g.w, r32.final_assign)
    q64 = orc.modularity_fp64(g.offsets, g.cols, g.w, r64.final_assign)
    agr = agreement(r32.final_assign, r64.final_assign)
    print(f"fp32 Q={r32.final_q:.9f} (recomputed {q32:.9f}) fp64 Q={r64.final_q:.9f} "
          f"(recomputed {q64:.9f}) |dQ|={abs(q32 - q64):.3e} agreement={agr:.4f}", flush=True)
    assert abs(r32.final_q - q32) <= 1e-6 * abs(q32)
    assert abs(r64.final_q - q64) <= 1e-6 * abs(q64)
    # observed on the MI355X at this size: |dQ| = 0 and agreement 1.0 (FP32 hash
    # sums of lognormal weights are order dependent, so a small margin)
    assert abs(q32 - q64) <= 1e-6
    assert agr >= 0.995


@pytest.mark.timeout(600)
def test_sbm_2p20_exact_bit_parity(ctx, mgl, orc):
    g = mgl.gen_sbm(1 << 20, 1024, exponent=2.1, min_size=64, deg_in=24, deg_out=4, seed=2,
                    weights=1)
    print(f"SBM n={g.n} nnz={g.nnz}", flush=True)
    res = ctx.louvain(g, precision=mgl.PREC_EXACT, keep_levels=True)
    ref = orc.louvain(g.offsets, g.cols, g.w.astype(np.float64), mode=0)
    assert len(res.levels) == len(ref.levels)
    for lg, lr in zip(res.levels, ref.levels):
        assert lg.n == lr["n"] and lg.iters == lr["iters"] and lg.q == lr["q"]
        np.testing.assert_array_equal(lg.qtraj, lr["qtraj"])
        np.testing.assert_array_equal(lg.assign, lr["assign"])
    np.testing.assert_array_equal(res.final_assign, ref.final)
    q64 = orc.modularity_fp64(g.offsets, g.cols, g.w, res.final_assign)
    print(f"levels={len(res.levels)} Q={res.final_q:.9f} fp64={q64:.9f}", flush=True)
    assert abs(res.final_q - q64) <= 1e-6 * abs(q64)


@pytest.mark.timeout(900)
def test_rmat24_full_run_properties(ctx, mgl, orc):
    g = mgl.gen_rmat(24, 32, seed=3)
    print(f"R-MAT 24: n={g.n} nnz={g.nnz}", flush=True)
    a = ctx.louvain(g, keep_levels=True)
    b = ctx.louvain(g)
    # deterministic
    assert a.final_q == b.final_q
    np.testing.assert_array_equal(a.final_assign, b.final_assign)
    # committed modularity never decreases; level sizes shrink
    for l in a.levels:
        assert np.all(np.diff(l.qtraj) >= 0)
    for l0, l1 in zip(a.levels, a.levels[1:]):
        assert l1.n < l0.n
        assert l1.q >= l0.q
    # the final labels are a dense relabelling of the last level's assignment
    assert a.final_assign.min() == 0
    q64 = orc.modularity_fp64(g.offsets, g.cols, None, a.final_assign)
    print(f"levels={len(a.levels)} iters={[l.iters for l in a.levels]} Q={a.final_q:.12f} "
          f"fp64={q64:.12f} seconds={a.seconds:.3f}", flush=True)
    assert abs(a.final_q - q64) <= 1e-6 * abs(q64)
    # community totals sum to 2m after a pass
    s = ctx.session(g)
    s.reset()
    s.pass_()
    at, kt = s.weights()
    assert at.sum() == g.nnz and kt.sum() == g.nnz
    s.close()


def test_session_count_matches_numpy_and_leaves_passes_unchanged(ctx, mgl):
    g = mgl.gen_rmat(13, 16, seed=17)
    # distinct (vertex, neighbour community != own) pairs of a random assignment
    rng = np.random.default_rng(4)
    c = rng.integers(0, g.n // 16, g.n).astype(np.int32) * 16
    s = ctx.session(g, fuse_modularity=False)
    s.set_assign(c)
    s.count()
    s.pass_()
    st = s.stats()
    rows = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.offsets))
    cu = c[g.cols].astype(np.int64)
    keep = cu != c[rows]
    want = len(np.unique(rows[keep] * g.n + cu[keep]))
    assert st["distinct"] == want
    assert 0 < st["lds_inserts"] <= st["lds_probes"]
    assert 0 <= st["lds_multi_probe"] <= st["lds_inserts"]
    s.close()
    # counting before each pass does not change what the passes compute
    a, b = ctx.session(g), ctx.session(g)
    a.reset()
    b.reset()
    for _ in range(6):
        b.count()
        qa, acc_a = a.pass_()
        qb, acc_b = b.pass_()
        assert qa == qb and acc_a == acc_b
        assert b.stats()["distinct"] > 0 and a.stats()["distinct"] == 0
    a.close()
    b.close()
