"""Pins the CPU oracle (oracle/louvain_oracle.cpp) with checks that do not
retype its formulas:

  * move rule     : exact-rational dQ by full modularity recomputation on tiny
                    random graphs (fractions), for v0/v1/v2 and both directions
  * known answer  : two triangles joined by an edge (hand-derived per pass)
  * modularity    : closed forms (one community, singletons, ring of cliques)
                    and networkx.community.modularity (fixture + live)
  * coarsening    : dict-based brute force, weight conservation, symmetry,
                    dense-rank renumbering, Q invariance across levels
  * u128 -> double: Python's correctly rounded int -> float
The reference ships no golden vectors or tests for this path, so these are
the pins.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def csr_from_edges(n, edges, w=None):
    from conftest import load_mgl
    return load_mgl().CSR.from_edges(n, edges, w)


def q_exact(n, edges, c):
    """Modularity with Fractions from the undirected edge list (unit weights,
    a self-loop counts once in the CSR, i.e. weight 1 in sum_in and k)."""
    # build directed entries like the CSR does
    ent = []
    for u, v in edges:
        ent.append((u, v))
        if u != v:
            ent.append((v, u))
    m2 = Fraction(len(ent))
    k = [0] * n
    for u, v in ent:
        k[u] += 1
    sin = sum(1 for u, v in ent if c[u] == c[v])
    tot = {}
    for v in range(n):
        tot[c[v]] = tot.get(c[v], 0) + k[v]
    return Fraction(sin) / m2 - sum(Fraction(t * t) for t in tot.values()) / (m2 * m2)


def random_graph(rng, n, p):
    edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
    return edges


@pytest.mark.parametrize("version", [0, 1, 2])
def test_move_rule_brute_force(orc, version):
    rng = np.random.default_rng(1234 + version)
    checked = 0
    for trial in range(60):
        n = int(rng.integers(4, 10))
        edges = random_graph(rng, n, 0.45)
        if not edges:
            continue
        g = csr_from_edges(n, edges)
        c = rng.integers(0, n, n).astype(np.int32)
        k = orc.vertex_weights(g.offsets)
        a = orc.community_totals(k, c, n)
        for up in (True, False):
            got = orc.move(g.offsets, g.cols, None, k, c, a, g.nnz, up, version)
            q0 = q_exact(n, edges, c)
            for v in range(n):
                ci = int(c[v])
                nbrs = g.cols[g.offsets[v]:g.offsets[v + 1]]
                cands = sorted({int(c[u]) for u in nbrs if c[u] != ci})
                if version == 1:
                    cands = [x for x in cands if (x < ci if up else x > ci)]
                gains = {}
                for x in cands:
                    c2 = c.copy()
                    c2[v] = x
                    gains[x] = q_exact(n, edges, c2) - q0
                vals = sorted(gains.values())
                # skip rows where FP64 rounding could reorder equal or zero gains
                amb = any(abs(float(gv)) < 1e-12 for gv in vals) or any(
                    abs(float(vals[i + 1] - vals[i])) < 1e-12 for i in range(len(vals) - 1))
                if amb:
                    continue
                dst = ci
                if gains:
                    x = max(gains, key=lambda y: gains[y])
                    if gains[x] > 0:
                        dst = x
                if version != 1:
                    dst = min(dst, ci) if up else max(dst, ci)
                assert got[v] == dst, (trial, v, up, gains, got[v])
                checked += 1
    assert checked > 200


def test_known_answer_two_triangles(orc):
    with open(os.path.join(GOLDEN, "kat_two_triangles.json")) as f:
        kat = json.load(f)
    g = csr_from_edges(6, kat["edges"])
    for ver, exp in kat["versions"].items():
        R = orc.louvain(g.offsets, g.cols, None, version=int(ver))
        assert len(R.levels) == len(exp["levels"])
        for lv, le in zip(R.levels, exp["levels"]):
            assert lv["n"] == le["n"] and lv["iters"] == le["iters"]
            for q, qe in zip(lv["qtraj"], le["qtraj"]):
                assert abs(q - Fraction(qe)) < 1e-15
            np.testing.assert_array_equal(lv["assign"], le["assign"])
        go, gc, gw = R.graphs[1]
        np.testing.assert_array_equal(go, exp["coarse"]["offsets"])
        np.testing.assert_array_equal(gc, exp["coarse"]["cols"])
        np.testing.assert_array_equal(gw, exp["coarse"]["w"])
        np.testing.assert_array_equal(R.final, exp["final"])


def test_modularity_closed_forms(orc):
    rng = np.random.default_rng(7)
    n = 40
    edges = random_graph(rng, n, 0.2)
    g = csr_from_edges(n, edges)
    k = orc.vertex_weights(g.offsets)
    m2 = g.nnz
    # one community: Q = 0 exactly
    c = np.zeros(n, np.int32)
    assert orc.modularity(g.offsets, g.cols, None, c, orc.community_totals(k, c, n), m2) == 0.0
    # singletons without self-loops: Q = -sum k^2 / (2m)^2
    c = np.arange(n, dtype=np.int32)
    q = orc.modularity(g.offsets, g.cols, None, c, k, m2)
    assert q == pytest.approx(-float(np.sum(k.astype(float) ** 2)) / m2 ** 2, abs=1e-15)
    # ring of r cliques K_s joined by single edges, clique partition:
    # Q = r*l/m' - r*(d/2m')^2, l = s(s-1)/2, d = s(s-1)+2, m' = r(l+1)
    for r, s_ in [(4, 5), (10, 3), (6, 8)]:
        e = []
        for i in range(r):
            base = i * s_
            e += [(base + a_, base + b_) for a_ in range(s_) for b_ in range(a_ + 1, s_)]
            e.append((base + s_ - 1, ((i + 1) % r) * s_))
        gg = csr_from_edges(r * s_, e)
        kk = orc.vertex_weights(gg.offsets)
        cc = (np.arange(r * s_) // s_ * s_).astype(np.int32)
        qq = orc.modularity(gg.offsets, gg.cols, None, cc, orc.community_totals(kk, cc, r * s_),
                            gg.nnz)
        l_, d_, mp = s_ * (s_ - 1) // 2, s_ * (s_ - 1) + 2, r * (s_ * (s_ - 1) // 2 + 1)
        assert qq == pytest.approx(r * l_ / mp - r * (d_ / (2 * mp)) ** 2, abs=1e-14)


def test_modularity_networkx_fixture(orc):
    with open(os.path.join(GOLDEN, "nx_modularity.json")) as f:
        fx = json.load(f)
    for case in fx["cases"]:
        n = case["n"]
        w = case.get("w")
        g = csr_from_edges(n, case["edges"], w)
        c = np.asarray(case["assign"], np.int32)
        if w is None:
            k = orc.vertex_weights(g.offsets)
            a = orc.community_totals(k, c, n)
            q = orc.modularity(g.offsets, g.cols, None, c, a, g.nnz)
        else:
            S = orc.scale_of(g.w)
            wi = orc.int_weights(g.w, S)
            k = orc.vertex_weights(g.offsets, wi)
            a = orc.community_totals(k, c, n)
            q = orc.modularity(g.offsets, g.cols, wi, c, a, int(wi.sum()))
        assert q == pytest.approx(case["q_networkx"], abs=1e-12), case["name"]


def test_modularity_networkx_live(orc):
    nx = pytest.importorskip("networkx")
    rng = np.random.default_rng(11)
    for t in range(5):
        n = 60
        edges = random_graph(rng, n, 0.1)
        G = nx.Graph()
        G.add_nodes_from(range(n))
        G.add_edges_from(edges)
        c = rng.integers(0, 6, n).astype(np.int32)
        comms = [set(np.flatnonzero(c == x).tolist()) for x in np.unique(c)]
        qn = nx.community.modularity(G, comms)
        g = csr_from_edges(n, edges)
        k = orc.vertex_weights(g.offsets)
        q = orc.modularity(g.offsets, g.cols, None, c, orc.community_totals(k, c, n), g.nnz)
        assert q == pytest.approx(qn, abs=1e-12)


def test_coarsen_brute_force(orc):
    rng = np.random.default_rng(3)
    for t in range(20):
        n = int(rng.integers(5, 40))
        edges = random_graph(rng, n, 0.25) + [(i, i) for i in range(n) if rng.random() < 0.1]
        g = csr_from_edges(n, edges)
        c = rng.integers(0, n, n).astype(np.int32)
        off2, col2, w2, rank = orc.coarsen(g.offsets, g.cols, None, c)
        used = sorted(set(c.tolist()))
        exp_rank = {x: i for i, x in enumerate(used)}
        for x in range(n):
            assert rank[x] == exp_rank.get(x, -1)
        agg = {}
        for u in range(n):
            for e in range(g.offsets[u], g.offsets[u + 1]):
                key = (exp_rank[c[u]], exp_rank[c[g.cols[e]]])
                agg[key] = agg.get(key, 0) + 1
        C = len(used)
        assert len(off2) == C + 1
        got = {}
        for r in range(C):
            cols = col2[off2[r]:off2[r + 1]]
            assert np.all(np.diff(cols) > 0)  # sorted, unique
            for j in range(off2[r], off2[r + 1]):
                got[(r, int(col2[j]))] = int(w2[j])
        assert got == agg
        assert int(w2.sum()) == g.nnz
        for (r, s_), x in got.items():
            assert got[(s_, r)] == x


def test_levels_conserve_modularity(mgl, orc):
    g = mgl.gen_rmat(12, 16, seed=4)
    R = orc.louvain(g.offsets, g.cols, None)
    assert len(R.levels) >= 2
    for l in range(len(R.levels) - 1):
        go, gc, gw = R.graphs[l + 1]
        n1 = len(go) - 1
        wi = gw.astype(np.int64)
        k = orc.vertex_weights(go, wi)
        q_start = orc.modularity(go, gc, wi, np.arange(n1, dtype=np.int32), k, int(wi.sum()))
        assert q_start == R.levels[l]["q"]
        # row sums of the coarse graph are the community totals
        assert int(wi.sum()) == g.nnz


def test_oracle_deterministic_and_monotone(mgl, orc):
    g = mgl.gen_rmat(14, 16, seed=1)
    A = orc.louvain(g.offsets, g.cols, None)
    B = orc.louvain(g.offsets, g.cols, None)
    np.testing.assert_array_equal(A.final, B.final)
    for la, lb in zip(A.levels, B.levels):
        np.testing.assert_array_equal(la["qtraj"], lb["qtraj"])
        # committed Q never decreases inside a level
        assert np.all(np.diff(la["qtraj"]) >= 0)


def test_exact_and_fp64_modes_agree_on_dyadic_weights(mgl, orc):
    g = mgl.gen_sbm(4000, 16, min_size=64, seed=9, weights=1)
    A = orc.louvain(g.offsets, g.cols, g.w.astype(np.float64), mode=0)
    B = orc.louvain(g.offsets, g.cols, g.w.astype(np.float64), mode=1)
    assert len(A.levels) == len(B.levels)
    np.testing.assert_array_equal(A.final, B.final)
    assert A.final_q == pytest.approx(B.final_q, abs=1e-12)


def test_u128_to_double_rounding(orc):
    rng = np.random.default_rng(5)
    vals = [0, 1, 2 ** 53 + 1, 2 ** 64 - 1, 2 ** 64, 2 ** 64 + 2 ** 11 + 1, 2 ** 127 + 2 ** 74,
            2 ** 100 + 2 ** 47, 2 ** 100 + 2 ** 47 + 1, 2 ** 100 + 3 * 2 ** 47]
    vals += [int(rng.integers(0, 2 ** 63)) * 2 ** int(rng.integers(0, 64)) + int(rng.integers(0, 2 ** 62))
             for _ in range(300)]
    for x in vals:
        x &= (1 << 128) - 1
        assert orc.lib.orc_u128_to_double(x >> 64, x & (2 ** 64 - 1)) == float(x), x
