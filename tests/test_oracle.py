"""Pins the CPU oracle (oracle/louvain_oracle.cpp) with checks that do not
retype its formulas:

  * move rule     : exact-rational dQ by full modularity recomputation on tiny
                    random graphs (fractions), for v0/v1/v2, both directions and
                    both tie rules; on graphs whose 2m is a power of two every
                    FP64 gain is exact, so ties and zero gains are checked too
  * known answers : two triangles joined by an edge; a path whose run has a
                    down-pass tie and a zero-gain move (hand-derived per pass)
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


def q_csr(off, col, w, c):
    """Modularity of assignment c on a CSR (directed entries, self-loops once)
    with fractions: sum of w over entries inside a community / 2m minus the
    squared community totals (row sums) / (2m)^2."""
    n = len(off) - 1
    wt = [Fraction(1)] * len(col) if w is None else [Fraction(x) for x in w]
    m2 = sum(wt)
    k = [sum(wt[off[v]:off[v + 1]], Fraction(0)) for v in range(n)]
    sin = sum((wt[e] for v in range(n) for e in range(off[v], off[v + 1])
               if c[v] == c[col[e]]), Fraction(0))
    tot = {}
    for v in range(n):
        tot[c[v]] = tot.get(c[v], 0) + k[v]
    return sin / m2 - sum(t * t for t in tot.values()) / (m2 * m2)


def rule_proposal(c, gains, up, version, tie):
    """The move rule applied to exact gains {v: {community: gain}}: start from
    (0, own community), higher gain wins, equal gains by id (the smaller on up
    passes or always with tie=1, else the larger); v1 filters candidates by
    direction first, v0/v2 clamp the result to the direction."""
    out = []
    for v, ci in enumerate(c):
        best, dst = Fraction(0), ci
        small = True if tie else up
        for x, gv in sorted(gains[v].items()):
            if version == 1 and not (x < ci if up else x > ci):
                continue
            if gv > best or (gv == best and (x < dst if small else x > dst)):
                best, dst = gv, x
        if version != 1:
            dst = min(dst, ci) if up else max(dst, ci)
        out.append(dst)
    return out


def exact_gains(off, col, w, c):
    """Delta-Q of moving each vertex alone into each neighbouring community."""
    q0 = q_csr(off, col, w, c)
    out = []
    for v in range(len(off) - 1):
        g = {}
        for e in range(off[v], off[v + 1]):
            x = int(c[col[e]])
            if x != c[v] and x not in g:
                c2 = list(c)
                c2[v] = x
                g[x] = q_csr(off, col, w, c2) - q0
        out.append(g)
    return out


def test_known_answer_path_ties(orc):
    with open(os.path.join(GOLDEN, "kat_path_ties.json")) as f:
        kat = json.load(f)
    g = csr_from_edges(kat["n"], kat["edges"])
    assert g.nnz == kat["two_m"]
    for rule in kat["rules"].values():
        # the derivation itself, from full modularity recomputations
        graphs = [(list(g.offsets), list(g.cols), None)]
        for lev in rule["levels"][1:]:
            gg = lev["graph"]
            graphs.append((gg["offsets"], gg["cols"], gg["w"]))
        for li, lev in enumerate(rule["levels"]):
            off, col, w = graphs[li]
            ident = list(range(lev["n"]))
            assert q_csr(off, col, w, ident) == Fraction(lev["q0"])
            for ps in lev["passes"]:
                ex = exact_gains(off, col, w, ps["c"])
                listed = [{int(x): Fraction(y) for x, y in ps["gains"][str(v)].items()}
                          for v in range(lev["n"])]
                assert ex == listed
                for ver in rule["versions"]:
                    assert rule_proposal(ps["c"], listed, ps["up"], ver, rule["tie"]) == \
                        ps["proposal"]
                assert q_csr(off, col, w, ps["proposal"]) == Fraction(ps["q"])
            if li + 1 < len(rule["levels"]):
                # the next level's graph is the coarsening by the committed assignment
                o2, c2, w2, _ = orc.coarsen(np.asarray(off), np.asarray(col),
                                            None if w is None else np.asarray(w, np.int64),
                                            np.asarray(lev["assign"], np.int32))
                nxt = rule["levels"][li + 1]["graph"]
                assert list(o2) == nxt["offsets"] and list(c2) == nxt["cols"]
                assert list(w2) == nxt["w"]
        # the oracle against the derivation
        for ver in rule["versions"]:
            R = orc.louvain(g.offsets, g.cols, None, version=ver, tie=rule["tie"])
            assert len(R.levels) == len(rule["levels"])
            for li, (lv, le) in enumerate(zip(R.levels, rule["levels"])):
                assert lv["n"] == le["n"] and lv["iters"] == le["iters"]
                assert list(lv["qtraj"]) == [float(Fraction(x)) for x in le["qtraj"]]
                np.testing.assert_array_equal(lv["assign"], le["assign"])
                if li >= 1:
                    go, gc, gw = R.graphs[li]
                    assert list(go) == le["graph"]["offsets"] and list(gc) == le["graph"]["cols"]
                    assert list(gw) == le["graph"]["w"]
            np.testing.assert_array_equal(R.final, rule["final"])
            assert R.final_q == float(Fraction(rule["final_q"]))
            # and pass by pass through the single-step entry point
            k = orc.vertex_weights(g.offsets)
            for ps in rule["levels"][0]["passes"]:
                c = np.asarray(ps["c"], np.int32)
                a = orc.community_totals(k, c, kat["n"])
                got = orc.move(g.offsets, g.cols, None, k, c, a, g.nnz, ps["up"], ver,
                               rule["tie"])
                np.testing.assert_array_equal(got, ps["proposal"])


@pytest.mark.parametrize("tie", [0, 1])
@pytest.mark.parametrize("version", [0, 1, 2])
def test_move_rule_exact_ties(orc, version, tie):
    # graphs with 2m in {8, 16, 32}: every division in the FP64 gain is by a
    # power of two, so the gains are exact and equal gains are equal doubles;
    # no row is skipped, ties and zero gains included
    rng = np.random.default_rng(77 + 10 * version + tie)
    checked = ties = zeros = 0
    for trial in range(150):
        n = int(rng.integers(4, 10))
        pairs = [(u, v) for u in range(n) for v in range(u + 1, n)]
        ne = int(rng.choice([4, 8, 16]))
        if ne > len(pairs):
            continue
        idx = rng.choice(len(pairs), ne, replace=False)
        edges = [pairs[i] for i in idx]
        g = csr_from_edges(n, edges)
        # coarse random assignments produce equal community totals often
        c = (rng.integers(0, n, n) // int(rng.integers(1, 3))).astype(np.int32)
        k = orc.vertex_weights(g.offsets)
        a = orc.community_totals(k, c, n)
        ex = exact_gains(list(g.offsets), list(g.cols), None, list(c))
        for up in (True, False):
            got = orc.move(g.offsets, g.cols, None, k, c, a, g.nnz, up, version, tie)
            want = rule_proposal(list(c), ex, up, version, tie)
            assert list(got) == want, (trial, up, ex)
            checked += n
            for gv in ex:
                vals = list(gv.values())
                ties += sum(1 for x in set(vals) if vals.count(x) > 1)
                zeros += vals.count(0)
    assert checked > 1000 and ties > 20 and zeros > 20, (checked, ties, zeros)


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
