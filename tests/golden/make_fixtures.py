"""Writes the fixtures under tests/golden/.

kat_two_triangles.json: the hand-derived known answer for two triangles
  {0,1,2}, {3,4,5} joined by edge 2-3 (unit weights, 2m = 14): per pass the
  move rule's proposal and Q as exact fractions, the coarse graph and the final
  partition. Written literally here (derived by hand, not computed).
nx_modularity.json: networkx.community.modularity of fixed partitions of small
  seeded graphs (networkx is a library routine independent of this build).
Run: python tests/golden/make_fixtures.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def kat():
    edges = [(0, 1), (0, 2), (1, 2), (3, 4), (3, 5), (4, 5), (2, 3)]
    # k = [2,2,3,3,2,2], 2m = 14, Q = sum_in/14 - sum_c a_c^2/196
    # pass 1 (up): v0/v2 -> [0,0,0,3,4,4]: communities {0,1,2},{3},{4,5}:
    #              Q = 8/14 - (49 + 9 + 16)/196 = 38/196
    #              (vertex 2 ties between 0 and 1 and takes 0; vertex 3 prefers 4 > 3
    #              and is clamped to 3)
    #              v1 -> [0,0,0,2,3,4]: communities {0,1,2},{3},{4},{5}:
    #              Q = 6/14 - (49 + 9 + 4 + 4)/196 = 18/196
    # pass 2 (down): -> [0,0,0,4,4,4]: Q = 12/14 - (49 + 49)/196 = 5/14
    # pass 3: no vertex moves, rejected, level Q = 5/14
    coarse = {"offsets": [0, 2, 4], "cols": [0, 1, 0, 1], "w": [6, 1, 1, 6]}
    lvl1 = {"n": 2, "iters": 1, "qtraj": ["5/14"], "assign": [0, 1]}
    out = {"edges": edges, "versions": {}}
    for ver, q1 in (("0", "38/196"), ("1", "18/196"), ("2", "38/196")):
        out["versions"][ver] = {
            "levels": [{"n": 6, "iters": 3, "qtraj": [q1, "5/14", "5/14"],
                        "assign": [0, 0, 0, 4, 4, 4]}, lvl1],
            "coarse": coarse,
            "final": [0, 0, 0, 1, 1, 1],
        }
    with open(os.path.join(HERE, "kat_two_triangles.json"), "w") as f:
        json.dump(out, f, indent=1)


def nx_fixture():
    import networkx as nx
    rng = np.random.default_rng(2024)
    cases = []
    for t in range(6):
        n = int(rng.integers(20, 80))
        p = float(rng.uniform(0.05, 0.2))
        edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < p]
        weighted = t % 2 == 1
        # dyadic weights (multiples of 1/16 in (0.5, 1.5]) keep the exact mode exact
        w = [float(rng.integers(9, 25)) / 16.0 for _ in edges] if weighted else None
        G = nx.Graph()
        G.add_nodes_from(range(n))
        for i, (u, v) in enumerate(edges):
            G.add_edge(u, v, weight=w[i] if weighted else 1.0)
        assign = rng.integers(0, int(rng.integers(2, 9)), n)
        comms = [set(np.flatnonzero(assign == x).tolist()) for x in np.unique(assign)]
        q = nx.community.modularity(G, comms, weight="weight")
        cases.append({"name": f"case{t}", "n": n, "edges": edges, "w": w,
                      "assign": assign.tolist(), "q_networkx": q})
    with open(os.path.join(HERE, "nx_modularity.json"), "w") as f:
        json.dump({"networkx": nx.__version__, "cases": cases}, f)


if __name__ == "__main__":
    kat()
    nx_fixture()
