"""GPU checks of the input/output paths: a Matrix Market file through the
reference loader semantics into the HIP path (against the oracle), and the
command-line front end (labels file equal to the library result)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def write_mtx(path, n, edges, banner="symmetric", weights=None):
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate real {banner}\n")
        f.write(f"{n} {n} {len(edges)}\n")
        for i, (u, v) in enumerate(edges):
            if weights is None:
                f.write(f"{u + 1} {v + 1}\n")
            else:
                f.write(f"{u + 1} {v + 1} {weights[i]}\n")


def test_mtx_symmetric_through_gpu(ctx, mgl, orc, tmp_path):
    rng = np.random.default_rng(8)
    n = 3000
    edges = [(int(u), int(v)) for u, v in rng.integers(0, n, (12000, 2)) if u > v]
    p = tmp_path / "g.mtx"
    write_mtx(p, n, edges)
    g = mgl.read_mtx(str(p))
    res = ctx.louvain(g, keep_levels=True)
    ref = orc.louvain(g.offsets, g.cols, None)
    assert len(res.levels) == len(ref.levels)
    for lg, lr in zip(res.levels, ref.levels):
        assert lg.q == lr["q"]
        np.testing.assert_array_equal(lg.assign, lr["assign"])


def test_mtx_general_weighted_through_gpu(ctx, mgl, orc, tmp_path):
    rng = np.random.default_rng(9)
    n = 2000
    pairs = rng.integers(0, n, (9000, 2))
    edges = [(int(u), int(v)) for u, v in pairs if u != v]
    w = [float(x) / 8 for x in rng.integers(1, 17, len(edges))]
    p = tmp_path / "w.mtx"
    write_mtx(p, n, edges, banner="general", weights=w)
    g = mgl.read_mtx(str(p))
    res = ctx.louvain(g, precision=mgl.PREC_EXACT, keep_levels=True)
    ref = orc.louvain(g.offsets, g.cols, g.w.astype(np.float64), mode=0)
    assert len(res.levels) == len(ref.levels)
    np.testing.assert_array_equal(res.final_assign, ref.final)
    assert res.final_q == ref.final_q


def _cli_tables(stdout):
    """Per-phase rows (loop, Q, dQ, time s, time ms) of the reference's table
    layout (louvain.cu:1536-1538, :1654)."""
    phases, cur = [], None
    for line in stdout.splitlines():
        if line.startswith("Phase "):
            cur = []
            phases.append(cur)
        elif line.startswith("| ") and cur is not None and not line.startswith("| Loop"):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            assert len(cells) == 5, line
            # the reference prints "%-10d" / "%-10f": left-aligned in 10 columns
            assert all(len(c) >= 1 for c in cells)
            cur.append((int(cells[0]), float(cells[1]), float(cells[2]), float(cells[3]),
                        float(cells[4])))
    return phases


def test_cli_table_layout(ctx, mgl, tmp_path):
    # the path graph of the hand-derived known answer: per phase a loop-0 row
    # with the singleton Q, then one row per iteration with new_Q and
    # dQ = new_Q - cur_Q (0 after a rejected proposal), times per iteration
    cli = os.path.join(ROOT, "mg-alphagcd_amd", "mgl_cli")
    p = tmp_path / "path.mtx"
    write_mtx(p, 5, [(2, 0), (4, 0), (2, 1), (4, 3)])
    r = subprocess.run([cli, "-path_of_graph", str(p)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    print(r.stdout)
    hdr = "| Loop       | Q          | dQ         | time(s)    | time(ms)   |"
    assert hdr in r.stdout
    assert "|------------|------------|------------|------------|------------|" in r.stdout
    ph = _cli_tables(r.stdout)
    assert len(ph) == 2
    q = [(0, -7 / 32, 0.0), (1, 5 / 32, 12 / 32), (2, 7 / 32, 2 / 32), (3, 7 / 32, 0.0)]
    assert [row[0] for row in ph[0]] == [0, 1, 2, 3]
    for row, (loop, qq, dq) in zip(ph[0], q):
        assert row[1] == pytest.approx(qq, abs=1e-6) and row[2] == pytest.approx(dq, abs=1e-6)
    assert ph[0][0][3:] == (0.0, 0.0)
    assert all(row[3] >= 0 and row[4] == pytest.approx(row[3] * 1000, abs=2e-3)
               for row in ph[0][1:])
    assert [row[0] for row in ph[1]] == [0, 1]
    assert ph[1][0][1] == pytest.approx(7 / 32, abs=1e-6) and ph[1][1][2] == 0.0
    for key in ("Main loop total execution time:", "The average execution time per loop:",
                "Coarsen graph execution time:", "Total time for clustering   :",
                "Total time for coarsen graph:", "TOTAL TIME                  :"):
        assert key in r.stdout
    # the level's Q trajectory and its singleton start through the library
    res = ctx.louvain(mgl.read_mtx(str(p)))
    assert res.levels[0].q0 == -7 / 32 and list(res.levels[0].qtraj) == [5 / 32, 7 / 32, 7 / 32]
    assert len(res.levels[0].iter_seconds) == 3 and np.all(res.levels[0].iter_seconds > 0)


def test_cli_labels_match_library(ctx, mgl, tmp_path):
    cli = os.path.join(ROOT, "mg-alphagcd_amd", "mgl_cli")
    out = tmp_path / "labels.txt"
    r = subprocess.run([cli, "-rmat", "13,16,5", "-version", "2", "-out", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "| Loop" in r.stdout and "Final modularity" in r.stdout
    labels = np.loadtxt(out, dtype=np.int32)
    res = ctx.louvain(mgl.gen_rmat(13, 16, seed=5))
    np.testing.assert_array_equal(labels, res.final_assign)


def _read_bin_assign(path):
    raw = open(path, "rb").read()
    assert raw[:8] == b"MGLASG1\0"
    n = int(np.frombuffer(raw[8:16], np.int64)[0])
    return np.frombuffer(raw[16:16 + 4 * n], np.int32)


def _read_bin_csr(path):
    raw = open(path, "rb").read()
    assert raw[:8] == b"MGLCSR1\0"
    n, m = (int(x) for x in np.frombuffer(raw[8:24], np.int64))
    o = 24
    off = np.frombuffer(raw[o:o + 8 * (n + 1)], np.int64)
    o += 8 * (n + 1)
    col = np.frombuffer(raw[o:o + 4 * m], np.int32)
    o += 4 * m
    w = np.frombuffer(raw[o:o + 8 * m], np.float64)
    assert o + 8 * m == len(raw)
    return off, col, w


@pytest.mark.parametrize("binary", [False, True])
def test_result_files_round_trip(ctx, mgl, tmp_path, binary):
    # mgl_write_result: levels table, composed assignment, dendrogram and coarse
    # graphs, read back and compared with the in-memory result
    g = mgl.gen_rmat(12, 16, seed=7)
    prefix = str(tmp_path / "run")
    res = ctx.louvain(g, keep_levels=True, keep_graphs=True, dump=prefix, dump_binary=binary)
    ext = ".bin" if binary else ".txt"
    rows = [l.split() for l in open(prefix + ".levels.txt") if not l.startswith("#")]
    assert len(rows) == len(res.levels) >= 2
    for row, lev in zip(rows, res.levels):
        assert int(row[1]) == lev.n and int(row[2]) == lev.iters
        assert float(row[3]) == lev.q
        np.testing.assert_array_equal(np.array(row[4:], float), lev.qtraj)
    rd = _read_bin_assign if binary else (lambda p: np.loadtxt(p, dtype=np.int32, ndmin=1))
    np.testing.assert_array_equal(rd(prefix + ".final" + ext), res.final_assign)
    for l, lev in enumerate(res.levels):
        np.testing.assert_array_equal(rd(f"{prefix}.level{l}{ext}"), lev.assign)
        if l == 0:
            continue
        if binary:
            off, col, w = _read_bin_csr(f"{prefix}.level{l}.csr")
        else:
            import scipy.io
            m = scipy.io.mmread(f"{prefix}.level{l}.mtx").tocsr()
            m.sort_indices()
            off, col, w = m.indptr, m.indices, m.data
        np.testing.assert_array_equal(off, lev.graph.offsets)
        np.testing.assert_array_equal(col, lev.graph.cols)
        np.testing.assert_array_equal(w, lev.graph_w64)
        # the dendrogram step: level l's vertices are the ranks of level l-1's ids
        prev = res.levels[l - 1].assign
        assert lev.n == len(np.unique(prev))


def test_cli_dump(ctx, mgl, tmp_path):
    cli = os.path.join(ROOT, "mg-alphagcd_amd", "mgl_cli")
    prefix = str(tmp_path / "cli")
    r = subprocess.run([cli, "-rmat", "12,16,7", "-dump", prefix, "-dump_binary", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = ctx.louvain(mgl.gen_rmat(12, 16, seed=7), keep_levels=True)
    np.testing.assert_array_equal(_read_bin_assign(prefix + ".final.bin"), res.final_assign)
    np.testing.assert_array_equal(_read_bin_assign(prefix + ".level0.bin"), res.levels[0].assign)
