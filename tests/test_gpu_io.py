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
