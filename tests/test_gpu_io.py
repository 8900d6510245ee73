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
