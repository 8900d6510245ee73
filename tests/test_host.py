"""CPU-only checks of the host side: the C ABI library loads and exports every
entry point declared in include/mgl.h, the graph generators and the .mtx
reader produce canonical CSRs with the reference loader's semantics, and
host-side argument handling fails cleanly. No kernel is launched."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "mgl.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mgl_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(mgl):
    lib = C.CDLL(mgl.LIB_PATH)
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def canonical(g):
    d = np.diff(g.offsets)
    assert g.offsets[0] == 0 and np.all(d >= 0) and g.offsets[-1] == g.nnz
    rows = np.repeat(np.arange(g.n, dtype=np.int64), d)
    key = rows * g.n + g.cols
    assert np.all(np.diff(key) > 0), "rows sorted by column, no duplicates"
    assert not np.any(rows == g.cols), "no self-loops"
    assert np.array_equal(np.sort(g.cols.astype(np.int64) * g.n + rows), key), "symmetric"
    if g.w is not None:
        # the weight of (u,v) equals the weight of (v,u)
        order = np.argsort(g.cols.astype(np.int64) * g.n + rows, kind="stable")
        assert np.array_equal(g.w[order], g.w)


def test_rmat_canonical_and_deterministic(mgl):
    a = mgl.gen_rmat(12, 16, seed=1)
    b = mgl.gen_rmat(12, 16, seed=1)
    c = mgl.gen_rmat(12, 16, seed=2)
    canonical(a)
    assert a.n == 4096 and 0 < a.nnz <= 2 * 16 * 4096
    np.testing.assert_array_equal(a.offsets, b.offsets)
    np.testing.assert_array_equal(a.cols, b.cols)
    assert not np.array_equal(a.cols[:100], c.cols[:100]) or a.nnz != c.nnz
    # skew: a = 0.57 puts the largest degrees on low ids
    d = np.diff(a.offsets)
    assert d[0] == d.max()


def test_rmat_limit_and_weights(mgl):
    g = mgl.gen_rmat(12, 8, seed=3, n_limit=3000, weights=1)
    canonical(g)
    assert g.n == 3000 and g.cols.max() < 3000
    assert np.all(g.w > 0.5) and np.all(g.w <= 1.5)
    # every weight is an integer multiple of 2^-24
    assert np.all(np.ldexp(g.w.astype(np.float64), 24) % 1 == 0)
    h = mgl.gen_rmat(12, 8, seed=3, n_limit=3000, weights=2)
    canonical(h)
    assert np.all(h.w > 0)
    lw = np.log(h.w.astype(np.float64))
    assert abs(lw.mean()) < 0.1 and abs(lw.std() - 1.0) < 0.1


def test_sbm(mgl):
    g = mgl.gen_sbm(20000, 40, exponent=2.1, min_size=64, deg_in=24, deg_out=4, seed=2)
    canonical(g)
    mean_deg = g.nnz / g.n
    assert 22 < mean_deg < 29, mean_deg


def test_mtx_symmetric_semantics(mgl, tmp_path):
    # symmetric banner: unit weights, every entry kept, off-diagonal mirrored,
    # diagonal once, duplicates kept (mmio_wrapper.cu:141-170)
    p = tmp_path / "s.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n% c\n4 4 5\n"
                 "2 1 7.0\n3 1 1\n3 3 2\n4 2 1\n4 2 1\n")
    g = mgl.read_mtx(str(p))
    assert g.w is None and g.n == 4
    rows = np.repeat(np.arange(4), np.diff(g.offsets))
    got = sorted(zip(rows.tolist(), g.cols.tolist()))
    assert got == sorted([(1, 0), (0, 1), (2, 0), (0, 2), (2, 2), (3, 1), (1, 3), (3, 1), (1, 3)])


def test_mtx_general_semantics(mgl, tmp_path):
    # general banner: only col < row kept, weights parsed, mirrored (:233-263)
    p = tmp_path / "g.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n3 3 5\n"
                 "1 2 5.0\n2 1 0.5\n3 1 2\n3 3 9\n3 2\n")
    g = mgl.read_mtx(str(p))
    rows = np.repeat(np.arange(3), np.diff(g.offsets))
    got = sorted(zip(rows.tolist(), g.cols.tolist(), g.w.tolist()))
    assert got == sorted([(1, 0, 0.5), (0, 1, 0.5), (2, 0, 2.0), (0, 2, 2.0), (2, 1, 1.0),
                          (1, 2, 1.0)])


def test_from_edges(mgl):
    g = mgl.CSR.from_edges(4, [(0, 1), (1, 2), (3, 3)], [1.0, 2.0, 4.0])
    np.testing.assert_array_equal(g.offsets, [0, 1, 3, 4, 5])
    np.testing.assert_array_equal(g.cols, [1, 0, 2, 1, 3])
    np.testing.assert_array_equal(g.w, [1, 1, 2, 2, 4])


def test_context_without_device_fails_cleanly(mgl):
    # in a container without a GPU, mgl_create reports an error instead of aborting
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(mgl.MglError):
        mgl.Context(0)


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_gain_prefilter_matches_full_evaluation(tmp_path):
    # the move kernels' exact gain prefilter (Argmax in mgl_tables.h), built for
    # the CPU: on random candidate sets and constructed near-ties of the exact
    # gain numerator, the same argmax and the same double as evaluating every
    # candidate's FP64 gain
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = tmp_path / "argmax_check"
    subprocess.run([hipcc, "-O2", "-std=c++17", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "mg-alphagcd_amd", "csrc"),
                    os.path.join(ROOT, "tests", "argmax_check.cpp"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    for seed in (1, 2, 3):
        r = subprocess.run([str(exe), str(seed), "300000"], capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stdout
        assert "mismatches 0" in r.stdout
