"""ctypes binding of the CPU oracle (oracle/liboracle.so). Test infrastructure:
only tests, the smoke check and bench.py's CPU-baseline leg load it."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "liboracle.so")
if not os.path.exists(PATH):
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                   stdout=subprocess.DEVNULL)
lib = C.CDLL(PATH)

_vp, _i64, _i32, _d, _i = C.c_void_p, C.c_int64, C.c_int32, C.c_double, C.c_int
P64, P32, PD = C.POINTER(_i64), C.POINTER(_i32), C.POINTER(_d)


def _s(name, res, args):
    f = getattr(lib, name)
    f.restype, f.argtypes = res, args


_s("orc_louvain", _vp, [_i64, P64, P32, PD, _i, _d, _i, _i, _i, _i])
_s("orc_res_levels", _i, [_vp])
_s("orc_res_n", _i64, [_vp, _i])
_s("orc_res_q", _d, [_vp, _i])
_s("orc_res_iters", _i, [_vp, _i])
_s("orc_res_qtraj", PD, [_vp, _i])
_s("orc_res_assign", P32, [_vp, _i])
_s("orc_res_final", P32, [_vp])
_s("orc_res_ngraphs", _i, [_vp])
_s("orc_res_graph_n", _i64, [_vp, _i])
_s("orc_res_graph_off", P64, [_vp, _i])
_s("orc_res_graph_col", P32, [_vp, _i])
_s("orc_res_graph_w", None, [_vp, _i, PD])
_s("orc_res_free", None, [_vp])
_s("orc_dyadic_scale", _i, [_i64, PD])
_s("orc_u128_to_double", _d, [C.c_uint64, C.c_uint64])
_s("orc_vertex_weights_i64", None, [_i64, P64, P64, P64])
_s("orc_move_i64", None, [_i64, P64, P32, P64, P64, P32, P64, _i64, _i, _i, _i, P32])
_s("orc_update_weights_i64", None, [_i64, P64, P32, P32, P64])
_s("orc_modularity_i64", _d, [_i64, P64, P32, P64, P32, P64, _i64])
_s("orc_move_f64", None, [_i64, P64, P32, PD, PD, P32, PD, _d, _i, _i, _i, P32])
_s("orc_modularity_f64", _d, [_i64, P64, P32, PD, P32, PD, _d])
_s("orc_coarsen_i64", _i64, [_i64, P64, P32, P64, P32, P64, P32, P64, P32])
_s("orc_pass_i64", _d, [_i64, P64, P32, P64, P64, P32, P64, _i64, _i, _i, P32, _d, _d])
_s("orc_num_threads", _i, [])


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _arr(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy()


class Result:
    pass


def louvain(offsets, cols, w=None, mode=0, threshold=1e-6, max_phases=1000, max_iter=0,
            version=2, tie=0):
    """Full multi-level run. mode 0 = exact (unit or dyadic weights), 1 = fp64.
    tie 0 = the reference's tie rule (alternating), 1 = always the smallest id.
    Returns levels [(n, q, iters, qtraj, assign)], graphs [(off, col, w)], final."""
    off = np.ascontiguousarray(offsets, np.int64)
    col = np.ascontiguousarray(cols, np.int32)
    wd = None if w is None else np.ascontiguousarray(w, np.float64)
    n = len(off) - 1
    h = lib.orc_louvain(n, _p(off, _i64), _p(col, _i32), None if wd is None else _p(wd, _d),
                        mode, threshold, max_phases, max_iter, version, tie)
    if not h:
        raise ValueError("oracle refused the weights for exact mode")
    R = Result()
    R.levels = []
    for l in range(lib.orc_res_levels(h)):
        nl, it = lib.orc_res_n(h, l), lib.orc_res_iters(h, l)
        R.levels.append(dict(n=nl, q=lib.orc_res_q(h, l), iters=it,
                             qtraj=_arr(lib.orc_res_qtraj(h, l), it, np.float64),
                             assign=_arr(lib.orc_res_assign(h, l), nl, np.int32)))
    R.graphs = []
    for l in range(lib.orc_res_ngraphs(h)):
        gn = lib.orc_res_graph_n(h, l)
        go = _arr(lib.orc_res_graph_off(h, l), gn + 1, np.int64)
        gm = int(go[-1]) if gn >= 0 and len(go) else 0
        gc = _arr(lib.orc_res_graph_col(h, l), gm, np.int32)
        gw = np.zeros(gm, np.float64)
        if gm:
            lib.orc_res_graph_w(h, l, _p(gw, _d))
        R.graphs.append((go, gc, gw))
    n0 = lib.orc_res_graph_n(h, 0)  # the input graph (kept as graphs[0])
    R.final = _arr(lib.orc_res_final(h), n0, np.int32)
    R.final_q = R.levels[-1]["q"] if R.levels else 0.0
    lib.orc_res_free(h)
    return R


def scale_of(w):
    wd = np.ascontiguousarray(w, np.float64)
    return lib.orc_dyadic_scale(len(wd), _p(wd, _d))


def int_weights(w, S):
    """Dyadic weights scaled to int64 units (exact)."""
    return np.ldexp(np.asarray(w, np.float64), S).astype(np.int64)


def vertex_weights(off, w=None):
    off = np.ascontiguousarray(off, np.int64)
    n = len(off) - 1
    k = np.zeros(max(n, 1), np.int64)
    lib.orc_vertex_weights_i64(n, _p(off, _i64), None if w is None else _p(w, _i64), _p(k, _i64))
    return k[:n]


def move(off, col, w, k, c, a, mass, up, version=2, tie=0):
    off = np.ascontiguousarray(off, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    k = np.ascontiguousarray(k, np.int64)
    c = np.ascontiguousarray(c, np.int32)
    a = np.ascontiguousarray(a, np.int64)
    n = len(off) - 1
    out = np.empty(max(n, 1), np.int32)
    lib.orc_move_i64(n, _p(off, _i64), _p(col, _i32), None if w is None else _p(w, _i64),
                     _p(k, _i64), _p(c, _i32), _p(a, _i64), int(mass), int(up), version, tie,
                     _p(out, _i32))
    return out[:n]


def move_f64(off, col, w, k, c, a, mass, up, version=2, tie=0):
    """FP64 move pass (weights, totals and 2m as doubles)."""
    off = np.ascontiguousarray(off, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    wd = None if w is None else np.ascontiguousarray(w, np.float64)
    k = np.ascontiguousarray(k, np.float64)
    c = np.ascontiguousarray(c, np.int32)
    a = np.ascontiguousarray(a, np.float64)
    n = len(off) - 1
    out = np.empty(max(n, 1), np.int32)
    lib.orc_move_f64(n, _p(off, _i64), _p(col, _i32), None if wd is None else _p(wd, _d),
                     _p(k, _d), _p(c, _i32), _p(a, _d), float(mass), int(up), version, tie,
                     _p(out, _i32))
    return out[:n]


def modularity(off, col, w, c, a, mass):
    off = np.ascontiguousarray(off, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    c = np.ascontiguousarray(c, np.int32)
    a = np.ascontiguousarray(a, np.int64)
    return lib.orc_modularity_i64(len(off) - 1, _p(off, _i64), _p(col, _i32),
                                  None if w is None else _p(w, _i64), _p(c, _i32), _p(a, _i64),
                                  int(mass))


def community_totals(k, c, n):
    a = np.zeros(n, np.int64)
    np.add.at(a, c, k)
    return a


def coarsen(off, col, w, c):
    off = np.ascontiguousarray(off, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    c = np.ascontiguousarray(c, np.int32)
    n = len(off) - 1
    m = int(off[-1])
    off2 = np.zeros(n + 1, np.int64)
    col2 = np.zeros(max(m, 1), np.int32)
    w2 = np.zeros(max(m, 1), np.int64)
    rank = np.zeros(max(n, 1), np.int32)
    C2 = lib.orc_coarsen_i64(n, _p(off, _i64), _p(col, _i32), None if w is None else _p(w, _i64),
                             _p(c, _i32), _p(off2, _i64), _p(col2, _i32), _p(w2, _i64),
                             _p(rank, _i32))
    m2 = int(off2[C2])
    return off2[:C2 + 1].copy(), col2[:m2].copy(), w2[:m2].copy(), rank[:n].copy()


def modularity_fp64(off, col, w, labels):
    """Independent FP64 modularity of a partition (sequential double sums):
    Q = sum_in/2m - sum_c tot_c^2/(2m)^2 with tot_c = sum of row weights."""
    off = np.ascontiguousarray(off, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    lab = np.ascontiguousarray(labels, np.int32)
    n = len(off) - 1
    wd = None if w is None else np.ascontiguousarray(w, np.float64)
    deg = np.diff(off)
    if wd is None:
        k = deg.astype(np.float64)
        mass = float(len(col))
    else:
        rows = np.repeat(np.arange(n, dtype=np.int64), deg)
        k = np.bincount(rows, weights=wd, minlength=n).astype(np.float64)
        mass = float(wd.sum())
    tot = np.zeros(max(n, 1), np.float64)
    np.add.at(tot, lab, k)
    a = np.zeros(max(n, 1), np.float64)
    a[:n] = tot[:n]
    return lib.orc_modularity_f64(n, _p(off, _i64), _p(col, _i32),
                                  None if wd is None else _p(wd, _d), _p(lab, _i32), _p(a, _d),
                                  mass)
