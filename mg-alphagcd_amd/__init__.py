"""mg-alphagcd_amd: Louvain community detection on one AMD Instinct MI355X.

Python binding of libmgl.so (C ABI in include/mgl.h). The compute path is the
library's gfx950 HIP kernels; this module only marshals arrays. Importing it
without a built libmgl.so raises; there is no CPU fallback.

The reference (NikodemGapski/MG-alphaGCD) is driven as
    mpirun -np 1 ./MG_GCD -path_of_graph X.mtx [-version v] [-threshold t]
                             [-max_phases p] [-max_iter i]       (src/main.cu:20-26)
which builds HostGraph (src/graph/host_graph.cpp:5-17) and GpuGraph
(src/graph/gpu_graph.cu:3-18) and calls X::run(HostGraph*, GpuGraph*,
threshold, max_iter, max_phases) (include/louvain/louvain.cuh:11). The same
names are mirrored here: HostGraph, GpuGraph and run(); run() returns the
result instead of printing only.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmgl.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()' "
        "or make -C mg-alphagcd_amd/csrc)")

_lib = C.CDLL(LIB_PATH)

# status codes (mgl_status)
OK, INVALID_ARG, INVALID_GRAPH, OOM, DEVICE_ERROR, UNSUPPORTED = range(6)
_STATUS = {0: "OK", 1: "INVALID_ARG", 2: "INVALID_GRAPH", 3: "OOM", 4: "DEVICE_ERROR",
           5: "UNSUPPORTED"}
PREC_AUTO, PREC_EXACT, PREC_FP64, PREC_FP32_HASH = 0, 1, 2, 3
TIE_REFERENCE, TIE_SMALLEST = 0, 1
LAYOUT_AUTO, LAYOUT_VERTEX, LAYOUT_BINNED = 0, 1, 2


class MglError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


class _Csr(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnz", C.c_int64), ("offsets", C.c_void_p),
                ("cols", C.c_void_p), ("w", C.c_void_p), ("w64", C.c_void_p)]


class _HostCsr(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnz", C.c_int64), ("offsets", C.POINTER(C.c_int64)),
                ("cols", C.POINTER(C.c_int32)), ("w", C.POINTER(C.c_float)),
                ("w64", C.POINTER(C.c_double))]


class _Opts(C.Structure):
    _fields_ = [("version", C.c_int), ("threshold", C.c_double), ("max_phases", C.c_int),
                ("max_iter", C.c_int), ("precision", C.c_int), ("tie_rule", C.c_int),
                ("fuse_modularity", C.c_int), ("keep_levels", C.c_int),
                ("keep_graphs", C.c_int), ("validate", C.c_int), ("tuning_file", C.c_char_p),
                ("use_graph", C.c_int), ("record_layout", C.c_int)]


class PassStats(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnz", C.c_int64), ("move_ms", C.c_double),
                ("update_ms", C.c_double), ("modularity_ms", C.c_double),
                ("pass_ms", C.c_double), ("dominant_ms", C.c_double),
                ("dominant_bin", C.c_int), ("dominant_vertices", C.c_int64),
                ("dominant_entries", C.c_int64), ("moved", C.c_int64),
                ("distinct", C.c_int64), ("dominant_distinct", C.c_int64),
                ("q", C.c_double), ("dominant_kernel", C.c_char * 96), ("hub_ms", C.c_double),
                ("hub_vertices", C.c_int64), ("hub_entries", C.c_int64),
                ("hub_distinct", C.c_int64), ("lds_inserts", C.c_int64),
                ("lds_probes", C.c_int64), ("lds_multi_probe", C.c_int64),
                ("n_bins", C.c_int), ("bin_ms", C.c_double * 16), ("hub_beside", C.c_int)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["dominant_kernel"] = self.dominant_kernel.decode()
        d["bin_ms"] = list(self.bin_ms)[:self.n_bins]
        return d


def _sig(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args
    return f


_vp, _i64, _i32, _d, _i = C.c_void_p, C.c_int64, C.c_int32, C.c_double, C.c_int
_sig("mgl_opts_default", None, [C.POINTER(_Opts)])
_sig("mgl_create", _i, [C.POINTER(_vp), _i])
_sig("mgl_destroy", None, [_vp])
_sig("mgl_last_error", C.c_char_p, [_vp])
_sig("mgl_louvain", _i, [_vp, C.POINTER(_Csr), C.POINTER(_Opts), C.POINTER(_vp)])
_sig("mgl_louvain_batched", _i, [_vp, C.POINTER(_Csr), _i32, C.POINTER(_Opts), C.POINTER(_vp)])
_sig("mgl_result_levels", _i, [_vp])
_sig("mgl_result_level_n", _i64, [_vp, _i])
_sig("mgl_result_level_q", _d, [_vp, _i])
_sig("mgl_result_level_iters", _i, [_vp, _i])
_sig("mgl_result_level_qtraj", C.POINTER(_d), [_vp, _i])
_sig("mgl_result_level_q0", _d, [_vp, _i])
_sig("mgl_result_level_iter_seconds", C.POINTER(_d), [_vp, _i])
_sig("mgl_result_level_assign", C.POINTER(_i32), [_vp, _i])
_sig("mgl_result_level_graph", _i, [_vp, _i, C.POINTER(_i64), C.POINTER(_i64),
                                     C.POINTER(C.POINTER(_i64)), C.POINTER(C.POINTER(_i32)),
                                     C.POINTER(C.POINTER(_d))])
_sig("mgl_result_final_assign", C.POINTER(_i32), [_vp])
_sig("mgl_result_vertices", _i64, [_vp])
_sig("mgl_result_level_seconds", _d, [_vp, _i, C.POINTER(_d)])
_sig("mgl_result_final_q", _d, [_vp])
_sig("mgl_result_precision", _i, [_vp])
_sig("mgl_result_seconds", _d, [_vp])
_sig("mgl_result_free", None, [_vp])
_sig("mgl_session_create", _i, [_vp, C.POINTER(_Csr), C.POINTER(_Opts), C.POINTER(_vp)])
_sig("mgl_session_destroy", None, [_vp])
_sig("mgl_session_reset", _i, [_vp, C.POINTER(_d)])
_sig("mgl_session_pass", _i, [_vp, C.POINTER(_d), C.POINTER(_i)])
_sig("mgl_session_stats", _i, [_vp, C.POINTER(PassStats)])
_sig("mgl_session_count", _i, [_vp])
_sig("mgl_session_bins", _i, [_vp, C.POINTER(_i64), C.POINTER(_i)])
_sig("mgl_session_set_assign", _i, [_vp, C.POINTER(_i32)])
_sig("mgl_session_move", _i, [_vp, _i, C.POINTER(_i32)])
_sig("mgl_session_modularity", _i, [_vp, C.POINTER(_d)])
_sig("mgl_session_get_weights", _i, [_vp, C.POINTER(_d), C.POINTER(_d)])
_sig("mgl_session_apply", _i, [_vp, C.POINTER(_i32)])
_sig("mgl_session_coarsen", _i, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64),
                                  C.POINTER(_i32), C.POINTER(_d), C.POINTER(_i32)])
_sig("mgl_gen_rmat", _i, [_i, _d, _d, _d, _d, C.c_uint64, _i, _i64, C.POINTER(_HostCsr)])
_sig("mgl_gen_sbm", _i, [_i64, _i64, _d, _i64, _d, _d, C.c_uint64, _i, C.POINTER(_HostCsr)])
_sig("mgl_read_mtx", _i, [C.c_char_p, C.POINTER(_HostCsr)])
_sig("mgl_write_assign", _i, [C.c_char_p, C.POINTER(_i32), _i64])
_sig("mgl_write_result", _i, [C.c_void_p, C.c_char_p, _i])
_sig("mgl_csr_free", None, [C.POINTER(_HostCsr)])
_sig("mgl_version_major", _i, [])
_sig("mgl_bin_bounds", _i, [_vp, C.POINTER(_i64), _i])
_sig("mgl_autotune", _i, [_vp, C.POINTER(_Csr), C.c_char_p])
_sig("mgl_device_lds_bytes", _i, [_vp])
_sig("mgl_membench", _i, [_vp, _i64, _i, C.POINTER(_d)])
_sig("mgl_memory_stats", _i, [C.POINTER(_i64), C.POINTER(_i64), _i])
_sig("mgl_modularity_dev", _i, [_vp, _i64, _i64, _vp, _vp, _vp, _vp, C.POINTER(_d)])
_sig("mgl_modularity_dev_noalloc", _i, [_vp, _i64, _i64, _vp, _vp, _vp, _vp, C.POINTER(_d)])
NOALLOC_MAX_VERTICES = 1 << 23


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


# ---------------------------------------------------------------------------
@dataclass
class CSR:
    """Host CSR graph: directed entries, an undirected edge in both rows.
    Weights: None (unit), float64 (passed as doubles, the reference's
    weight_t) or any other dtype, converted to float32."""
    offsets: np.ndarray  # int64 [n+1]
    cols: np.ndarray     # int32 [nnz]
    w: Optional[np.ndarray] = None  # float32 or float64 [nnz], or None

    def __post_init__(self):
        self.offsets = np.ascontiguousarray(self.offsets, dtype=np.int64)
        self.cols = np.ascontiguousarray(self.cols, dtype=np.int32)
        if self.w is not None:
            dt = np.float64 if np.asarray(self.w).dtype == np.float64 else np.float32
            self.w = np.ascontiguousarray(self.w, dtype=dt)
        # the C ABI reads nnz entries of each array: short arrays are refused here
        if self.offsets.ndim != 1 or len(self.offsets) < 1:
            raise ValueError("offsets must hold n+1 >= 1 entries")
        m = int(self.offsets[-1])
        if m < 0 or len(self.cols) < m:
            raise ValueError(f"cols holds {len(self.cols)} entries, offsets[-1] = {m}")
        if self.w is not None and len(self.w) < m:
            raise ValueError(f"w holds {len(self.w)} entries, offsets[-1] = {m}")

    @property
    def n(self) -> int:
        return len(self.offsets) - 1

    @property
    def nnz(self) -> int:
        return int(self.offsets[-1]) if len(self.offsets) else 0

    def _c(self) -> _Csr:
        w32 = self.w if self.w is not None and self.w.dtype == np.float32 else None
        w64 = self.w if self.w is not None and self.w.dtype == np.float64 else None
        return _Csr(self.n, self.nnz, self.offsets.ctypes.data, self.cols.ctypes.data,
                    w32.ctypes.data if w32 is not None else None,
                    w64.ctypes.data if w64 is not None else None)

    @staticmethod
    def from_edges(n: int, edges, weights=None, w64: bool = False) -> "CSR":
        """Symmetric CSR from undirected edges (u, v[, w]); self-loops kept once.
        Weights are stored as float32, or as float64 with w64=True."""
        e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
        u, v = e[:, 0], e[:, 1]
        ww = None if weights is None else np.asarray(weights, dtype=np.float64)
        loop = u == v
        ru = np.concatenate([u, v[~loop]])
        rv = np.concatenate([v, u[~loop]])
        rw = None if ww is None else np.concatenate([ww, ww[~loop]])
        order = np.lexsort((rv, ru))
        ru, rv = ru[order], rv[order]
        off = np.zeros(n + 1, dtype=np.int64)
        np.add.at(off, ru + 1, 1)
        off = np.cumsum(off)
        return CSR(off, rv.astype(np.int32),
                   None if rw is None else rw[order].astype(np.float64 if w64 else np.float32))


def _take_host_csr(h: _HostCsr) -> CSR:
    n, m = h.n, h.nnz
    off = np.ctypeslib.as_array(h.offsets, shape=(n + 1,)).copy()
    cols = np.ctypeslib.as_array(h.cols, shape=(max(m, 1),))[:m].copy()
    w = None
    if h.w64:
        w = np.ctypeslib.as_array(h.w64, shape=(max(m, 1),))[:m].copy()
    elif h.w:
        w = np.ctypeslib.as_array(h.w, shape=(max(m, 1),))[:m].copy()
    _lib.mgl_csr_free(C.byref(h))
    return CSR(off, cols, w)


def _chk(st: int):
    if st != OK:
        raise MglError(st, "host call failed")


def gen_rmat(scale: int, edge_factor: float = 16, a=0.57, b=0.19, c=0.19, seed: int = 1,
             weights: int = 0, n_limit: int = 0) -> CSR:
    """Seeded R-MAT graph, symmetrised, no self-loops or duplicates.
    weights: 0 unit, 1 uniform (0.5,1.5], 2 lognormal(0,1)."""
    h = _HostCsr()
    _chk(_lib.mgl_gen_rmat(scale, edge_factor, a, b, c, seed, weights, n_limit, C.byref(h)))
    return _take_host_csr(h)


def gen_sbm(n: int, n_comm: int, exponent=2.1, min_size=64, deg_in=24.0, deg_out=4.0,
            seed: int = 2, weights: int = 1) -> CSR:
    """Planted partition with power-law community sizes."""
    h = _HostCsr()
    _chk(_lib.mgl_gen_sbm(n, n_comm, exponent, min_size, deg_in, deg_out, seed, weights,
                          C.byref(h)))
    return _take_host_csr(h)


def memory_stats(reset_peak: bool = False):
    """(live, peak) bytes of device memory held by the library's calls."""
    live, peak = C.c_int64(), C.c_int64()
    _chk(_lib.mgl_memory_stats(C.byref(live), C.byref(peak), int(reset_peak)))
    return live.value, peak.value


def gen_batch_graph(i: int) -> CSR:
    """Graph i of the batched workload: vertex count log-uniform in [1,000,
    200,000] (seed 4,000,000 + i); even i a planted partition (8-256
    communities, power-law sizes, intra / inter mean degree 16 / 3), odd i an
    R-MAT graph (edge factor uniform in [8, 32]); weights uniform (0.5, 1.5]."""
    import math
    rng = np.random.default_rng(4_000_000 + i)
    n = int(round(math.exp(rng.uniform(math.log(1_000), math.log(200_000)))))
    seed = 4_000_000 + i
    if i % 2 == 0:
        k = int(rng.integers(8, 257))
        k = max(1, min(k, n // 8))
        return gen_sbm(n, k, exponent=2.1, min_size=8, deg_in=16, deg_out=3, seed=seed,
                       weights=1)
    ef = float(rng.uniform(8, 32))
    scale = max(1, math.ceil(math.log2(n)))
    return gen_rmat(scale, ef, 0.57, 0.19, 0.19, seed=seed, weights=1, n_limit=n)


def read_mtx(path: str) -> CSR:
    """Matrix Market file with the reference loader's semantics (general
    banner: weights as float64, the reference's weight_t)."""
    h = _HostCsr()
    _chk(_lib.mgl_read_mtx(path.encode(), C.byref(h)))
    return _take_host_csr(h)


# ---------------------------------------------------------------------------
@dataclass
class Level:
    n: int
    q: float
    iters: int
    qtraj: np.ndarray
    q0: float = 0.0  # Q of the singleton start ("Loop 0")
    iter_seconds: Optional[np.ndarray] = None  # device time of each iteration
    assign: Optional[np.ndarray] = None
    graph: Optional[CSR] = None  # input graph of this level (levels >= 1), weights as float64
    graph_w64: Optional[np.ndarray] = None
    seconds: float = 0.0
    coarsen_seconds: float = 0.0


@dataclass
class Result:
    levels: List[Level] = field(default_factory=list)
    final_assign: Optional[np.ndarray] = None
    final_q: float = 0.0
    precision: int = 0
    seconds: float = 0.0


def _read_result(r, dump: Optional[str] = None, dump_binary: bool = False) -> Result:
    if dump:
        st = _lib.mgl_write_result(r, dump.encode(), 1 if dump_binary else 0)
        if st != OK:
            _lib.mgl_result_free(r)
            raise MglError(st, f"mgl_write_result({dump!r}) failed")
    out = Result()
    L = _lib.mgl_result_levels(r)
    for l in range(L):
        n = _lib.mgl_result_level_n(r, l)
        it = _lib.mgl_result_level_iters(r, l)
        qp = _lib.mgl_result_level_qtraj(r, l)
        qt = np.ctypeslib.as_array(qp, shape=(it,)).copy() if (it > 0 and qp) else np.zeros(0)
        ap = _lib.mgl_result_level_assign(r, l)
        assign = np.ctypeslib.as_array(ap, shape=(n,)).copy() if (ap and n > 0) else None
        tp = _lib.mgl_result_level_iter_seconds(r, l)
        ts = np.ctypeslib.as_array(tp, shape=(it,)).copy() if (it > 0 and tp) else np.zeros(0)
        lev = Level(n=n, q=_lib.mgl_result_level_q(r, l), iters=it, qtraj=qt,
                    q0=_lib.mgl_result_level_q0(r, l), iter_seconds=ts, assign=assign)
        cs = C.c_double()
        lev.seconds = _lib.mgl_result_level_seconds(r, l, C.byref(cs))
        lev.coarsen_seconds = cs.value
        gn, gm = C.c_int64(), C.c_int64()
        po, pc, pw = C.POINTER(_i64)(), C.POINTER(_i32)(), C.POINTER(_d)()
        if l >= 1 and _lib.mgl_result_level_graph(r, l, C.byref(gn), C.byref(gm), C.byref(po),
                                                  C.byref(pc), C.byref(pw)) == OK:
            off = np.ctypeslib.as_array(po, shape=(gn.value + 1,)).copy()
            m = gm.value
            cols = np.ctypeslib.as_array(pc, shape=(m,)).copy() if m else np.zeros(0, np.int32)
            w64 = np.ctypeslib.as_array(pw, shape=(m,)).copy() if m else np.zeros(0)
            lev.graph = CSR(off, cols, None)
            lev.graph_w64 = w64
        out.levels.append(lev)
    n0 = _lib.mgl_result_vertices(r)
    fp = _lib.mgl_result_final_assign(r)
    out.final_assign = (np.ctypeslib.as_array(fp, shape=(n0,)).copy() if (fp and n0 > 0)
                        else np.zeros(0, np.int32))
    out.final_q = _lib.mgl_result_final_q(r)
    out.precision = _lib.mgl_result_precision(r)
    out.seconds = _lib.mgl_result_seconds(r)
    _lib.mgl_result_free(r)
    return out


def make_opts(version=2, threshold=1e-6, max_phases=1000, max_iter=0, precision=PREC_AUTO,
              tie_rule=TIE_REFERENCE, fuse_modularity=True, keep_levels=False,
              keep_graphs=False, validate=1, tuning_file=None, use_graph=True,
              record_layout=0) -> _Opts:
    o = _Opts()
    _lib.mgl_opts_default(C.byref(o))
    o.version, o.threshold, o.max_phases, o.max_iter = version, threshold, max_phases, max_iter
    o.precision, o.tie_rule = precision, tie_rule
    o.fuse_modularity, o.keep_levels, o.keep_graphs = int(fuse_modularity), int(keep_levels), int(keep_graphs)
    o.validate = int(validate)
    o.tuning_file = tuning_file.encode() if tuning_file else None
    o.use_graph = int(use_graph)
    o.record_layout = int(record_layout)
    return o


class Context:
    """One device and its streams (mgl_create)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        st = _lib.mgl_create(C.byref(h), device)
        if st != OK:
            raise MglError(st, f"mgl_create(device={device}) failed")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.mgl_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, st):
        if st != OK:
            raise MglError(st, _lib.mgl_last_error(self._h).decode(errors="replace"))

    def louvain(self, g: CSR, dump: Optional[str] = None, dump_binary: bool = False,
                **kw) -> Result:
        """Full multi-level run; with dump=PREFIX the result files are written
        (mgl_write_result: levels, final assignment, dendrogram, coarse graphs)."""
        o = make_opts(**kw)
        cg = g._c()
        r = C.c_void_p()
        self._err(_lib.mgl_louvain(self._h, C.byref(cg), C.byref(o), C.byref(r)))
        return _read_result(r, dump, dump_binary)

    def louvain_batched(self, graphs: Sequence[CSR], **kw) -> List[Result]:
        o = make_opts(**kw)
        arr = (_Csr * len(graphs))(*[g._c() for g in graphs])
        res = (C.c_void_p * len(graphs))()
        self._err(_lib.mgl_louvain_batched(self._h, arr, len(graphs), C.byref(o), res))
        return [_read_result(C.c_void_p(res[i])) for i in range(len(graphs))]

    def session(self, g: CSR, **kw) -> "Session":
        return Session(self, g, **kw)

    def bin_bounds(self):
        buf = (C.c_int64 * 16)()
        nb = _lib.mgl_bin_bounds(self._h, buf, 16)
        return [buf[i] for i in range(nb)]

    def membench(self, nbytes: int, iters: int = 5) -> float:
        g = C.c_double()
        self._err(_lib.mgl_membench(self._h, nbytes, iters, C.byref(g)))
        return g.value

    def lds_bytes(self) -> int:
        return _lib.mgl_device_lds_bytes(self._h)

    def modularity_dev(self, n: int, nnz: int, d_offsets: int, d_cols: int, d_w: int,
                       d_assign: int) -> float:
        """Modularity of a partition of a graph already in device memory
        (device addresses as integers; d_w = 0 for unit weights)."""
        q = C.c_double()
        self._err(_lib.mgl_modularity_dev(self._h, n, nnz, d_offsets, d_cols, d_w or None,
                                          d_assign, C.byref(q)))
        return q.value

    def modularity_dev_noalloc(self, n: int, nnz: int, d_offsets: int, d_cols: int, d_w: int,
                               d_assign: int) -> float:
        """modularity_dev without any device allocation (graphs of at most
        NOALLOC_MAX_VERTICES vertices; overlapping calls on one device race)."""
        q = C.c_double()
        self._err(_lib.mgl_modularity_dev_noalloc(self._h, n, nnz, d_offsets, d_cols,
                                                  d_w or None, d_assign, C.byref(q)))
        return q.value

    def autotune(self, g: CSR, path: str):
        cg = g._c()
        self._err(_lib.mgl_autotune(self._h, C.byref(cg), path.encode()))


class Session:
    """Level-0 graph resident in HBM, driven pass by pass or stage by stage."""

    def __init__(self, ctx: Context, g: CSR, **kw):
        self.ctx, self.g = ctx, g
        o = make_opts(**kw)
        cg = g._c()
        h = C.c_void_p()
        ctx._err(_lib.mgl_session_create(ctx._h, C.byref(cg), C.byref(o), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.mgl_session_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self) -> float:
        q = C.c_double()
        self.ctx._err(_lib.mgl_session_reset(self._h, C.byref(q)))
        return q.value

    def pass_(self):
        q, acc = C.c_double(), C.c_int()
        self.ctx._err(_lib.mgl_session_pass(self._h, C.byref(q), C.byref(acc)))
        return q.value, bool(acc.value)

    def count(self):
        """Counters of the next pass's move (distinct pairs, LDS probes), untimed,
        state unchanged; they appear in stats() after that pass."""
        self.ctx._err(_lib.mgl_session_count(self._h))

    def stats(self) -> dict:
        s = PassStats()
        self.ctx._err(_lib.mgl_session_stats(self._h, C.byref(s)))
        return s.as_dict()

    def bins(self):
        buf = (C.c_int64 * 16)()
        nb = C.c_int(16)
        self.ctx._err(_lib.mgl_session_bins(self._h, buf, C.byref(nb)))
        return [buf[i] for i in range(nb.value)]

    def set_assign(self, c):
        c = np.ascontiguousarray(c, dtype=np.int32)
        self.ctx._err(_lib.mgl_session_set_assign(self._h, _p(c, C.c_int32)))

    def move(self, up: bool) -> np.ndarray:
        out = np.empty(self.g.n, dtype=np.int32)
        self.ctx._err(_lib.mgl_session_move(self._h, int(up), _p(out, C.c_int32)))
        return out

    def modularity(self) -> float:
        q = C.c_double()
        self.ctx._err(_lib.mgl_session_modularity(self._h, C.byref(q)))
        return q.value

    def weights(self):
        a = np.empty(self.g.n, dtype=np.float64)
        k = np.empty(self.g.n, dtype=np.float64)
        self.ctx._err(_lib.mgl_session_get_weights(self._h, _p(a, C.c_double), _p(k, C.c_double)))
        return a, k

    def apply(self, c_new):
        c_new = np.ascontiguousarray(c_new, dtype=np.int32)
        self.ctx._err(_lib.mgl_session_apply(self._h, _p(c_new, C.c_int32)))

    def coarsen(self):
        n, m = self.g.n, self.g.nnz
        off = np.empty(n + 2, dtype=np.int64)
        col = np.empty(max(m, 1), dtype=np.int32)
        w = np.empty(max(m, 1), dtype=np.float64)
        rank = np.empty(max(n, 1), dtype=np.int32)
        n2, m2 = C.c_int64(), C.c_int64()
        self.ctx._err(_lib.mgl_session_coarsen(self._h, C.byref(n2), C.byref(m2),
                                               _p(off, C.c_int64), _p(col, C.c_int32),
                                               _p(w, C.c_double), _p(rank, C.c_int32)))
        C2, E2 = n2.value, m2.value
        return off[:C2 + 1].copy(), col[:E2].copy(), w[:E2].copy(), rank[:n].copy()


# ---------------------------------------------------------------------------
# Mirror of the reference's entry objects (HostGraph / GpuGraph / run).
class HostGraph:
    """host_graph.h:21 HostGraph(path): loads a .mtx file (mass = sum of weights = 2m)."""

    def __init__(self, path_or_csr):
        self.csr = read_mtx(path_or_csr) if isinstance(path_or_csr, str) else path_or_csr
        w = self.csr.w
        self.mass = float(self.csr.nnz if w is None else np.sum(w.astype(np.float64)))


class GpuGraph:
    """gpu_graph.cuh:30 GpuGraph(n_pes, HostGraph*): the device context holding it."""

    def __init__(self, host: HostGraph, device: int = 0):
        self.host = host
        self.ctx = Context(device)


def run(host: HostGraph, gpu: GpuGraph, threshold: float = 1e-6, max_iter: int = 500,
        max_phases: int = 1000, version: int = 2, **kw) -> Result:
    """X::run(HostGraph*, GpuGraph*, threshold, max_iter, max_phases)
    (louvain.cuh:11). Like the reference, max_iter is not applied unless
    honour_max_iter=True is passed."""
    honour = kw.pop("honour_max_iter", False)
    return gpu.ctx.louvain(host.csr, version=version, threshold=threshold,
                           max_phases=max_phases, max_iter=max_iter if honour else 0, **kw)
