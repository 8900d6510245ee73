/* mgl.h - C ABI of the MI355X Louvain engine (libmgl.so).
 *
 * The reference (NikodemGapski/MG-alphaGCD) has no plugin or FFI: its Louvain
 * path sits behind the MG_GCD command line (src/main.cu:15-135, flags
 * :20-26) and the in-process call
 *     X::run(HostGraph*, GpuGraph*, double threshold, int max_iter, int max_phases)
 * (include/louvain/louvain.cuh:11, louvain_gl.cuh:11, louvain_baseline.cuh:11)
 * selected by -version (src/main.cu:118-131). Its internal seams are
 * BIN::bin_create (include/bin/BIN.cuh:9-25), the binned move dispatcher
 * calculate_eicj_and_move_vertex_bin (src/louvain/louvain.cu:1032), the weight
 * update (:1299-1396), compute_modularity (:105-238) and
 * coarsen_graph_mg::coarsen_graph (include/coarsen_graph/coarsen_graph_mg.cuh:9).
 * Each entry point below names the reference interface it replaces.
 *
 * Conventions: plain pointers and sizes, no exceptions, no exit(); every call
 * returns an mgl_status and the context keeps the last error string. One
 * context = one device and its two streams (the move pass runs its hub path on
 * the second); a context is not thread safe, distinct contexts may be used from
 * distinct threads. Inputs are borrowed for the
 * duration of a call; results are owned by the library until freed.
 */
#ifndef MGL_H
#define MGL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  MGL_OK = 0,
  MGL_INVALID_ARG = 1,
  MGL_INVALID_GRAPH = 2,
  MGL_OOM = 3,
  MGL_DEVICE_ERROR = 4,
  MGL_UNSUPPORTED = 5
} mgl_status;

/* Precision of the sums (weights, community totals, modularity). */
enum {
  MGL_PREC_AUTO = 0,     /* EXACT when the weights allow it, else FP64 */
  MGL_PREC_EXACT = 1,    /* integer / dyadic fixed-point sums, bit-exact, order independent */
  MGL_PREC_FP64 = 2,     /* double weights and sums (reference weight_t = double) */
  MGL_PREC_FP32_HASH = 3 /* float edge weights and hash slots; FP64 for 2m, a_c and Q */
};

/* Tie rule between candidate communities of equal gain. */
enum {
  MGL_TIE_REFERENCE = 0, /* smallest id on up passes, largest on down passes (louvain.cu:544-554) */
  MGL_TIE_SMALLEST = 1   /* always the smallest id */
};

/* Host CSR. Directed entries; an undirected edge appears in both rows.
 * Replaces HostGraph's arrays (include/graph/host_graph.h:21-43: uint32
 * offsets and columns, weight_t = double weights, include/common.h:85-87).
 * Weights come as doubles (w64, the reference's weight_t), as floats (w), or
 * not at all (unit weights); w64 takes precedence when both are set. */
typedef struct {
  int64_t n;              /* vertices */
  int64_t nnz;            /* directed CSR entries = offsets[n] */
  const int64_t* offsets; /* [n+1], offsets[0] = 0, non-decreasing */
  const int32_t* cols;    /* [nnz], 0 <= col < n */
  const float* w;         /* [nnz] or NULL */
  const double* w64;      /* [nnz] or NULL */
} mgl_csr;

/* Options; MG_GCD flags -threshold, -max_phases, -max_iter, -version
 * (src/main.cu:20-26) plus build-only switches. */
typedef struct {
  int version;         /* 0 = plain atomic, 1 = global bi-probing, 2 = local bi-probing (default) */
  double threshold;    /* 1e-6 */
  int max_phases;      /* 1000 */
  int max_iter;        /* <= 0: unlimited (the reference ignores -max_iter) */
  int precision;       /* MGL_PREC_* */
  int tie_rule;        /* MGL_TIE_* */
  int fuse_modularity; /* 1 (default): sum_in of a proposal comes from the next move pass */
  int keep_levels;     /* 1: keep per-level assignments in the result */
  int keep_graphs;     /* 1: keep per-level coarse CSRs in the result */
  int validate;        /* 1 (default): check offsets, column range and symmetry */
  const char* tuning_file; /* bin thresholds written by mgl_autotune, or NULL for built-in defaults */
  int use_graph;       /* 1 (default): the inner loop of a level is captured once as a hipGraph and replayed */
  int record_layout;   /* MGL_LAYOUT_*: where the per-vertex records gathered by the move kernels live */
} mgl_opts;

/* Record layout of a level (results are identical in every layout). */
enum {
  MGL_LAYOUT_AUTO = 0,   /* degree-bin order when the level's largest graph has more than
                            2^19 vertices, else vertex order */
  MGL_LAYOUT_VERTEX = 1, /* record of v at position v */
  MGL_LAYOUT_BINNED = 2  /* records in degree-bin order: hub records share cache lines */
};

typedef struct mgl_ctx mgl_ctx;
typedef struct mgl_result mgl_result;
typedef struct mgl_session mgl_session;

void mgl_opts_default(mgl_opts* o);

/* Replaces the MPI/NVSHMEM bootstrap and per-node device selection
 * (src/main.cu:30-109): one context per device. */
mgl_status mgl_create(mgl_ctx** ctx, int device);
void mgl_destroy(mgl_ctx* ctx);
const char* mgl_last_error(const mgl_ctx* ctx);

/* Full multi-level Louvain on one graph. Replaces louvain::run /
 * louvain_gl::run / louvain_baseline::run (src/louvain/louvain.cu:1401-1715,
 * louvain_gl.cu:1387-1705, louvain_baseline.cu:1374-1682), which mutate the
 * caller's graph and print only; here the input is borrowed and the levels,
 * their modularity, assignments and coarse graphs are returned. */
mgl_status mgl_louvain(mgl_ctx* ctx, const mgl_csr* g, const mgl_opts* o, mgl_result** out);

/* Many independent graphs in one launch sequence (no reference counterpart:
 * the reference runs one graph per process). results[i] belongs to graphs[i];
 * each equals what mgl_louvain returns for that graph alone. */
mgl_status mgl_louvain_batched(mgl_ctx* ctx, const mgl_csr* graphs, int32_t n_graphs,
                               const mgl_opts* o, mgl_result** results);

/* Result accessors. Pointers stay valid until mgl_result_free. */
int mgl_result_levels(const mgl_result* r);
int64_t mgl_result_level_n(const mgl_result* r, int level);
double mgl_result_level_q(const mgl_result* r, int level);
int mgl_result_level_iters(const mgl_result* r, int level);
const double* mgl_result_level_qtraj(const mgl_result* r, int level); /* [iters] */
/* Modularity of the level's singleton start: the reference's "Loop 0" row
 * (louvain.cu:1509-1538). */
double mgl_result_level_q0(const mgl_result* r, int level);
/* Device time of each inner iteration of the level in seconds [iters]: the
 * reference's per-loop time column (louvain.cu:1650-1655), taken from the
 * device clock at each decision (in a batch, the iteration of the union). */
const double* mgl_result_level_iter_seconds(const mgl_result* r, int level);
const int32_t* mgl_result_level_assign(const mgl_result* r, int level); /* NULL unless keep_levels */
/* Coarse graph that is the INPUT of `level` (level >= 1); weights as doubles in
 * the graph's units (exact mode: integers or dyadic values). NULL unless keep_graphs. */
mgl_status mgl_result_level_graph(const mgl_result* r, int level, int64_t* n, int64_t* nnz,
                                  const int64_t** offsets, const int32_t** cols,
                                  const double** w);
/* Wall time of the level's inner loop; *coarsen_seconds (optional) receives
 * the time of the level's end (bookkeeping and coarsening). */
double mgl_result_level_seconds(const mgl_result* r, int level, double* coarsen_seconds);
const int32_t* mgl_result_final_assign(const mgl_result* r); /* dense labels of level-0 vertices */
int64_t mgl_result_vertices(const mgl_result* r); /* level-0 vertices (length of the final assignment) */
double mgl_result_final_q(const mgl_result* r);
int mgl_result_precision(const mgl_result* r); /* precision actually used */
double mgl_result_seconds(const mgl_result* r); /* device wall time of the run */
void mgl_result_free(mgl_result* r);

/* ---- Device-resident session: one graph kept in HBM, driven pass by pass.
 * Used to time the hot path with inputs already resident, and to check each
 * stage on its own. */
typedef struct {
  int64_t n, nnz;
  double move_ms;      /* last pass: move kernels */
  double update_ms;    /* last pass: community-weight update */
  double modularity_ms;/* last pass: sum_in / sum a^2 reduction (+ E-pass when not fused) */
  double pass_ms;      /* last pass: whole inner-loop body incl. commit */
  double dominant_ms;  /* last pass: duration of the longest single-kernel move launch
                          among those that ran alone (not beside the hub path) */
  int dominant_bin;    /* bin id of that launch */
  int64_t dominant_vertices, dominant_entries; /* vertices / CSR entries that launch processed */
  int64_t moved;       /* vertices whose proposal differs from their community */
  int64_t distinct;    /* (vertex, neighbour community) pairs aggregated in the move */
  int64_t dominant_distinct; /* the same, for the dominant launch */
  double q;            /* committed modularity */
  char dominant_kernel[96]; /* name of the dominant launch's kernel (as profilers print it) */
  double hub_ms;       /* the hub path (rows beyond the LDS-table bins: several kernels), or 0 */
  int64_t hub_vertices, hub_entries, hub_distinct;
  /* LDS hash tables of the move pass: key inserts, compare-and-swap probes, and
   * inserts that needed more than one probe (first-probe hit rate = 1 - multi/inserts) */
  int64_t lds_inserts, lds_probes, lds_multi_probe;
  /* last pass: duration of each bin's move launch (the hub bin: its whole path), 0 if empty */
  int n_bins;
  double bin_ms[16];
  /* last pass: bins 1..hub_beside ran beside the hub path on the second stream
   * (their bin_ms span includes the time they shared the device), 0 if none */
  int hub_beside;
} mgl_pass_stats;

mgl_status mgl_session_create(mgl_ctx* ctx, const mgl_csr* g, const mgl_opts* o,
                              mgl_session** s);
void mgl_session_destroy(mgl_session* s);
/* Singleton state of level 0: c = identity, a = k, Q computed
 * (louvain.cu:1492-1541). */
mgl_status mgl_session_reset(mgl_session* s, double* q0);
/* One inner-loop iteration of louvain.cu:1548-1660: move, weight update,
 * modularity, commit-or-reject. *accepted = 1 when committed. */
mgl_status mgl_session_pass(mgl_session* s, double* q, int* accepted);
/* Counts what the next mgl_session_pass's move aggregates, without timing it
 * or changing the state it starts from: distinct (vertex, community) pairs per
 * bin and the LDS hash-probe statistics, reported in that pass's statistics.
 * No counter runs inside a timed pass (counting takes global atomics). */
mgl_status mgl_session_count(mgl_session* s);
mgl_status mgl_session_stats(const mgl_session* s, mgl_pass_stats* st);
/* Stage entry points on the session's level-0 graph. */
/* Replaces BIN::bin_create (src/bin/BIN.cu:468-499): returns the bin sizes. */
mgl_status mgl_session_bins(mgl_session* s, int64_t* bin_sizes, int* n_bins);
/* Sets the committed assignment c (host array [n]) and recomputes a. */
mgl_status mgl_session_set_assign(mgl_session* s, const int32_t* c);
/* Replaces calculate_eicj_and_move_vertex_bin (louvain.cu:1032-1296):
 * proposal from the committed state into c_new[n]. */
mgl_status mgl_session_move(mgl_session* s, int up, int32_t* c_new);
/* Replaces compute_modularity (louvain.cu:105-238) for the committed c. */
mgl_status mgl_session_modularity(mgl_session* s, double* q);
/* Replaces compute_modularity (louvain.cu:105-238) for any partition of a
 * graph that is already in device memory: d_offsets [n+1], d_cols [nnz],
 * d_w [nnz] float weights or NULL (unit weights: exact integer sums, Q bit-equal
 * to the exact modes'), d_assign [n] community ids in [0, n). Device pointers
 * of the context's device; runs on the context's stream and returns once *q
 * is set. Out-of-range ids, columns or rows are skipped on the device and
 * reported as MGL_INVALID_GRAPH. The community totals and sums live in device
 * memory this call allocates and frees. */
mgl_status mgl_modularity_dev(mgl_ctx* ctx, int64_t n, int64_t nnz, const int64_t* d_offsets,
                              const int32_t* d_cols, const float* d_w, const int32_t* d_assign,
                              double* q);
/* The same as mgl_modularity_dev without any device allocation. No
 * counterpart in the reference: for callers that cannot allocate (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc). Its
 * working memory is a fixed array in device memory, one per device, held for
 * the life of the process (64 MiB + 256 B): graphs of at most
 * MGL_NOALLOC_MAX_VERTICES vertices, larger ones are refused with
 * MGL_INVALID_ARG before anything runs. The array is cleared at the start of
 * each call; calls that overlap on one device (from any context) race on it
 * and must be serialised by the caller. */
#define MGL_NOALLOC_MAX_VERTICES (1 << 23)
mgl_status mgl_modularity_dev_noalloc(mgl_ctx* ctx, int64_t n, int64_t nnz,
                                      const int64_t* d_offsets, const int32_t* d_cols,
                                      const float* d_w, const int32_t* d_assign, double* q);
/* Community totals a[n] (as doubles, in graph units) of the committed state. */
mgl_status mgl_session_get_weights(mgl_session* s, double* a, double* k);
/* Replaces compute_community_weight(_local_atomic) (louvain.cu:1299-1396,
 * louvain_baseline.cu:1319-1372): applies the moves c -> c_new (host [n]) to a,
 * then makes c_new the committed assignment. */
mgl_status mgl_session_apply(mgl_session* s, const int32_t* c_new);
/* Replaces coarsen_graph_mg::coarsen_graph (coarsen_graph_mg.cu:2054-2240):
 * coarsens by the committed c. Output buffers sized n+1 / nnz by the caller;
 * rank[n] receives the dense rank of each community id (-1 if unused). */
mgl_status mgl_session_coarsen(mgl_session* s, int64_t* n2, int64_t* nnz2, int64_t* off2,
                               int32_t* col2, double* w2, int32_t* rank);

/* ---- Synthetic graphs (host, multithreaded). Output CSR is symmetric, without
 * self-loops or duplicates, rows sorted; free with mgl_csr_free. */
typedef struct {
  int64_t n, nnz;
  int64_t* offsets;
  int32_t* cols;
  float* w;     /* NULL for unit weights (generators: float weights) */
  double* w64;  /* NULL, or the weights as parsed doubles (mgl_read_mtx) */
} mgl_host_csr;
/* weights: 0 unit, 1 uniform (0.5,1.5] (1.5 - U[0,1) in float), 2 lognormal(0,1) as float */
mgl_status mgl_gen_rmat(int scale, double edge_factor, double a, double b, double c,
                        uint64_t seed, int weights, int64_t n_limit, mgl_host_csr* out);
/* Planted partition: n vertices in n_comm communities whose sizes follow a
 * power law (exponent, min size), mean intra / inter degree. */
mgl_status mgl_gen_sbm(int64_t n, int64_t n_comm, double exponent, int64_t min_size,
                       double deg_in, double deg_out, uint64_t seed, int weights,
                       mgl_host_csr* out);
/* Reads a Matrix Market file with the reference loader's semantics
 * (src/mmio/mmio_wrapper.cu:76-295): "symmetric" banner = unit weights, every
 * off-diagonal entry mirrored, diagonal kept once; "general" banner = only
 * entries with col < row kept, weights parsed as doubles (1.0 if absent; the
 * reference's atof into weight_t, :243-250) into w64, mirrored; w is NULL.
 * Rows are sorted by column (the reference's order within a row is not fixed). */
mgl_status mgl_read_mtx(const char* path, mgl_host_csr* out);
/* Writes assignment / CSR files (text). */
mgl_status mgl_write_assign(const char* path, const int32_t* assign, int64_t n);
/* Writes a result to files named from prefix (no reference counterpart: the
 * reference prints the per-level table only, louvain.cu:1536-1538,1654,1670-1703):
 *   prefix.levels.txt      level, vertices, iterations, Q, Q trajectory
 *   prefix.final.{txt,bin} composed assignment of the input vertices
 *   prefix.level<L>.{txt,bin}  the dendrogram: assignment of level L's vertices
 *                          (community ids of level L; level L+1's vertex of a
 *                          community is its rank among the used ids)   [keep_levels]
 *   prefix.level<L>.{mtx,csr} coarse graph that is the input of level L >= 1:
 *                          Matrix Market "coordinate real general" with every
 *                          entry, or binary CSR                          [keep_graphs]
 * binary = 0: text; 1: binary, little-endian: an 8-byte tag ("MGLASG1" or
 * "MGLCSR1" and a zero byte), int64 n (and int64 nnz), then int32 labels, or
 * int64 offsets[n+1], int32 cols[nnz], float64 weights[nnz]. */
mgl_status mgl_write_result(const mgl_result* r, const char* prefix, int binary);
void mgl_csr_free(mgl_host_csr* g);

/* Build-time / tuning helpers. */
int mgl_version_major(void);
/* Number of move bins and their upper degree bounds (inclusive). */
int mgl_bin_bounds(const mgl_ctx* ctx, int64_t* upper, int max_bins);
/* LDS bytes one workgroup may use on the context's device (sizes the tables). */
int mgl_device_lds_bytes(const mgl_ctx* ctx);
/* Device memory the library's calls hold now and held at most (process-wide,
 * all devices), in bytes; reset_peak = 1 restarts the peak from the current
 * value. */
mgl_status mgl_memory_stats(int64_t* live_bytes, int64_t* peak_bytes, int reset_peak);
/* Copies `bytes` between two device buffers `iters` times with a streaming
 * copy kernel and returns the best read+write GB/s (HBM calibration). */
mgl_status mgl_membench(mgl_ctx* ctx, int64_t bytes, int iters, double* gbps);
/* Times the move pass of g under candidate bin layouts and hub-path shapes and
 * writes the fastest to path as a tuning file (text: '#' comment lines,
 * "hub_agg_threads N", "hub_agg_per_cu N", "hub_bucket N", then one bin per
 * line, "upper shape lanes_or_threads size"). A malformed tuning file makes
 * the call that names it fail with MGL_INVALID_ARG. */
mgl_status mgl_autotune(mgl_ctx* ctx, const mgl_csr* g, const char* path);

#ifdef __cplusplus
}
#endif
#endif /* MGL_H */
