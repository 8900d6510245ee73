// Host-side declarations of the kernel launchers (mgl_kernels.hip, mgl_coarsen.hip).
#pragma once
#include "mgl_common.h"

// ---- move-bin layout -------------------------------------------------------
// A bin holds vertices whose degree lies in (upper[b-1], upper[b]]. Bin 0 is
// degree 0 (isolated vertices never move and are never launched). Each bin is
// served by one kernel shape:
//   SHAPE_GROUP : S lanes of a wavefront per vertex, LDS hash of H slots each
//   SHAPE_BLOCK : one workgroup of BS threads per vertex, LDS hash of H slots
//   SHAPE_GLOBAL: one workgroup of 1024 threads per vertex, hash in HBM sized
//                 2*nextpow2(degree) per vertex
enum Shape { SHAPE_NONE = 0, SHAPE_GROUP = 1, SHAPE_BLOCK = 2, SHAPE_GLOBAL = 3 };
struct BinSpec {
  int64_t upper;  // inclusive degree bound (INT64_MAX for the last bin)
  int shape;
  int s;          // lanes per vertex (GROUP) or threads per block (BLOCK/GLOBAL)
  int lg;         // log2 of hash slots per vertex (GROUP/BLOCK)
};
#define MGL_MAX_BINS 16

struct BinLayout {
  int nb = 0;
  BinSpec b[MGL_MAX_BINS];
};
BinLayout default_bins();

// ---- move ------------------------------------------------------------------
struct MoveArgs {
  const eid* off;
  const vid* col;
  const void* w;
  const void* k;      // K[n]
  const vid* c;       // assignment the moves start from
  const void* a;      // K[n] community totals of c
  vid* c_new;         // proposal
  void* vin;          // K[n]: per-vertex weight to own community incl. self-loop (nullable)
  const int32_t* gid; // per-vertex graph id (nullable = graph 0)
  GState* gs;
  int up;
  int version;
  int tie_small;
  unsigned long long* stats; // [0] distinct pairs (nullable)
};

// Launch the move kernel of one bin. perm points at the bin's vertices.
// For SHAPE_GLOBAL, toff[i] is the slot offset of the i-th vertex's table
// (toff[nb] = total) and keys/vals point at scratch of that many slots.
// Returns 0, or -1 when no kernel instance matches the bin's shape.
int launch_move_bin(int mode, const BinSpec& bs, const MoveArgs& A, const vid* perm,
                    int64_t nb, const int64_t* toff, vid* gkeys, void* gvals,
                    hipStream_t st);
void move_kernels_allow_lds(int bytes);

// ---- per-vertex row reductions (vertex weights, modularity sum_in) ----------
struct RowArgs {
  const eid* off;
  const vid* col;
  const void* w;
  const vid* c;       // for OP_INTRA
  void* k;            // OP_WEIGHT: K[n] row sums
  void* vin;          // K[n]: OP_WEIGHT -> self-loop weight; OP_INTRA -> intra weight
  const int32_t* gid;
  const GState* gs;   // OP_INTRA skips inactive graphs when non-null
};
enum { OP_WEIGHT = 0, OP_INTRA = 1 };
void launch_rows_bin(int mode, int op, const BinSpec& bs, const RowArgs& A, const vid* perm,
                     int64_t nb, hipStream_t st);

// ---- binning ---------------------------------------------------------------
// counts[nbins] += histogram of degree bins; perm filled per bin using cursor[].
void launch_bin_count(const eid* off, int64_t n, const BinLayout& L, unsigned long long* counts,
                      hipStream_t st);
void launch_bin_scatter(const eid* off, int64_t n, const BinLayout& L,
                        unsigned long long* cursor /* [nb] starts at bin offsets */, vid* perm,
                        hipStream_t st);
void launch_table_sizes(const eid* off, const vid* perm, int64_t nb, int64_t* sizes,
                        hipStream_t st);

// ---- community-weight update, reductions, control --------------------------
void launch_update(int mode, int64_t n, const vid* c, const vid* cn, const void* k, void* a,
                   const int32_t* gid, GState* gs, hipStream_t st);
// Adds sum(vin) (if vin) and sum(a^2) into each graph's accumulators.
void launch_reduce(int mode, int64_t n, const void* vin, const void* a, const int32_t* gid,
                   GState* gs, hipStream_t st);
// Level start: active = run[g] and cleared accumulators; then, after the
// singleton reduction, q_cur = Q(singletons).
void launch_level_begin(int ng, GState* gs, const int32_t* run, hipStream_t st);
void launch_level_init(int ng, GState* gs, int exact, hipStream_t st);
void launch_decide(int ng, GState* gs, double thr, int max_iter, int fused, int exact,
                   double* traj, int traj_max, int pre, hipStream_t st);
void launch_fixup(int64_t n, vid* c, vid* p, vid* nx, const int32_t* gid, const GState* gs,
                  hipStream_t st);
void launch_iota(vid* x, int64_t n, hipStream_t st);
void launch_copy_k_to_a(int mode, int64_t n, const void* k, void* a, hipStream_t st);
void launch_copy_vid(const vid* src, vid* dst, int64_t n, hipStream_t st);
void launch_fill_bytes0(void* p, size_t bytes, hipStream_t st);
void launch_cast_weights(int to_mode, const float* w, void* out, int64_t nnz, int scale,
                         hipStream_t st);
void launch_gid_fill(const int64_t* voff, int ng, int32_t* gid, hipStream_t st);

// ---- coarsening (mgl_coarsen.hip) ------------------------------------------
struct DevScan {
  // exclusive scan of n values (int32 or int64 input) into int64 output
  static size_t temp_bytes(int64_t n);
};
void scan_exclusive_i32(const int32_t* in, int64_t* out, int64_t n, void* temp, hipStream_t st);
void scan_exclusive_i64(const int64_t* in, int64_t* out, int64_t n, void* temp, hipStream_t st);

struct CoarsenBufs {
  // all sized by the level-0 graph
  int32_t* flag;      // [n]
  int64_t* rank;      // [n+1]
  vid* newc;          // [n]
  uint64_t* k0; uint64_t* k1;  // [nnz] keys ping-pong
  void* v0; void* v1;          // [nnz] values ping-pong (8 B each)
  int64_t* hist;      // radix histograms
  int32_t* head;      // [nnz]
  int64_t* uidx;      // [nnz+1]
  int64_t* rowcnt;    // [n+1]
  void* scan_tmp;
  int64_t hist_cap;
};
// Coarsens the level graph by committed assignment c. Vertices of graphs with
// keep[g] == 0 are dropped. ebase[g] = first compacted entry of graph g.
// Writes the coarse CSR into (off2, col2, w2) and returns C, E2 through host
// pointers (one synchronisation).
int coarsen_level(int mode_in, int mode_out, int64_t n, const eid* off, const vid* col,
                  const void* w, const vid* c, const int32_t* gid, const int32_t* keep,
                  const int64_t* ebase, int64_t e_keep, CoarsenBufs& B, eid* off2, vid* col2,
                  void* w2, int64_t* C_out, int64_t* E2_out, hipStream_t st);
