// Host-side declarations of the kernel launchers (mgl_kernels.hip, mgl_coarsen.hip).
#pragma once
#include "mgl_common.h"

// ---- move-bin layout -------------------------------------------------------
// A bin holds vertices whose degree lies in (upper[b-1], upper[b]]. Bin 0 is
// degree 0 (isolated vertices never move and are never launched). Each bin is
// served by one kernel shape:
//   SHAPE_GROUP : S lanes of a wavefront per vertex, LDS hash of 2^lg slots each
//   SHAPE_BLOCK : one workgroup of BS threads per vertex, LDS hash of `size` slots
//   SHAPE_GLOBAL: hub path (mgl_hub.hip): rows split into chunks, candidates
//                 bucketed by hash, one workgroup per bucket
//   SHAPE_LANE  : one lane per vertex, the row merged in registers (degree <= 16)
enum Shape { SHAPE_NONE = 0, SHAPE_GROUP = 1, SHAPE_BLOCK = 2, SHAPE_GLOBAL = 3, SHAPE_LANE = 4 };
struct BinSpec {
  int64_t upper;  // inclusive degree bound (INT64_MAX for the last bin)
  int shape;
  int s;          // lanes per vertex (GROUP) or threads per block (BLOCK)
  int size;       // GROUP: log2 of slots per vertex; BLOCK: slots per vertex
};
#define MGL_MAX_BINS 16

struct BinLayout {
  int nb = 0;
  BinSpec b[MGL_MAX_BINS];
};
// Built-in layout for a device whose workgroups may use lds_bytes of LDS and a
// mode whose hash slots take slot_bytes (+4 for the key).
BinLayout default_bins(int lds_bytes, int slot_bytes);

// ---- move ------------------------------------------------------------------
struct MoveArgs {
  const eid* off;     // row offsets by record position (the row of v at off[rpos[v]]; by
                      // vertex id when rpos is null)
  const vid* col;     // columns as record positions (rpos[u] when rpos is set)
  const void* w;      // weights of those entries
  const void* k;      // K[n] vertex totals, by record position like off
  const vid* c;       // assignment the moves start from
  const void* a;      // K[n] community totals of c
  const void* rec;    // Rec<AT>[n]: {c[u], a[c[u]]} of the assignment the moves start from,
                      // stored at position rpos[u]
  const vid* rpos;    // [n] record position of each vertex (nullable = identity)
  vid* c_new;         // proposal
  void* vin;          // K[n]: per-vertex weight to own community incl. self-loop (nullable)
  const int32_t* gid; // per-vertex graph id (nullable = graph 0)
  GState* gs;
  int up;
  int version;
  int tie_small;
  unsigned long long* stats; // [0] distinct pairs (nullable)
  unsigned long long* probes; // [3] LDS hash inserts, probes, multi-probe inserts (nullable)
  int64_t pos0;       // >= 0: perm[i] has record position pos0 + i (records in degree-bin
                      // order, bins not compacted); < 0: look the position up in rpos
  int* err;           // set to 1 when an LDS table insert finds no free slot (nullable)
};

// Launch the move kernel of one GROUP or BLOCK bin. perm points at the bin's
// vertices. Returns 0, or -1 when no kernel instance matches the bin's shape.
int launch_move_bin(int mode, int rec32, const BinSpec& bs, const MoveArgs& A, const vid* perm,
                    int64_t nb, hipStream_t st);
// Name of the kernel launch_move_bin would launch for a bin (as profilers
// print it); cnt names the counting variant.
int move_kernel_name(int mode, int rec32, const BinSpec& bs, char* buf, int len,
                     bool cnt = false);
// Per-mode instantiation units of the move kernels (mgl_move_*.hip): launch, or
// with name != nullptr only name, the kernel of a bin; -1 for a mode they lack.
#define MGL_MOVE_UNIT_DECL(tag)                                                          \
  int move_bin_##tag(int mode, int rec32, bool cnt, const BinSpec& bs, const MoveArgs& A, \
                     const vid* perm, int64_t nb, hipStream_t st, char* name, int len);   \
  void move_allow_lds_##tag(int bytes);
MGL_MOVE_UNIT_DECL(unit)
MGL_MOVE_UNIT_DECL(i32)
MGL_MOVE_UNIT_DECL(int)
MGL_MOVE_UNIT_DECL(f32)
#undef MGL_MOVE_UNIT_DECL
// rec[rpos[v]] = {c[v], a[c[v]]}; rec32 selects 8-byte records (unit mode, 2m < 2^31)
// list (optional): the n vertex ids to process instead of 0..n-1
// rpos / rinv (optional): record position of each vertex and its inverse; with
// rinv and no list, position i is written from vertex rinv[i] (sequential writes)
// first: items before it are skipped (positions of isolated vertices, whose
// records no move kernel reads)
void launch_build_rec(int mode, int rec32, int64_t n, const vid* c, const void* a, void* rec,
                      hipStream_t st, const vid* list = nullptr, const vid* rpos = nullptr,
                      const vid* rinv = nullptr, int64_t first = 0);
// Record layout of a level: rpos[perm[i]] = i and rinv[i] = perm[i] for i < n;
// colr[e] = rpos[col[e]] for e < nnz.
void launch_record_layout(const vid* perm, int64_t n, vid* rpos, vid* rinv, hipStream_t st);
// Rows of a level in record-position order (records in degree-bin order):
// deg[0..n] (degree of the vertex at each position, then 0), whose exclusive
// scan gives offr; then the rows of positions p0 .. p0+np-1 (one bin; upper =
// its degree bound picks the lanes per row) are placed at offr with columns
// as record positions and their weights (wb bytes each, 0: none).
void launch_pos_degrees(const eid* off, const vid* rinv, int64_t n, int64_t* deg,
                        hipStream_t st);
void launch_place_rows(int64_t upper, const eid* off, const vid* col, const void* w, int wb,
                       const vid* rinv, const vid* rpos, const eid* offr, int64_t p0, int64_t np,
                       vid* colr, void* wr, hipStream_t st);
void move_kernels_allow_lds(int bytes);

// ---- hub path (mgl_hub.hip) ------------------------------------------------
// One bucket of the segmented hub path, fixed for a level (built on the host
// with the plan): its hub, and where that hub's candidates and bucket offsets
// are. Chunk c of the hub starts at tmp position base + c * hub_chunk_edges().
struct HubBucket {
  int64_t base;  // tmp position of the hub's first chunk
  int64_t cb0;   // first cbo entry of the hub
  int32_t hub;   // hub index
  vid v;         // hub vertex
  int32_t lb;    // bucket index within the hub
  int32_t P;     // buckets of the hub
  int32_t nch;   // chunks of the hub
  int32_t g;     // graph of the hub (batches)
};
// One chunk of a hub's row, fixed for a level: CSR entries [e0, e1), their
// compact position in tmp, the chunk's bucket offsets in cbo (segmented form)
// and its hub.
struct HubChunk {
  int64_t e0, e1;
  int64_t pos0;  // tmp position of entry e0
  int64_t cq;    // first of the chunk's P + 1 entries in cbo
  int32_t hub;   // hub index
  int32_t P;     // buckets of the hub
  vid v;         // hub vertex
  int32_t g;     // graph of the hub
};
// Per hub and pass (k_hub_head): the inputs of its gains, so that the hub
// kernels load them in one step instead of a chain of dependent loads. ki and
// acik hold the bits of the mode's total type (int64 or double).
struct HubGain {
  int64_t ki;    // the hub's total
  int64_t acik;  // its community's total without the hub
  vid ci;        // its community
  int32_t pad;
};
struct HubPlan {
  int64_t m = 0, nchunks = 0, nbuckets = 0, ehub = 0;
  const HubChunk* chunks = nullptr; // [nchunks]
  HubGain* gin = nullptr;           // [m]
  const vid* hub_v = nullptr;       // [m]
  const int64_t* hub_bk = nullptr;  // [m+1] bucket ranges
  const int64_t* bk_hub = nullptr;  // [nbuckets] owning hub
  vid* tmp_c = nullptr;             // [ehub] neighbour community or EMPTY
  void* tmp_a = nullptr;            // [ehub] its community total
  void* bkt_a = nullptr;            // [ehub] bucketed totals
  void* tmp_w = nullptr;            // [ehub] weights (weighted modes)
  vid* bkt_c = nullptr;             // [ehub] bucketed candidates
  void* bkt_w = nullptr;
  int64_t* bcount = nullptr;        // [nbuckets]
  int64_t* boff = nullptr;          // [nbuckets+1]
  int64_t* bcur = nullptr;          // [nbuckets]
  void* hub_eici = nullptr;         // K[m]
  void* hub_self = nullptr;         // K[m]
  double* bgain = nullptr;          // [nbuckets]
  vid* bdst = nullptr;              // [nbuckets]
  vid* gk = nullptr;                // [2*ehub] HBM tables for oversized buckets
  void* gv = nullptr;
  void* ga = nullptr;
  void* scan_tmp = nullptr;
  uint32_t lds_slots = 0;           // LDS table slots of the aggregate kernel
  // Segmented form (every hub of at most hub_seg_max() chunks): each chunk's
  // candidates stay in its region of tmp, grouped by bucket, and cbo holds the
  // chunk's P_i + 1 bucket offsets; the aggregate reads a bucket's segments
  // from every chunk of its hub, so no scatter pass and no global scan.
  int segmented = 0;
  int32_t* cbo = nullptr;               // [sum over hubs of chunks x (P_i + 1)]
  unsigned long long* galloc = nullptr; // [1] HBM-table slots handed out in this pass
  const HubBucket* bmeta = nullptr;     // [nbuckets] (segmented form)
  int agg_bs = 1024;                    // threads per aggregate workgroup: 256, 512 or 1024
  int count_bs = 256;                   // threads per count workgroup (segmented form)
};
int launch_hub(int mode, int rec32, const MoveArgs& A, const HubPlan& H, hipStream_t st);
void hub_allow_lds(int bytes);
int hub_chunk_edges();
int hub_max_buckets();
int hub_seg_max();  // chunks per hub of the segmented form
int hub_agg_cap();  // candidates of one bucket on the aggregate's owner path
void launch_row_ranges(const eid* off, const vid* perm, int64_t nb, int64_t* out, hipStream_t st);
// out[i] = src[idx[i]] for i < n
void launch_gather_i32(const int32_t* src, const vid* idx, int64_t n, int32_t* out, hipStream_t st);
// out[i] = src[idx[i]] for 8-byte values
void launch_gather_u64(const void* src, const vid* idx, int64_t n, void* out, hipStream_t st);

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

// ---- community-weight update, reductions, control --------------------------
void launch_update(int mode, int64_t n, const vid* c, const vid* cn, const void* k, void* a,
                   const int32_t* gid, GState* gs, hipStream_t st, const vid* list = nullptr);
// Adds sum(vin) (if vin) and sum(a^2) into each graph's accumulators.
// with a list, its ids must be increasing (graphs stay contiguous)
void launch_reduce(int mode, int64_t n, const void* vin, const void* a, const int32_t* gid,
                   GState* gs, hipStream_t st, const vid* list = nullptr);
// Level start: active = run[g] and cleared accumulators; then, after the
// singleton reduction, q_cur = Q(singletons).
void launch_level_begin(int ng, GState* gs, const int32_t* run, hipStream_t st);
void launch_level_init(int ng, GState* gs, int exact, hipStream_t st);
// nactive (optional) receives the number of graphs still active afterwards;
// ttraj (optional, [ng * traj_max]) the device clock ticks of each iteration.
void launch_decide(int ng, GState* gs, double thr, int max_iter, int fused, int exact,
                   double* traj, int traj_max, int pre, hipStream_t st, int* nactive = nullptr,
                   int64_t* ttraj = nullptr);
void launch_fixup(int64_t n, vid* c, vid* p, vid* nx, const int32_t* gid, const GState* gs,
                  hipStream_t st, const vid* list = nullptr);
// active-set compaction of a batch
void launch_filter_bin(const vid* perm, int64_t nb, const int32_t* gid, const GState* gs,
                       vid* out, unsigned long long* cnt, hipStream_t st);
void launch_active_list(int64_t n, const eid* off, const int32_t* gid, const GState* gs,
                        int32_t* flag, int64_t* pos, vid* out, void* scan_tmp, hipStream_t st);
void launch_iota(vid* x, int64_t n, hipStream_t st);
void launch_copy_k_to_a(int mode, int64_t n, const void* k, void* a, hipStream_t st);
void launch_copy_vid(const vid* src, vid* dst, int64_t n, hipStream_t st);
void launch_fill_bytes0(void* p, size_t bytes, hipStream_t st);
void launch_cast_weights(int to_mode, const float* w, void* out, int64_t nnz, int scale,
                         hipStream_t st);
void launch_cast_weights_f64(int to_mode, const double* w, void* out, int64_t nnz, int scale,
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
  // per-vertex arrays sized by the level-0 graph, per-entry arrays by cap
  int32_t* flag;      // [n]
  int64_t* rank;      // [n+1]
  vid* newc;          // [n]
  uint64_t* k0; uint64_t* k1;  // [cap] keys ping-pong
  void* v0; void* v1;          // [cap] values ping-pong (8 B each)
  int64_t* hist;      // radix histograms
  int32_t* head;      // [cap]
  int64_t* uidx;      // [cap+1]
  int64_t* rowcnt;    // [n+1]
  void* scan_tmp;
  int64_t hist_cap;
  int64_t cap;        // kept entries one group may hold
};
// Coarsens the level graph by committed assignment c. Vertices of graphs with
// keep[g] == 0 are dropped. ebase[g] = first compacted entry of graph g.
// Writes the coarse CSR into (off2, col2, w2) and returns C, E2 through host
// pointers (one synchronisation).
int coarsen_level(int mode_in, int mode_out, int64_t n, const eid* off, const vid* col,
                  const void* w, const vid* c, const int32_t* gid, const int32_t* keep,
                  const int64_t* ebase, int64_t e_keep, CoarsenBufs& B, eid* off2, vid* col2,
                  void* w2, int64_t* C_out, int64_t* E2_out, hipStream_t st);
// The same in two steps, so that a batch can be coarsened in groups of
// consecutive graphs whose kept entries fit B.cap (each graph's entries stay in
// its graph): the renumbering of the whole level (rank[0..n], rank[n] = C,
// newc), then for each group of fine rows [v_lo, v_hi) and entries [e_lo,
// e_hi), e_keep of them kept (compacted from kbase on), its coarse rows
// [r_lo, r_hi) = [rank[v_lo], rank[v_hi]) with entries written from U0 on
// (*U_out of them). off2[r_hi] is written by the group.
int coarsen_renumber(int64_t n, const vid* c, const int32_t* gid, const int32_t* keep,
                     CoarsenBufs& B, int64_t* C_out, hipStream_t st);
int coarsen_group(int mode_in, int mode_out, int64_t v_lo, int64_t v_hi, int64_t e_lo,
                  int64_t e_hi, const eid* off, const vid* col, const void* w, const int32_t* gid,
                  const int64_t* delta, int64_t kbase, int64_t e_keep, int64_t r_lo,
                  int64_t r_hi, int64_t U0, CoarsenBufs& B, eid* off2, vid* col2, void* w2,
                  int64_t* U_out, hipStream_t st);
