// Host engine of the MI355X Louvain path: level loop, device-side inner-loop
// control, coarsening between levels, batching, sessions and the C ABI.
//
// Reference driver: src/louvain/louvain.cu:1401-1715 (louvain::run), the same
// loop in louvain_gl.cu:1387-1705 and louvain_baseline.cu:1374-1682.
//   phase loop   : while phase < max_phases && Q - Q_old > thr   (:1473)
//   level start  : c = id, k = row sums, a = k, Q(singletons)    (:1492-1533)
//   inner loop   : while new_Q - cur_Q > thr: move, flip up/down, update a,
//                  Q(c_new), commit if the gain exceeds thr, else reject (:1548-1660)
//   level end    : stop if the level gained <= thr, else coarsen (:1674-1679)
//
// Differences of form (results are the same):
//   * Every decision of the inner loop is taken on the device (one thread per
//     graph), so a batch of graphs runs the same launch sequence and the host
//     only polls an "any graph active" word.
//   * Fused modularity (default): sum_in of proposal p_t is produced by the
//     move pass that computes p_{t+1}, which reads the same c[u] gathers. One
//     CSR pass per iteration instead of two; the extra move of the final,
//     rejected iteration replaces the reference's modularity pass.
#include "mgl_kernels.h"
#include "../../include/mgl.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

// ---------------------------------------------------------------------------
// Bin layout and hub-path settings of a run: the built-in ones, or a tuning
// file's (mgl_opts.tuning_file), which applies to the call or session that
// names it only.
struct Tuning {
  BinLayout bins;
  bool have_bins = false;     // false: built-in layout for the device
  int64_t hub_bucket = 3072;  // hub candidates per bucket on average (capped by the tables)
  bool hub_bucket_clamp = true; // false: hub_bucket as given (oversized buckets take the
                                // scanned-LDS or HBM-table fallback; used by tests)
  int hub_segmented = 1;      // 0: the hub path always scatters into buckets
  int hub_agg_threads = 512;  // threads per hub-aggregate workgroup (256, 512, 1024)
  int hub_agg_per_cu = 4;     // hub-aggregate workgroups sharing a CU's LDS (1..8)
  int hub_count_threads = 256; // threads per hub-count workgroup (256, 512, 1024) of the
                               // segmented form; every workgroup takes 8,192 entries
  bool table_check = true;    // false: LDS tables smaller than a bin's rows are accepted
                              // (an overflow then fails the call; used by tests)
  int64_t coarsen_group = 0;  // > 0: kept entries of one coarsening group (else the default)
  int hub_stream = 1;         // 1: the hub path runs on a second stream beside the other bins
  int64_t hub_fork_upper = 512; // the hub path starts once the bins of rows longer than this
                                // are launched (0: after every bin)
};

struct mgl_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  // second stream of the move pass (the hub path beside the other bins) and
  // the events that fork it from st and join it back
  hipStream_t st2 = nullptr;
  hipEvent_t evfork = nullptr, evjoin = nullptr;
  std::string err;
  int lds_max = 65536;
  double clock_hz = 1.0e8;  // rate of the device clock read by wall_clock64()
  Tuning tune;  // settings of calls without a tuning file (the autotuner varies them)
};

#define FAIL(ctx, code, ...)                                    \
  do {                                                          \
    char _b[512];                                               \
    snprintf(_b, sizeof(_b), __VA_ARGS__);                      \
    (ctx)->err = _b;                                            \
    return (code);                                              \
  } while (0)

#define HIPC(ctx, call)                                                                 \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess) {                                                             \
      (void)hipGetLastError();                                                          \
      FAIL(ctx, _e == hipErrorOutOfMemory ? MGL_OOM : MGL_DEVICE_ERROR, "%s: %s (%s:%d)", \
           #call, hipGetErrorString(_e), __FILE__, __LINE__);                           \
    }                                                                                   \
  } while (0)

// Device memory held by the engine's buffers, process-wide (mgl_memory_stats).
static std::atomic<int64_t> g_dev_live{0}, g_dev_peak{0};
static void dev_bytes_add(int64_t d) {
  const int64_t now = g_dev_live.fetch_add(d) + d;
  int64_t pk = g_dev_peak.load();
  while (now > pk && !g_dev_peak.compare_exchange_weak(pk, now)) {
  }
}

// Device buffer owned by the engine.
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) {
      (void)hipFree(p);
      dev_bytes_add(-(int64_t)bytes);
    }
    p = nullptr;
    bytes = 0;
  }
  hipError_t ensure(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    release();
    if (b == 0) b = 16;
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) {
      bytes = b;
      dev_bytes_add((int64_t)b);
    } else {
      p = nullptr;
    }
    return e;
  }
  template <class T> T* as() const { return (T*)p; }
};

// Pinned host buffer owned by the engine: the small reads of device state
// between launches (graph states, the error word, the active count) land here
// in one DMA each, without a pageable bounce buffer.
struct HBuf {
  void* p = nullptr;
  size_t bytes = 0;
  HBuf() = default;
  HBuf(const HBuf&) = delete;
  HBuf& operator=(const HBuf&) = delete;
  ~HBuf() {
    if (p) (void)hipHostFree(p);
  }
  hipError_t ensure(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipHostMalloc(&p, b, hipHostMallocDefault);
    if (e == hipSuccess) bytes = b;
    else p = nullptr;
    return e;
  }
};

static size_t wsize(int mode) {  // bytes of one stored edge weight
  switch (mode) {
    case MODE_UNIT: return 0;
    case MODE_INT: return 8;
    case MODE_I32: return 4;
    case MODE_F64: return 8;
    case MODE_F32: return 4;
  }
  return 8;
}
static bool mode_exact(int mode) { return mode_is_int(mode); }
// Weights of a coarse level: exact unit/integer levels whose graphs all have
// 2m < 2^31 (rec32) carry int32 weights, since every coarse entry, degree and
// community total is at most 2m; otherwise int64. Float modes keep theirs.
static int coarse_mode(int mode, bool rec32) {
  if (mode == MODE_UNIT || mode == MODE_I32) return rec32 ? MODE_I32 : MODE_INT;
  return mode;
}
static int slot_bytes(int mode) {  // hash-slot value (partial sum)
  switch (mode) {
    case MODE_UNIT: return 4;
    case MODE_INT: return 8;
    case MODE_I32: return 4;
    case MODE_F64: return 8;
    case MODE_F32: return 4;
  }
  return 8;
}
static int total_bytes(int mode, bool rec32) {  // community total carried in a record
  return ((mode == MODE_UNIT || mode == MODE_I32) && rec32) ? 4 : 8;
}
// LDS bytes of one table slot: key + partial sum + community total
static int slot_total_bytes(int mode, bool rec32) {
  return 4 + slot_bytes(mode) + total_bytes(mode, rec32);
}

// Largest prime <= x (x >= 2): BLOCK tables are prime-sized so that every
// double-hashing step is coprime with the table size.
static int prime_at_most(int x) {
  for (; x > 2; --x) {
    bool p = (x % 2) != 0;
    for (int d = 3; p && (int64_t)d * d <= x; d += 2) p = (x % d) != 0;
    if (p) return x;
  }
  return 2;
}

// Bin layout for a mode, with LDS tables shrunk to what the device allows.
// BLOCK bins of at most 16 entries per thread use the owner-based kernel
// (key + partial sum per slot); longer ones also keep the total in the slot.
static BinLayout layout_for(const mgl_ctx* ctx, const Tuning& tu, int mode, bool rec32) {
  const int sb = 4 + slot_bytes(mode);
  BinLayout L = tu.have_bins ? tu.bins : default_bins(ctx->lds_max, sb);
  for (int i = 0; i < L.nb; ++i) {
    BinSpec& b = L.b[i];
    if (b.shape == SHAPE_BLOCK) {
      const bool owner = (b.upper + b.s - 1) / b.s <= 16;
      const int64_t cap =
          (int64_t)(ctx->lds_max - 2048) / (owner ? sb : slot_total_bytes(mode, rec32));
      if (b.size > cap) b.size = (int)cap;
      b.size = prime_at_most(b.size);
      // every distinct community of a row needs a slot
      if (b.upper > b.size) b.upper = b.size;
    }
  }
  // keep bounds increasing after any clamp and end with the unbounded hub bin
  for (int i = 1; i < L.nb; ++i)
    if (L.b[i].upper < L.b[i - 1].upper) L.b[i].upper = L.b[i - 1].upper;
  L.b[L.nb - 1].upper = INT64_MAX;
  L.b[L.nb - 1].shape = SHAPE_GLOBAL;
  return L;
}

// ---------------------------------------------------------------------------
// One level's graph (a union of the batch's graphs still running).
struct LevelGraph {
  int64_t n = 0, nnz = 0;
  int mode = MODE_UNIT;
  DBuf off, col, w, gid;
  std::vector<int64_t> voff;  // host, per-graph vertex ranges [G+1]
  std::vector<int64_t> eoff;  // host, per-graph entry ranges [G+1]
};

struct LevelOut {
  int64_t n = 0;
  double q = 0;
  int iters = 0;
  double seconds = 0;        // inner loop of the level (wall clock)
  double coarsen_seconds = 0;
  std::vector<double> qtraj;
  double q0 = 0;                 // Q of the singleton start
  std::vector<double> iter_secs; // device time of each iteration
  std::vector<int32_t> assign;
  int64_t gn = 0, gnnz = 0;
  std::vector<int64_t> goff;
  std::vector<int32_t> gcol;
  std::vector<double> gw;
  bool has_graph = false;
};

struct mgl_result {
  int precision = 0;
  double seconds = 0;
  std::vector<LevelOut> levels;
  std::vector<int32_t> final_assign;
  double final_q = 0;
};

static const int TRAJ_MAX = 4096;
// Kept entries one coarsening group of a batch may hold: the sort's keys,
// values, heads and unique indices take about 44 bytes per entry, so a batch
// is coarsened in groups of consecutive graphs of at most this many entries
// (a larger single graph is a group of its own).
static const int64_t COARSEN_GROUP_ENTRIES = (int64_t)1 << 30;
// Levels with more vertices than this keep their records in degree-bin order
// (an 8-byte record array of 2^19 vertices fills 4 MB, one XCD's L2).
static const int64_t LAYOUT_AUTO_MIN_N = (int64_t)1 << 19;

// ---------------------------------------------------------------------------
// Engine: runs the multi-level loop over a union of G graphs.
struct Engine {
  mgl_ctx* ctx;
  mgl_opts o;
  int G = 1;
  int prec = MGL_PREC_EXACT;
  int scale = 0;          // dyadic scale of level-0 weights (MODE_INT)
  BinLayout L;            // for the current level's mode
  // level-0 sizes (capacity of all per-vertex / per-entry buffers)
  int64_t n0 = 0, nnz0 = 0;
  std::vector<int64_t> vbase0;  // per-graph level-0 vertex ranges [G+1]
  // per-vertex state
  DBuf bc, bp, bn, bk, ba, bvin, bperm, blab, bfin, bgid0, brec, bperm2, blist, bcnt2;
  int64_t nlist = -1;  // >= 0: per-vertex passes walk blist (active vertices of a batch)
  // record layout of the current level: rpos[v] = position of v's record,
  // rinv = its inverse; the move kernels read the level's rows placed in that
  // order (offr: offsets by position, colr: columns as record positions, wr:
  // their weights), so each bin's rows are contiguous; pdeg and pscan are
  // the scratch of offr's scan
  DBuf brpos, brinv, bcolr, boffr, bwr, bpdeg, bpscan, bkr;  // bkr: k by position
  bool rlay = false;
  bool rec32 = false;  // 8-byte records (unit mode, every graph's 2m < 2^31)
  vid* c = nullptr; vid* p = nullptr; vid* nx = nullptr;
  // per-graph
  DBuf bgs, brun, btraj, bttraj, bkeep, bdelta, bgather, bcounts;
  std::vector<GState> hgs;
  HBuf hpin;  // staging of the reads below
  // bins of the current level
  std::vector<int64_t> bin_n, bin_off;
  // hub plan of the current level (last bin)
  DBuf hchunk, hv, hbk, hbkhub, htmpc, htmpw, hbktc, hbktw, hbcount, hboff, hbcur, heici, hself,
      hbgain, hbdst, hgk, hgv, hscan, htmpa, hbkta, hga, hcbo, hgalloc, hbmeta, hgin;
  HubPlan hub;
  // coarsening buffers
  DBuf cflag, crank, cnewc, ck0, ck1, cv0, cv1, chist, chead, cuidx, crowcnt, cscan;
  CoarsenBufs CB{};
  // level graphs (ping-pong)
  LevelGraph lg[3];  // [0] = input (kept), [1]/[2] = coarse levels in turn
  int cur = 0;
  // host-side phase state per graph
  std::vector<double> Q_host, Q_old;
  std::vector<int> phase, running;
  std::vector<mgl_result*> res;
  // timing
  hipEvent_t ev[2 * MGL_MAX_BINS + 8];
  bool ev_ok = false;
  mgl_pass_stats stats{};
  bool timing = false;    // HIP events around each bin and stage
  bool counting = false;  // per-bin distinct pairs and LDS probe counters (global atomics)
  int shared_from = 0;    // last move: bins 1..shared_from ran beside the hub path (0: none)

  Tuning tune;  // this run's bin layout and hub settings

  explicit Engine(mgl_ctx* c_) : ctx(c_), tune(c_->tune) {
    for (auto& e : ev) e = nullptr;
  }
  ~Engine() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  hipStream_t st() const { return ctx->st; }
  bool rec32_now() { return rec32 && (LG().mode == MODE_UNIT || LG().mode == MODE_I32); }
  const vid* vlist() const { return nlist >= 0 ? blist.as<vid>() : nullptr; }
  int64_t vcount() { return nlist >= 0 ? nlist : LG().n; }
  // device word: graphs still active after the last decision
  int* nactive() { return (int*)(bcounts.as<unsigned long long>() + 3 * MGL_MAX_BINS); }
  // device word: set by a move kernel whose LDS table overflowed
  int* err_word() { return (int*)(bcounts.as<unsigned long long>() + 4 * MGL_MAX_BINS + 3); }
  mgl_status err_overflow() {
    FAIL(ctx, MGL_DEVICE_ERROR,
         "an LDS hash table of a move kernel overflowed (the bin layout's table is smaller "
         "than a row's distinct communities)");
  }
  mgl_status check_err() {
    HIPC(ctx, hpin.ensure(16));
    HIPC(ctx, hipMemcpyAsync(hpin.p, err_word(), 4, hipMemcpyDeviceToHost, st()));
    HIPC(ctx, hipStreamSynchronize(st()));
    int e;
    memcpy(&e, hpin.p, 4);
    if (e) return err_overflow();
    return MGL_OK;
  }
  LevelGraph& LG() { return lg[cur]; }

  mgl_status alloc_state() {
    const int64_t n = n0 > 0 ? n0 : 1;
    const int64_t e = nnz0 > 0 ? nnz0 : 1;
    HIPC(ctx, bc.ensure(4 * n));
    HIPC(ctx, bp.ensure(4 * n));
    HIPC(ctx, bn.ensure(4 * n));
    HIPC(ctx, bk.ensure(8 * n));
    HIPC(ctx, ba.ensure(8 * n));
    HIPC(ctx, bvin.ensure(8 * n));
    HIPC(ctx, bperm.ensure(4 * n));
    HIPC(ctx, brec.ensure(16 * n));
    if (G > 1) {
      HIPC(ctx, bperm2.ensure(4 * n));
      HIPC(ctx, blist.ensure(4 * n));
      HIPC(ctx, bcnt2.ensure(8 * MGL_MAX_BINS));
    }
    HIPC(ctx, blab.ensure(4 * n));
    HIPC(ctx, bfin.ensure(4 * n));
    HIPC(ctx, bgs.ensure(sizeof(GState) * G));
    HIPC(ctx, brun.ensure(4 * G));
    HIPC(ctx, bkeep.ensure(4 * G));
    HIPC(ctx, bdelta.ensure(8 * G));
    HIPC(ctx, bgather.ensure(8 * (G + 1) * 2));
    HIPC(ctx, btraj.ensure(sizeof(double) * (size_t)G * TRAJ_MAX));
    HIPC(ctx, bttraj.ensure(sizeof(int64_t) * (size_t)G * TRAJ_MAX));
    // [0, MAX) bin counts, [MAX, 2 MAX) scatter cursors, [2 MAX, 3 MAX) distinct
    // pairs per bin, [3 MAX] active graphs, [4 MAX, 4 MAX + 3) LDS probe counters,
    // [4 MAX + 3] LDS-table overflow flag
    HIPC(ctx, bcounts.ensure(8 * (MGL_MAX_BINS * 4 + 4)));
    launch_fill_bytes0(bcounts.p, 8 * (MGL_MAX_BINS * 4 + 4), st());
    c = bc.as<vid>(); p = bp.as<vid>(); nx = bn.as<vid>();
    if (!ev_ok) {
      // timing only (read after the pass's stream synchronisation): no
      // system-scope fence at each record, which would write back and
      // invalidate the caches between the move kernels of a pass
      for (auto& x : ev) HIPC(ctx, hipEventCreateWithFlags(&x, hipEventDisableSystemFence));
      ev_ok = true;
    }
    return MGL_OK;
  }

  int64_t max_graph_nnz0 = 0;  // entries of the largest input graph

  mgl_status alloc_coarsen() {
    const int64_t n = n0 > 0 ? n0 : 1;
    const int64_t grp = tune.coarsen_group > 0 ? tune.coarsen_group : COARSEN_GROUP_ENTRIES;
    const int64_t e = std::max<int64_t>(
        1, std::max<int64_t>(max_graph_nnz0, std::min<int64_t>(nnz0, grp)));
    HIPC(ctx, cflag.ensure(4 * n));
    HIPC(ctx, crank.ensure(8 * (n + 1)));
    HIPC(ctx, cnewc.ensure(4 * n));
    HIPC(ctx, ck0.ensure(8 * e));
    HIPC(ctx, ck1.ensure(8 * e));
    const int64_t nblk = (e + 4095) / 4096;
    const int64_t hist_cap = 256 * nblk * 2 + 2;
    HIPC(ctx, chist.ensure(8 * hist_cap));
    HIPC(ctx, chead.ensure(4 * e));
    HIPC(ctx, cuidx.ensure(8 * (e + 1)));
    HIPC(ctx, crowcnt.ensure(8 * (n + 1)));
    const int64_t smax = std::max<int64_t>(std::max<int64_t>(n + 1, e + 1), 256 * nblk + 1);
    HIPC(ctx, cscan.ensure(DevScan::temp_bytes(smax)));
    CB.flag = cflag.as<int32_t>();
    CB.rank = crank.as<int64_t>();
    CB.newc = cnewc.as<vid>();
    CB.k0 = ck0.as<uint64_t>();
    CB.k1 = ck1.as<uint64_t>();
    CB.v0 = cv0.p;
    CB.v1 = cv1.p;
    CB.hist = chist.as<int64_t>();
    CB.head = chead.as<int32_t>();
    CB.uidx = cuidx.as<int64_t>();
    CB.rowcnt = crowcnt.as<int64_t>();
    CB.scan_tmp = cscan.p;
    CB.hist_cap = hist_cap;
    CB.cap = e;
    return MGL_OK;
  }

  // -------------------------------------------------------------------------
  // Binning of the current level (BIN::bin_create, BIN.cu:468-499).
  mgl_status make_bins() {
    LevelGraph& g = LG();
    L = layout_for(ctx, tune, g.mode, rec32_now());
    const int nb = L.nb;
    bin_n.assign(nb, 0);
    bin_off.assign(nb + 1, 0);
    rlay = false;
    if (g.n == 0) return MGL_OK;
    unsigned long long* cnt = bcounts.as<unsigned long long>();
    launch_fill_bytes0(cnt, 8 * MGL_MAX_BINS * 2, st());
    launch_bin_count(g.off.as<eid>(), g.n, L, cnt, st());
    std::vector<unsigned long long> h(MGL_MAX_BINS);
    HIPC(ctx, hipMemcpyAsync(h.data(), cnt, 8 * MGL_MAX_BINS, hipMemcpyDeviceToHost, st()));
    HIPC(ctx, hipStreamSynchronize(st()));
    for (int i = 0; i < nb; ++i) {
      bin_n[i] = (int64_t)h[i];
      bin_off[i + 1] = bin_off[i] + bin_n[i];
    }
    std::vector<unsigned long long> cur0(MGL_MAX_BINS, 0);
    for (int i = 0; i < nb; ++i) cur0[i] = (unsigned long long)bin_off[i];
    unsigned long long* cursor = cnt + MGL_MAX_BINS;
    HIPC(ctx, hipMemcpyAsync(cursor, cur0.data(), 8 * MGL_MAX_BINS, hipMemcpyHostToDevice, st()));
    HIPC(ctx, hipStreamSynchronize(st()));
    launch_bin_scatter(g.off.as<eid>(), g.n, L, cursor, bperm.as<vid>(), st());
    // records in degree-bin order once a graph's record array outgrows an L2
    // slice; in a batch of smaller graphs each graph's records stay together
    // in vertex order (every row's neighbours lie in its own graph)
    int64_t gmax = g.n;
    if (g.voff.size() > 2) {
      gmax = 0;
      for (size_t i = 0; i + 1 < g.voff.size(); ++i)
        gmax = std::max<int64_t>(gmax, g.voff[i + 1] - g.voff[i]);
    }
    rlay = o.record_layout == MGL_LAYOUT_BINNED ||
           (o.record_layout == MGL_LAYOUT_AUTO && gmax > LAYOUT_AUTO_MIN_N);
    if (rlay) {
      const int wb = (int)wsize(g.mode);
      HIPC(ctx, brpos.ensure(4 * (size_t)n0));
      HIPC(ctx, brinv.ensure(4 * (size_t)n0));
      HIPC(ctx, bcolr.ensure(4 * (size_t)std::max<int64_t>(nnz0, 1)));
      HIPC(ctx, boffr.ensure(8 * ((size_t)n0 + 1)));
      HIPC(ctx, bpdeg.ensure(8 * ((size_t)n0 + 1)));
      HIPC(ctx, bpscan.ensure(DevScan::temp_bytes(n0 + 1)));
      if (wb) HIPC(ctx, bwr.ensure((size_t)wb * (size_t)std::max<int64_t>(nnz0, 1)));
      launch_record_layout(bperm.as<vid>(), g.n, brpos.as<vid>(), brinv.as<vid>(), st());
      launch_pos_degrees(g.off.as<eid>(), brinv.as<vid>(), g.n, bpdeg.as<int64_t>(), st());
      scan_exclusive_i64(bpdeg.as<int64_t>(), boffr.as<int64_t>(), g.n + 1, bpscan.p, st());
      for (int b = 1; b < nb; ++b)
        launch_place_rows(L.b[b].upper, g.off.as<eid>(), g.col.as<vid>(), g.w.p, wb,
                          brinv.as<vid>(), brpos.as<vid>(), boffr.as<eid>(), bin_off[b], bin_n[b],
                          bcolr.as<vid>(), wb ? bwr.p : nullptr, st());
    }
    // hub plan of the last bin: chunks of each row, hash buckets per hub
    hub = HubPlan();
    const int last = nb - 1;
    if (L.b[last].shape == SHAPE_GLOBAL && bin_n[last] > 0) {
      const int64_t m = bin_n[last];
      const vid* hperm = bperm.as<vid>() + bin_off[last];
      std::vector<int64_t> ranges(2 * m);
      std::vector<vid> hubv(m);
      if (rlay) {
        // the hub rows as placed: positions bin_off[last] .. bin_off[last] + m
        std::vector<int64_t> ho(m + 1);
        HIPC(ctx, hipMemcpyAsync(ho.data(), boffr.as<int64_t>() + bin_off[last], 8 * (m + 1),
                                 hipMemcpyDeviceToHost, st()));
        HIPC(ctx, hipStreamSynchronize(st()));
        for (int64_t i = 0; i < m; ++i) {
          ranges[2 * i] = ho[i];
          ranges[2 * i + 1] = ho[i + 1];
        }
      } else {
        DBuf rr;
        HIPC(ctx, rr.ensure(16 * m));
        launch_row_ranges(g.off.as<eid>(), hperm, m, rr.as<int64_t>(), st());
        HIPC(ctx, hipMemcpyAsync(ranges.data(), rr.p, 16 * m, hipMemcpyDeviceToHost, st()));
        HIPC(ctx, hipStreamSynchronize(st()));  // before rr is released
      }
      HIPC(ctx, hipMemcpyAsync(hubv.data(), hperm, 4 * m, hipMemcpyDeviceToHost, st()));
      HIPC(ctx, hipStreamSynchronize(st()));
      const int64_t CH = hub_chunk_edges();
      // owner tables of the aggregate: (key, sum) slots in a share of the
      // CU's LDS, so that hub_agg_per_cu buckets share a CU; prime-sized. The
      // segmented aggregate also holds its hub's segment table (two 4-byte
      // entries per chunk) in static LDS.
      const int asb = slot_bytes(g.mode) == 4 ? 8 : 4 + slot_bytes(g.mode);
      const int seg = 8 * hub_seg_max() + 64;
      const uint32_t Hlds = (uint32_t)prime_at_most(
          (int)((ctx->lds_max / tune.hub_agg_per_cu - 512 - seg) / asb));
      // a bucket's candidates must fit the owner path (at most 7/8 of the
      // table and hub_agg_cap() per workgroup) with a quarter to spare for the
      // spread of the hash partition
      int64_t BK = tune.hub_bucket;
      if (tune.hub_bucket_clamp) {
        BK = std::min<int64_t>(BK, (int64_t)Hlds * 7 / 8 * 4 / 5);
        BK = std::min<int64_t>(BK, (int64_t)hub_agg_cap() * 4 / 5);
        if (BK < 64) BK = 64;
      }
      std::vector<int32_t> hgid(m, 0);
      if (g.gid.p) {
        // graph of each hub vertex (batches)
        DBuf hg;
        HIPC(ctx, hg.ensure(4 * m));
        launch_gather_i32(g.gid.as<int32_t>(), hperm, m, hg.as<int32_t>(), st());
        HIPC(ctx, hipMemcpyAsync(hgid.data(), hg.p, 4 * m, hipMemcpyDeviceToHost, st()));
        HIPC(ctx, hipStreamSynchronize(st()));
      }
      std::vector<HubChunk> chunk;
      std::vector<int64_t> hbkv(m + 1, 0), bkhub;
      std::vector<HubBucket> bmeta;
      int64_t pos = 0, ncbo = 0, max_nch = 0;
      for (int64_t i = 0; i < m; ++i) {
        const int64_t e0 = ranges[2 * i], e1 = ranges[2 * i + 1], d = e1 - e0;
        int64_t P = (d + BK - 1) / BK;
        if (P < 1) P = 1;
        if (P > hub_max_buckets()) P = hub_max_buckets();
        hbkv[i + 1] = hbkv[i] + P;
        const int64_t nch = (d + CH - 1) / CH;
        for (int64_t c = 0; c < nch; ++c) {
          const int64_t ce = e0 + c * CH;
          chunk.push_back(HubChunk{ce, std::min(ce + CH, e1), pos + c * CH, ncbo + c * (P + 1),
                                   (int32_t)i, (int32_t)P, hubv[i], hgid[i]});
        }
        for (int64_t j = 0; j < P; ++j) {
          bkhub.push_back(i);
          bmeta.push_back(HubBucket{pos, ncbo, (int32_t)i, hubv[i], (int32_t)j, (int32_t)P,
                                    (int32_t)nch, hgid[i]});
        }
        pos += d;
        ncbo += nch * (P + 1);
        max_nch = std::max(max_nch, nch);
      }
      // segmented form unless a hub has more chunks than the aggregate's LDS
      // segment table holds (degree above hub_seg_max() chunks)
      hub.segmented = (tune.hub_segmented && max_nch <= hub_seg_max()) ? 1 : 0;
      hub.m = m;
      hub.nchunks = (int64_t)chunk.size();
      hub.nbuckets = hbkv[m];
      hub.ehub = pos;
      const size_t sb = (size_t)slot_bytes(g.mode);
      const size_t e = (size_t)std::max<int64_t>(pos, 1);
      HIPC(ctx, hchunk.ensure(sizeof(HubChunk) * std::max<size_t>(chunk.size(), 1)));
      HIPC(ctx, hgin.ensure(sizeof(HubGain) * m));
      HIPC(ctx, hv.ensure(4 * m));
      HIPC(ctx, hbk.ensure(8 * (m + 1)));
      HIPC(ctx, hbkhub.ensure(8 * hub.nbuckets));
      const size_t ab = (size_t)total_bytes(g.mode, rec32_now());
      HIPC(ctx, htmpc.ensure(4 * e));
      HIPC(ctx, hbktc.ensure(4 * e));
      HIPC(ctx, htmpa.ensure(ab * e));
      HIPC(ctx, hbkta.ensure(ab * e));
      HIPC(ctx, hga.ensure(ab * 2 * e));
      if (g.mode != MODE_UNIT) {
        HIPC(ctx, htmpw.ensure(sb * e));
        HIPC(ctx, hbktw.ensure(sb * e));
      }
      HIPC(ctx, hbcount.ensure(8 * hub.nbuckets));
      HIPC(ctx, hcbo.ensure(4 * std::max<int64_t>(ncbo, 1)));
      HIPC(ctx, hgalloc.ensure(8));
      HIPC(ctx, hboff.ensure(8 * (hub.nbuckets + 1)));
      HIPC(ctx, hbcur.ensure(8 * hub.nbuckets));
      HIPC(ctx, heici.ensure(8 * m));
      HIPC(ctx, hself.ensure(8 * m));
      HIPC(ctx, hbgain.ensure(8 * hub.nbuckets));
      HIPC(ctx, hbdst.ensure(4 * hub.nbuckets));
      HIPC(ctx, hgk.ensure(4 * 2 * e));
      HIPC(ctx, hgv.ensure(sb * 2 * e));
      HIPC(ctx, hscan.ensure(DevScan::temp_bytes(hub.nbuckets + 1)));
      HIPC(ctx, hipMemcpyAsync(hchunk.p, chunk.data(), sizeof(HubChunk) * chunk.size(),
                               hipMemcpyHostToDevice, st()));
      HIPC(ctx, hipMemcpyAsync(hv.p, hperm, 4 * m, hipMemcpyDeviceToDevice, st()));
      HIPC(ctx, hipMemcpyAsync(hbk.p, hbkv.data(), 8 * (m + 1), hipMemcpyHostToDevice, st()));
      HIPC(ctx, hipMemcpyAsync(hbkhub.p, bkhub.data(), 8 * hub.nbuckets, hipMemcpyHostToDevice, st()));
      HIPC(ctx, hbmeta.ensure(sizeof(HubBucket) * hub.nbuckets));
      HIPC(ctx, hipMemcpyAsync(hbmeta.p, bmeta.data(), sizeof(HubBucket) * hub.nbuckets,
                               hipMemcpyHostToDevice, st()));
      HIPC(ctx, hipStreamSynchronize(st()));
      hub.chunks = hchunk.as<HubChunk>();
      hub.gin = hgin.as<HubGain>();
      hub.hub_v = hv.as<vid>();
      hub.hub_bk = hbk.as<int64_t>();
      hub.bk_hub = hbkhub.as<int64_t>();
      hub.tmp_c = htmpc.as<vid>();
      hub.tmp_a = htmpa.p;
      hub.bkt_a = hbkta.p;
      hub.ga = hga.p;
      hub.tmp_w = htmpw.p;
      hub.bkt_c = hbktc.as<vid>();
      hub.bkt_w = hbktw.p;
      hub.bcount = hbcount.as<int64_t>();
      hub.boff = hboff.as<int64_t>();
      hub.bcur = hbcur.as<int64_t>();
      hub.hub_eici = heici.p;
      hub.hub_self = hself.p;
      hub.bgain = hbgain.as<double>();
      hub.bdst = hbdst.as<vid>();
      hub.gk = hgk.as<vid>();
      hub.gv = hgv.p;
      hub.scan_tmp = hscan.p;
      hub.cbo = hcbo.as<int32_t>();
      hub.galloc = hgalloc.as<unsigned long long>();
      hub.bmeta = hbmeta.as<HubBucket>();
      hub.lds_slots = Hlds;
      hub.agg_bs = tune.hub_agg_threads;
      hub.count_bs = tune.hub_count_threads;
    }
    return MGL_OK;
  }

  MoveArgs move_args(const vid* from, vid* to, int up, void* vin) {
    LevelGraph& g = LG();
    MoveArgs A;
    A.off = rlay ? boffr.as<eid>() : g.off.as<eid>();
    A.col = rlay ? bcolr.as<vid>() : g.col.as<vid>();
    A.w = (rlay && wsize(g.mode)) ? bwr.p : g.w.p;
    A.k = rlay ? bkr.p : bk.p;
    A.c = from;
    A.a = ba.p;
    A.rec = brec.p;
    A.rpos = rlay ? brpos.as<vid>() : nullptr;
    A.c_new = to;
    A.vin = vin;
    A.gid = g.gid.as<int32_t>();
    A.gs = bgs.as<GState>();
    A.up = up;
    A.version = o.version;
    A.tie_small = o.tie_rule == MGL_TIE_SMALLEST;
    A.stats = nullptr;
    A.probes = nullptr;
    A.pos0 = -1;
    A.err = err_word();
    return A;
  }

  // MoveArgs of the hub bin (its counters, no record positions by bin).
  MoveArgs hub_args(MoveArgs A, int b) {
    if (counting) {
      A.stats = bcounts.as<unsigned long long>() + 2 * MGL_MAX_BINS + b;
      A.probes = bcounts.as<unsigned long long>() + 4 * MGL_MAX_BINS;
    }
    A.pos0 = -1;
    return A;
  }

  // All move kernels of the level (calculate_eicj_and_move_vertex_bin,
  // louvain.cu:1032-1296), largest bins first.
  mgl_status move(const vid* from, vid* to, int up, void* vin, bool time_bins) {
    LevelGraph& g = LG();
    MoveArgs A = move_args(from, to, up, vin);
    const int r32 = rec32_now() ? 1 : 0;
    // in degree-bin order the isolated vertices (bin 0) take the first
    // positions; no CSR entry points to them and they are never moved
    const int64_t first = (rlay && nlist < 0 && L.nb > 1) ? bin_off[1] : 0;
    launch_build_rec(g.mode, r32, vcount(), from, ba.p, brec.p, st(), vlist(),
                     rlay ? brpos.as<vid>() : nullptr, rlay ? brinv.as<vid>() : nullptr, first);
    // the hub path (several dependent kernels, partly latency bound) can run
    // on a second stream beside the bins of short rows (degree <= 128: many
    // small, latency-bound launches): it is forked once the larger bins are
    // done and joined before anything reads the proposals; the bins write
    // disjoint proposals and only read the records
    const int hb = L.nb - 1;
    const bool fork = tune.hub_stream && ctx->st2 && L.b[hb].shape == SHAPE_GLOBAL &&
                      bin_n[hb] > 0;
    int fork_at = 0;  // the hub path starts when the bins above this one are done
    if (fork)
      for (int b = hb - 1; b >= 1; --b)
        if (L.b[b].upper <= tune.hub_fork_upper) { fork_at = b; break; }
    shared_from = fork ? fork_at : 0;
    int ne = 0;
    for (int b = L.nb - 1; b >= 1; --b) {
      if (fork && b == fork_at) {
        HIPC(ctx, hipEventRecord(ctx->evfork, st()));
        HIPC(ctx, hipStreamWaitEvent(ctx->st2, ctx->evfork, 0));
        if (time_bins) HIPC(ctx, hipEventRecord(ev[2 * hb], ctx->st2));
        if (launch_hub(g.mode, r32, hub_args(A, hb), hub, ctx->st2) != 0)
          FAIL(ctx, MGL_UNSUPPORTED, "no hub-path kernels for mode %d", g.mode);
        if (time_bins) HIPC(ctx, hipEventRecord(ev[2 * hb + 1], ctx->st2));
        HIPC(ctx, hipEventRecord(ctx->evjoin, ctx->st2));
      }
      if (bin_n[b] == 0) continue;
      if (fork && b == hb) continue;  // launched beside the short-row bins
      hipStream_t sb = st();
      // per-bin count of aggregated (vertex, neighbour community) pairs
      if (counting) {
        A.stats = bcounts.as<unsigned long long>() + 2 * MGL_MAX_BINS + b;
        A.probes = bcounts.as<unsigned long long>() + 4 * MGL_MAX_BINS;
      }
      A.pos0 = (rlay && nlist < 0 && L.b[b].shape != SHAPE_GLOBAL) ? bin_off[b] : -1;
      if (time_bins) HIPC(ctx, hipEventRecord(ev[2 * b], sb));
      const int rc = (L.b[b].shape == SHAPE_GLOBAL)
                         ? launch_hub(g.mode, r32, A, hub, sb)
                         : launch_move_bin(g.mode, r32, L.b[b], A, bperm.as<vid>() + bin_off[b],
                                           bin_n[b], sb);
      if (rc != 0) FAIL(ctx, MGL_UNSUPPORTED, "no move kernel for bin %d (s=%d size=%d)", b, L.b[b].s, L.b[b].size);
      if (time_bins) HIPC(ctx, hipEventRecord(ev[2 * b + 1], sb));
      ++ne;
    }
    if (fork && fork_at == 0) {
      // no short-row bin: the hub path runs after the others, on its stream
      HIPC(ctx, hipEventRecord(ctx->evfork, st()));
      HIPC(ctx, hipStreamWaitEvent(ctx->st2, ctx->evfork, 0));
      if (time_bins) HIPC(ctx, hipEventRecord(ev[2 * hb], ctx->st2));
      if (launch_hub(g.mode, r32, hub_args(A, hb), hub, ctx->st2) != 0)
        FAIL(ctx, MGL_UNSUPPORTED, "no hub-path kernels for mode %d", g.mode);
      if (time_bins) HIPC(ctx, hipEventRecord(ev[2 * hb + 1], ctx->st2));
      HIPC(ctx, hipEventRecord(ctx->evjoin, ctx->st2));
    }
    if (fork) HIPC(ctx, hipStreamWaitEvent(st(), ctx->evjoin, 0));
    HIPC(ctx, hipGetLastError());
    return MGL_OK;
  }

  mgl_status rows(int op, const vid* cc, void* kout, void* vin, bool skip_inactive) {
    LevelGraph& g = LG();
    RowArgs R;
    R.off = g.off.as<eid>();
    R.col = g.col.as<vid>();
    R.w = g.w.p;
    R.c = cc;
    R.k = kout;
    R.vin = vin;
    R.gid = g.gid.as<int32_t>();
    R.gs = skip_inactive ? bgs.as<GState>() : nullptr;
    for (int b = L.nb - 1; b >= 1; --b)
      launch_rows_bin(g.mode, op, L.b[b], R, bperm.as<vid>() + bin_off[b], bin_n[b], st());
    HIPC(ctx, hipGetLastError());
    return MGL_OK;
  }

  // a[x] = sum of k over members of x (used after an externally set c).
  mgl_status recompute_a(const vid* cc);

  // Level start (louvain.cu:1492-1533) for the graphs with running[g] = 1.
  mgl_status level_start() {
    LevelGraph& g = LG();
    std::vector<int32_t> run(G);
    for (int i = 0; i < G; ++i) run[i] = running[i];
    HIPC(ctx, hipMemcpyAsync(brun.p, run.data(), 4 * G, hipMemcpyHostToDevice, st()));
    HIPC(ctx, hipStreamSynchronize(st()));
    launch_level_begin(G, bgs.as<GState>(), brun.as<int32_t>(), st());
    nlist = -1;
    mgl_status s = make_bins();
    if (s) return s;
    launch_fill_bytes0(bk.p, 8 * (size_t)std::max<int64_t>(g.n, 1), st());
    launch_fill_bytes0(bvin.p, 8 * (size_t)std::max<int64_t>(g.n, 1), st());
    s = rows(OP_WEIGHT, nullptr, bk.p, bvin.p, false);
    if (s) return s;
    if (rlay) {
      // the move kernels read k by record position, like the row offsets
      HIPC(ctx, bkr.ensure(8 * (size_t)std::max<int64_t>(n0, 1)));
      launch_gather_u64(bk.p, brinv.as<vid>(), g.n, bkr.p, st());
    }
    launch_copy_k_to_a(g.mode, g.n, bk.p, ba.p, st());
    launch_iota(c, g.n, st());
    launch_iota(p, g.n, st());
    launch_iota(nx, g.n, st());
    launch_reduce(g.mode, g.n, bvin.p, ba.p, g.gid.as<int32_t>(), bgs.as<GState>(), st());
    launch_level_init(G, bgs.as<GState>(), mode_exact(g.mode) ? 1 : 0, st());
    HIPC(ctx, hipGetLastError());
    return MGL_OK;
  }

  // First proposal of a level in fused mode: p = move(c), a = a(p), sum a^2(p).
  mgl_status pre_move() {
    LevelGraph& g = LG();
    mgl_status s = move(c, p, 1, nullptr, false);
    if (s) return s;
    launch_update(g.mode, g.n, c, p, bk.p, ba.p, g.gid.as<int32_t>(), bgs.as<GState>(), st());
    launch_reduce(g.mode, g.n, nullptr, ba.p, g.gid.as<int32_t>(), bgs.as<GState>(), st());
    launch_decide(G, bgs.as<GState>(), o.threshold, o.max_iter, 1, mode_exact(g.mode), nullptr, 0,
                  1, st());
    HIPC(ctx, hipGetLastError());
    return MGL_OK;
  }

  // One inner iteration. it = 1-based index of the proposal being decided.
  mgl_status iteration(int it) {
    LevelGraph& g = LG();
    const bool ex = mode_exact(g.mode);
    const int32_t* gid = g.gid.as<int32_t>();
    GState* gs = bgs.as<GState>();
    double* traj = btraj.as<double>();
    if (timing) HIPC(ctx, hipEventRecord(ev[2 * MGL_MAX_BINS + 0], st()));
    if (o.fuse_modularity) {
      const int up = ((it + 1) & 1) ? 1 : 0;  // direction of proposal it+1
      mgl_status s = move(p, nx, up, bvin.p, timing);
      if (s) return s;
      if (timing) HIPC(ctx, hipEventRecord(ev[2 * MGL_MAX_BINS + 1], st()));
      launch_update(g.mode, vcount(), p, nx, bk.p, ba.p, gid, gs, st(), vlist());
      if (timing) HIPC(ctx, hipEventRecord(ev[2 * MGL_MAX_BINS + 2], st()));
      launch_reduce(g.mode, vcount(), bvin.p, ba.p, gid, gs, st(), vlist());
      launch_fill_bytes0(nactive(), 4, st());
      launch_decide(G, gs, o.threshold, o.max_iter, 1, ex, traj, TRAJ_MAX, 0, st(), nactive(),
                    bttraj.as<int64_t>());
      if (timing) HIPC(ctx, hipEventRecord(ev[2 * MGL_MAX_BINS + 3], st()));
      launch_fixup(vcount(), c, p, nx, gid, gs, st(), vlist());
      vid* t = c; c = p; p = nx; nx = t;
    } else {
      const int up = (it & 1) ? 1 : 0;
      mgl_status s = move(c, p, up, nullptr, timing);
      if (s) return s;
      if (timing) HIPC(ctx, hipEventRecord(ev[2 * MGL_MAX_BINS + 1], st()));
      launch_update(g.mode, vcount(), c, p, bk.p, ba.p, gid, gs, st(), vlist());
      if (timing) HIPC(ctx, hipEventRecord(ev[2 * MGL_MAX_BINS + 2], st()));
      s = rows(OP_INTRA, p, nullptr, bvin.p, true);
      if (s) return s;
      launch_reduce(g.mode, vcount(), bvin.p, ba.p, gid, gs, st(), vlist());
      launch_fill_bytes0(nactive(), 4, st());
      launch_decide(G, gs, o.threshold, o.max_iter, 0, ex, traj, TRAJ_MAX, 0, st(), nactive(),
                    bttraj.as<int64_t>());
      if (timing) HIPC(ctx, hipEventRecord(ev[2 * MGL_MAX_BINS + 3], st()));
      launch_fixup(vcount(), c, p, nullptr, gid, gs, st(), vlist());
    }
    if (timing) HIPC(ctx, hipEventRecord(ev[2 * MGL_MAX_BINS + 4], st()));
    HIPC(ctx, hipGetLastError());
    return MGL_OK;
  }

  // The graph states into hgs; with check_error also the error word (one
  // synchronisation for both), reported as check_err does.
  mgl_status read_gs(bool check_error = false) {
    const size_t gb = sizeof(GState) * G;
    HIPC(ctx, hpin.ensure(gb + 16));
    HIPC(ctx, hipMemcpyAsync(hpin.p, bgs.p, gb, hipMemcpyDeviceToHost, st()));
    if (check_error)
      HIPC(ctx, hipMemcpyAsync((char*)hpin.p + gb, err_word(), 4, hipMemcpyDeviceToHost, st()));
    HIPC(ctx, hipStreamSynchronize(st()));
    hgs.resize(G);
    memcpy(hgs.data(), hpin.p, gb);
    if (check_error) {
      int e;
      memcpy(&e, (char*)hpin.p + gb, 4);
      if (e) return err_overflow();
    }
    return MGL_OK;
  }

  mgl_status any_active(bool& any) {
    mgl_status s = read_gs();
    if (s) return s;
    any = false;
    for (int i = 0; i < G; ++i) any |= hgs[i].active != 0;
    return MGL_OK;
  }

  mgl_status read_nactive(int& na) {
    HIPC(ctx, hpin.ensure(16));
    HIPC(ctx, hipMemcpyAsync(hpin.p, nactive(), 4, hipMemcpyDeviceToHost, st()));
    HIPC(ctx, hipStreamSynchronize(st()));
    memcpy(&na, hpin.p, 4);
    return MGL_OK;
  }

  // Batch compaction: keep only the vertices of graphs still active in the
  // move bins, and list the active non-isolated vertices (increasing ids) for
  // the per-vertex passes, so finished graphs stop costing launch work.
  mgl_status compact() {
    LevelGraph& g = LG();
    unsigned long long* cnt = bcnt2.as<unsigned long long>();
    launch_fill_bytes0(cnt, 8 * MGL_MAX_BINS, st());
    for (int b = 1; b < L.nb; ++b)
      if (bin_n[b] > 0)
        launch_filter_bin(bperm.as<vid>() + bin_off[b], bin_n[b], g.gid.as<int32_t>(),
                          bgs.as<GState>(), bperm2.as<vid>() + bin_off[b], cnt + b, st());
    launch_active_list(g.n, g.off.as<eid>(), g.gid.as<int32_t>(), bgs.as<GState>(), CB.flag,
                       CB.rank, blist.as<vid>(), CB.scan_tmp, st());
    std::vector<unsigned long long> h(MGL_MAX_BINS);
    HIPC(ctx, hipMemcpyAsync(h.data(), cnt, 8 * MGL_MAX_BINS, hipMemcpyDeviceToHost, st()));
    int64_t nl = 0;
    HIPC(ctx, hipMemcpyAsync(&nl, CB.rank + g.n, 8, hipMemcpyDeviceToHost, st()));
    HIPC(ctx, hipStreamSynchronize(st()));
    for (int b = 1; b < L.nb; ++b) bin_n[b] = (int64_t)h[b];
    std::swap(bperm.p, bperm2.p);
    std::swap(bperm.bytes, bperm2.bytes);
    nlist = nl;
    HIPC(ctx, hipGetLastError());
    return MGL_OK;
  }

  // Inner loop of a level (louvain.cu:1548-1660). With use_graph the
  // iterations are captured once per level into a hipGraph holding R of them
  // (R = 6: the proposal buffers rotate with period 3 and the up/down
  // direction alternates with period 2) and replayed until no graph of the
  // batch is active; iterations after a graph's last decision are no-ops for
  // it. In a batch, whenever the number of active graphs has halved the
  // vertex sets are compacted (and the hipGraph captured again).
  mgl_status run_level() {
    mgl_status s;
    if (o.fuse_modularity) {
      s = pre_move();
      if (s) return s;
    }
    int last_na = 0;
    for (int i = 0; i < G; ++i) last_na += running[i] ? 1 : 0;
    const bool can_compact = G > 1 && CB.flag != nullptr;
    if (!o.use_graph) {
      for (int it = 1;; ++it) {
        if (it > (1 << 22)) FAIL(ctx, MGL_DEVICE_ERROR, "inner loop did not terminate");
        s = iteration(it);
        if (s) return s;
        int na;
        s = read_nactive(na);
        if (s) return s;
        if (na == 0) break;
        if (can_compact && 2 * na <= last_na) {
          s = compact();
          if (s) return s;
          last_na = na;
        }
      }
      return MGL_OK;
    }
    const int R = o.fuse_modularity ? 6 : 2;
    int64_t done = 0;
    while (true) {
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      HIPC(ctx, hipStreamBeginCapture(st(), hipStreamCaptureModeThreadLocal));
      for (int j = 0; j < R; ++j) {
        s = iteration(1 + j);
        if (s) {
          (void)hipStreamEndCapture(st(), &graph);
          if (graph) (void)hipGraphDestroy(graph);
          (void)hipGetLastError();
          return s;
        }
      }
      HIPC(ctx, hipStreamEndCapture(st(), &graph));
      hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        HIPC(ctx, e);
      }
      mgl_status rs = MGL_OK;
      bool finished = false, recapture = false;
      while (true) {
        if (done * R > (1 << 22)) {
          ctx->err = "inner loop did not terminate";
          rs = MGL_DEVICE_ERROR;
          break;
        }
        e = hipGraphLaunch(exec, st());
        if (e != hipSuccess) {
          (void)hipGetLastError();
          ctx->err = std::string("hipGraphLaunch: ") + hipGetErrorString(e);
          rs = MGL_DEVICE_ERROR;
          break;
        }
        ++done;
        int na;
        rs = read_nactive(na);
        if (rs) break;
        if (na == 0) { finished = true; break; }
        if (can_compact && 2 * na <= last_na) {
          last_na = na;
          recapture = true;
          break;
        }
      }
      (void)hipGraphExecDestroy(exec);
      (void)hipGraphDestroy(graph);
      if (rs) return rs;
      if (finished) return MGL_OK;
      if (recapture) {
        s = compact();
        if (s) return s;
      }
    }
  }

  mgl_status level_end(int level);
  mgl_status final_q_fp64(std::vector<double>& q);
  mgl_status run(std::vector<mgl_result*>& out);
  mgl_status setup_union(const mgl_csr* graphs, int ng);
};

// ---------------------------------------------------------------------------
__global__ void k_scatter_a(int64_t n, const vid* c, const int64_t* k, int64_t* a) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    atomicAdd((unsigned long long*)&a[c[v]], (unsigned long long)k[v]);
}
__global__ void k_scatter_a_d(int64_t n, const vid* c, const double* k, double* a) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&a[c[v]], k[v]);
}
mgl_status Engine::recompute_a(const vid* cc) {
  LevelGraph& g = LG();
  launch_fill_bytes0(ba.p, 8 * (size_t)std::max<int64_t>(g.n, 1), st());
  const unsigned grid = (unsigned)std::min<int64_t>(8192, (g.n + 255) / 256 + 1);
  if (mode_exact(g.mode))
    k_scatter_a<<<grid, 256, 0, st()>>>(g.n, cc, bk.as<int64_t>(), ba.as<int64_t>());
  else
    k_scatter_a_d<<<grid, 256, 0, st()>>>(g.n, cc, bk.as<double>(), ba.as<double>());
  HIPC(ctx, hipGetLastError());
  return MGL_OK;
}

__global__ void k_global_labels(int64_t n0, const vid* fin, const int32_t* gid0,
                                const int64_t* vbase, vid* out) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n0;
       v += (int64_t)gridDim.x * blockDim.x)
    out[v] = fin[v] + (vid)vbase[gid0 ? gid0[v] : 0];
}

// FP64 modularity of the final labels on the input graph (float modes).
mgl_status Engine::final_q_fp64(std::vector<double>& q) {
  cur = 0;
  LevelGraph& g = LG();
  std::vector<int32_t> run(G, 1);
  HIPC(ctx, hipMemcpyAsync(brun.p, run.data(), 4 * G, hipMemcpyHostToDevice, st()));
  HIPC(ctx, hipStreamSynchronize(st()));
  launch_level_begin(G, bgs.as<GState>(), brun.as<int32_t>(), st());
  mgl_status s = make_bins();
  if (s) return s;
  // global label of every input vertex
  DBuf dv;
  HIPC(ctx, dv.ensure(8 * (G + 1)));
  HIPC(ctx, hipMemcpyAsync(dv.p, vbase0.data(), 8 * (G + 1), hipMemcpyHostToDevice, st()));
  const unsigned gr = (unsigned)std::min<int64_t>(8192, (n0 + 255) / 256 + 1);
  k_global_labels<<<gr, 256, 0, st()>>>(n0, bfin.as<vid>(), bgid0.as<int32_t>(), dv.as<int64_t>(),
                                        c);
  launch_fill_bytes0(bk.p, 8 * (size_t)std::max<int64_t>(g.n, 1), st());
  launch_fill_bytes0(bvin.p, 8 * (size_t)std::max<int64_t>(g.n, 1), st());
  s = rows(OP_WEIGHT, nullptr, bk.p, bvin.p, false);
  if (s) return s;
  s = recompute_a(c);
  if (s) return s;
  launch_fill_bytes0(bvin.p, 8 * (size_t)std::max<int64_t>(g.n, 1), st());
  s = rows(OP_INTRA, c, nullptr, bvin.p, false);
  if (s) return s;
  launch_reduce(g.mode, g.n, bvin.p, ba.p, g.gid.as<int32_t>(), bgs.as<GState>(), st());
  s = read_gs();
  if (s) return s;
  q.resize(G);
  for (int i = 0; i < G; ++i) {
    const double M = hgs[i].mass;
    q[i] = (M == 0) ? 0.0 : hgs[i].sin_d / M - hgs[i].sa2_d / (M * M);
  }
  return MGL_OK;
}

// label composition kernels
// fin[v0] = rank_all[c[lab[v0]]] - rank_all[voff[g]]  (dense label inside graph g)
__global__ void k_label_final(int64_t n0, const vid* lab, const vid* c, const int64_t* rank,
                              const int32_t* gid0, const int32_t* run, const int64_t* gbase,
                              vid* fin) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n0;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int g = gid0 ? gid0[v] : 0;
    if (!run[g]) continue;
    fin[v] = (vid)(rank[c[lab[v]]] - gbase[g]);
  }
}
// lab[v0] = newc[lab[v0]] for graphs that continue
__global__ void k_label_next(int64_t n0, vid* lab, const vid* newc, const int32_t* gid0,
                             const int32_t* keep) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n0;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int g = gid0 ? gid0[v] : 0;
    if (!keep[g]) continue;
    lab[v] = newc[lab[v]];
  }
}
__global__ void k_gather_i64(const int64_t* src, const int64_t* idx, int m, int64_t* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) out[i] = src[idx[i]];
}
__global__ void k_flag_run(int64_t n, const vid* c, const int32_t* gid, const int32_t* run,
                           int32_t* flag) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    if (!run[gid ? gid[v] : 0]) continue;
    flag[c[v]] = 1;
  }
}

static mgl_status copy_level_graph_to_host(mgl_ctx* ctx, LevelGraph& g, int64_t v0, int64_t v1,
                                           int scale, LevelOut& lo) {
  const int64_t e0 = g.eoff.empty() ? 0 : 0;
  (void)e0;
  std::vector<int64_t> off(v1 - v0 + 1);
  HIPC(ctx, hipMemcpyAsync(off.data(), g.off.as<eid>() + v0, 8 * (v1 - v0 + 1),
                           hipMemcpyDeviceToHost, ctx->st));
  HIPC(ctx, hipStreamSynchronize(ctx->st));
  const int64_t eb = off[0], ee = off[v1 - v0];
  lo.gn = v1 - v0;
  lo.gnnz = ee - eb;
  lo.goff.resize(off.size());
  for (size_t i = 0; i < off.size(); ++i) lo.goff[i] = off[i] - eb;
  lo.gcol.resize(lo.gnnz);
  lo.gw.resize(lo.gnnz);
  if (lo.gnnz) {
    HIPC(ctx, hipMemcpyAsync(lo.gcol.data(), g.col.as<vid>() + eb, 4 * lo.gnnz,
                             hipMemcpyDeviceToHost, ctx->st));
    if (g.mode == MODE_UNIT) {
      std::fill(lo.gw.begin(), lo.gw.end(), 1.0);
    } else if (g.mode == MODE_I32) {
      std::vector<int32_t> w(lo.gnnz);
      HIPC(ctx, hipMemcpyAsync(w.data(), g.w.as<int32_t>() + eb, 4 * lo.gnnz,
                               hipMemcpyDeviceToHost, ctx->st));
      HIPC(ctx, hipStreamSynchronize(ctx->st));
      for (int64_t i = 0; i < lo.gnnz; ++i) lo.gw[i] = std::ldexp((double)w[i], -scale);
    } else if (g.mode == MODE_INT) {
      std::vector<int64_t> w(lo.gnnz);
      HIPC(ctx, hipMemcpyAsync(w.data(), g.w.as<int64_t>() + eb, 8 * lo.gnnz,
                               hipMemcpyDeviceToHost, ctx->st));
      HIPC(ctx, hipStreamSynchronize(ctx->st));
      for (int64_t i = 0; i < lo.gnnz; ++i) lo.gw[i] = std::ldexp((double)w[i], -scale);
    } else if (g.mode == MODE_F64) {
      HIPC(ctx, hipMemcpyAsync(lo.gw.data(), g.w.as<double>() + eb, 8 * lo.gnnz,
                               hipMemcpyDeviceToHost, ctx->st));
    } else {
      std::vector<float> w(lo.gnnz);
      HIPC(ctx, hipMemcpyAsync(w.data(), g.w.as<float>() + eb, 4 * lo.gnnz,
                               hipMemcpyDeviceToHost, ctx->st));
      HIPC(ctx, hipStreamSynchronize(ctx->st));
      for (int64_t i = 0; i < lo.gnnz; ++i) lo.gw[i] = w[i];
    }
    HIPC(ctx, hipStreamSynchronize(ctx->st));
    for (int64_t i = 0; i < lo.gnnz; ++i) lo.gcol[i] -= (int32_t)v0;
  }
  lo.has_graph = true;
  return MGL_OK;
}

// Level end (louvain.cu:1664-1693): record the level, decide per graph whether
// to coarsen, build the next level's union graph.
mgl_status Engine::level_end(int level) {
  LevelGraph& g = LG();
  mgl_status s = read_gs(true);
  if (s) return s;
  // trajectories: only the columns any graph used
  int tmax = 1;
  for (int i = 0; i < G; ++i)
    if (running[i]) tmax = std::max(tmax, std::min(hgs[i].iters, TRAJ_MAX));
  std::vector<double> traj((size_t)G * TRAJ_MAX);
  HIPC(ctx, hipMemcpy2DAsync(traj.data(), sizeof(double) * TRAJ_MAX, btraj.p,
                             sizeof(double) * TRAJ_MAX, sizeof(double) * tmax, G,
                             hipMemcpyDeviceToHost, st()));
  std::vector<int64_t> ttraj((size_t)G * TRAJ_MAX);
  HIPC(ctx, hipMemcpy2DAsync(ttraj.data(), sizeof(int64_t) * TRAJ_MAX, bttraj.p,
                             sizeof(int64_t) * TRAJ_MAX, sizeof(int64_t) * tmax, G,
                             hipMemcpyDeviceToHost, st()));
  std::vector<int32_t> hc;
  if (o.keep_levels && g.n > 0) {
    hc.resize(g.n);
    HIPC(ctx, hipMemcpyAsync(hc.data(), c, 4 * g.n, hipMemcpyDeviceToHost, st()));
  }
  HIPC(ctx, hipStreamSynchronize(st()));
  std::vector<int32_t> keep(G, 0);
  for (int i = 0; i < G; ++i) {
    if (!running[i]) continue;
    LevelOut lo;
    lo.n = g.voff[i + 1] - g.voff[i];
    lo.q = hgs[i].q_cur;
    lo.iters = hgs[i].iters;
    const int nt = std::min(lo.iters, TRAJ_MAX);
    lo.qtraj.assign(traj.begin() + (size_t)i * TRAJ_MAX, traj.begin() + (size_t)i * TRAJ_MAX + nt);
    lo.q0 = hgs[i].q0;
    lo.iter_secs.resize(nt);
    for (int t = 0; t < nt; ++t)
      lo.iter_secs[t] = (double)ttraj[(size_t)i * TRAJ_MAX + t] / ctx->clock_hz;
    if (o.keep_levels) {
      lo.assign.resize(lo.n);
      for (int64_t v = 0; v < lo.n; ++v) lo.assign[v] = hc[g.voff[i] + v] - (int32_t)g.voff[i];
    }
    if (o.keep_graphs && level > 0) {
      s = copy_level_graph_to_host(ctx, g, g.voff[i], g.voff[i + 1], scale, lo);
      if (s) return s;
    }
    res[i]->levels.push_back(std::move(lo));
    Q_host[i] = hgs[i].q_cur;
    const bool gained = (Q_host[i] - Q_old[i]) > o.threshold;  // louvain.cu:1674
    // the reference coarsens, increments the phase and re-tests phase < max_phases;
    // a coarsening whose level would never run is skipped here
    keep[i] = (gained && phase[i] + 1 < o.max_phases) ? 1 : 0;
  }
  // final labels of every graph that ran this level: dense rank of c within the graph
  {
    std::vector<int32_t> run(G);
    for (int i = 0; i < G; ++i) run[i] = running[i];
    HIPC(ctx, hipMemcpyAsync(brun.p, run.data(), 4 * G, hipMemcpyHostToDevice, st()));
    if (g.n > 0) {
      launch_fill_bytes0(CB.flag, 4 * (size_t)g.n, st());
      const unsigned gr = (unsigned)std::min<int64_t>(8192, (g.n + 255) / 256 + 1);
      k_flag_run<<<gr, 256, 0, st()>>>(g.n, c, g.gid.as<int32_t>(), brun.as<int32_t>(), CB.flag);
      scan_exclusive_i32(CB.flag, CB.rank, g.n, CB.scan_tmp, st());
      std::vector<int64_t> idx(G);
      for (int i = 0; i < G; ++i) idx[i] = g.voff[i];
      int64_t* didx = bgather.as<int64_t>();
      int64_t* dout = didx + (G + 1);
      HIPC(ctx, hipMemcpyAsync(didx, idx.data(), 8 * G, hipMemcpyHostToDevice, st()));
      k_gather_i64<<<(G + 255) / 256, 256, 0, st()>>>(CB.rank, didx, G, dout);
      const unsigned g0 = (unsigned)std::min<int64_t>(8192, (n0 + 255) / 256 + 1);
      k_label_final<<<g0, 256, 0, st()>>>(n0, blab.as<vid>(), c, CB.rank, bgid0.as<int32_t>(),
                                          brun.as<int32_t>(), dout, bfin.as<vid>());
      HIPC(ctx, hipGetLastError());
    }
    HIPC(ctx, hipStreamSynchronize(st()));
  }
  bool any_keep = false;
  for (int i = 0; i < G; ++i) any_keep |= keep[i] != 0;
  for (int i = 0; i < G; ++i) {
    if (!running[i]) continue;
    if (!keep[i]) running[i] = 0;
    else phase[i]++;
  }
  if (!any_keep) return MGL_OK;
  // ---- coarsen the kept graphs (coarsen_graph_mg::coarsen_graph) ----
  HIPC(ctx, hipMemcpyAsync(bkeep.p, keep.data(), 4 * G, hipMemcpyHostToDevice, st()));
  std::vector<int64_t> delta(G, 0);
  int64_t e_keep = 0;
  for (int i = 0; i < G; ++i) {
    if (!keep[i]) continue;
    delta[i] = e_keep - g.eoff[i];
    e_keep += g.eoff[i + 1] - g.eoff[i];
  }
  HIPC(ctx, hipMemcpyAsync(bdelta.p, delta.data(), 8 * G, hipMemcpyHostToDevice, st()));
  HIPC(ctx, hipStreamSynchronize(st()));
  const int nxt = (cur == 1) ? 2 : 1;
  LevelGraph& ng = lg[nxt];
  ng.mode = coarse_mode(g.mode, rec32);
  // the placed rows belong to this level's move kernels; the next level makes
  // its own (make_bins), so they leave room for the coarsening
  bcolr.release();
  bwr.release();
  bpdeg.release();
  // upper bounds: C <= n, E2 <= e_keep
  HIPC(ctx, ng.off.ensure(8 * (g.n + 2)));
  HIPC(ctx, ng.col.ensure(4 * std::max<int64_t>(e_keep, 1)));
  HIPC(ctx, ng.w.ensure(std::max<size_t>(wsize(ng.mode), 4) * std::max<int64_t>(e_keep, 1)));
  if (g.mode != MODE_UNIT) {
    // value ping-pong buffers of the sort, sized for one group
    const int64_t ve = std::max<int64_t>(1, std::min<int64_t>(e_keep, CB.cap));
    HIPC(ctx, cv0.ensure(8 * ve));
    HIPC(ctx, cv1.ensure(8 * ve));
    CB.v0 = cv0.p;
    CB.v1 = cv1.p;
  }
  const int32_t* gidp = g.gid.as<int32_t>();
  const int64_t* deltap = G > 1 ? bdelta.as<int64_t>() : nullptr;
  int64_t C = 0, E2 = 0;
  int rc = coarsen_renumber(g.n, c, gidp, bkeep.as<int32_t>(), CB, &C, st());
  if (rc != 0) FAIL(ctx, MGL_DEVICE_ERROR, "coarsening failed (%d)", rc);
  // per-graph ranges of the coarse union: rank at each graph's first fine vertex
  std::vector<int64_t> idx(G + 1);
  for (int i = 0; i <= G; ++i) idx[i] = g.voff[i];
  int64_t* didx = bgather.as<int64_t>();
  int64_t* dout = didx + (G + 1);
  HIPC(ctx, hipMemcpyAsync(didx, idx.data(), 8 * (G + 1), hipMemcpyHostToDevice, st()));
  k_gather_i64<<<(G + 1 + 255) / 256, 256, 0, st()>>>(CB.rank, didx, G + 1, dout);
  ng.voff.resize(G + 1);
  HIPC(ctx, hipMemcpyAsync(ng.voff.data(), dout, 8 * (G + 1), hipMemcpyDeviceToHost, st()));
  HIPC(ctx, hipStreamSynchronize(st()));
  // groups of consecutive graphs whose kept entries fit the sort's buffers
  int64_t kbase = 0;
  for (int i0 = 0; i0 < G;) {
    int i1 = i0;
    int64_t ek = 0;
    while (i1 < G) {
      const int64_t ke = keep[i1] ? g.eoff[i1 + 1] - g.eoff[i1] : 0;
      if (i1 > i0 && ek + ke > CB.cap) break;
      ek += ke;
      ++i1;
    }
    int64_t U = 0;
    rc = coarsen_group(g.mode, ng.mode, g.voff[i0], g.voff[i1], g.eoff[i0], g.eoff[i1],
                       g.off.as<eid>(), g.col.as<vid>(), g.w.p, gidp, deltap, kbase, ek,
                       ng.voff[i0], ng.voff[i1], E2, CB, ng.off.as<eid>(), ng.col.as<vid>(),
                       ng.w.p, &U, st());
    if (rc != 0) FAIL(ctx, MGL_DEVICE_ERROR, "coarsening of graphs %d..%d failed (%d)", i0, i1 - 1, rc);
    E2 += U;
    kbase += ek;
    i0 = i1;
  }
  HIPC(ctx, hipMemcpyAsync(ng.off.as<eid>() + C, &E2, 8, hipMemcpyHostToDevice, st()));
  HIPC(ctx, hipStreamSynchronize(st()));
  HIPC(ctx, hipGetLastError());
  ng.n = C;
  ng.nnz = E2;
  // entry ranges: coarse offsets at the graph boundaries
  HIPC(ctx, hipMemcpyAsync(didx, ng.voff.data(), 8 * (G + 1), hipMemcpyHostToDevice, st()));
  k_gather_i64<<<(G + 1 + 255) / 256, 256, 0, st()>>>(ng.off.as<int64_t>(), didx, G + 1, dout);
  ng.eoff.resize(G + 1);
  HIPC(ctx, hipMemcpyAsync(ng.eoff.data(), dout, 8 * (G + 1), hipMemcpyDeviceToHost, st()));
  HIPC(ctx, hipStreamSynchronize(st()));
  if (G > 1) {
    HIPC(ctx, ng.gid.ensure(4 * std::max<int64_t>(C, 1)));
    HIPC(ctx, hipMemcpyAsync(didx, ng.voff.data(), 8 * (G + 1), hipMemcpyHostToDevice, st()));
    launch_gid_fill(didx, G, ng.gid.as<int32_t>(), st());
  }
  // labels of continuing graphs move to the coarse ids
  const unsigned g0 = (unsigned)std::min<int64_t>(8192, (n0 + 255) / 256 + 1);
  k_label_next<<<g0, 256, 0, st()>>>(n0, blab.as<vid>(), CB.newc, bgid0.as<int32_t>(),
                                     bkeep.as<int32_t>());
  HIPC(ctx, hipGetLastError());
  cur = nxt;
  return MGL_OK;
}

__global__ void k_local_iota(int64_t n0, const int32_t* gid0, const int64_t* vbase, vid* fin) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n0;
       v += (int64_t)gridDim.x * blockDim.x)
    fin[v] = (vid)(v - vbase[gid0 ? gid0[v] : 0]);
}

mgl_status Engine::run(std::vector<mgl_result*>& out) {
  auto t0 = std::chrono::steady_clock::now();
  Q_host.assign(G, 0.0);
  Q_old.assign(G, -1.0);
  phase.assign(G, 0);
  running.assign(G, 1);
  launch_iota(blab.as<vid>(), n0, st());
  // final labels before any level (max_phases = 0): every vertex on its own
  if (n0 > 0) {
    DBuf dv;
    HIPC(ctx, dv.ensure(8 * (G + 1)));
    HIPC(ctx, hipMemcpyAsync(dv.p, vbase0.data(), 8 * (G + 1), hipMemcpyHostToDevice, st()));
    const unsigned gr = (unsigned)std::min<int64_t>(8192, (n0 + 255) / 256 + 1);
    k_local_iota<<<gr, 256, 0, st()>>>(n0, bgid0.as<int32_t>(), dv.as<int64_t>(),
                                       bfin.as<vid>());
    HIPC(ctx, hipStreamSynchronize(st()));
  }
  int level = 0;
  while (true) {
    bool any = false;
    for (int i = 0; i < G; ++i) {
      // louvain.cu:1473: phase < max_phases && Q - Q_old > thr
      if (running[i] && !(phase[i] < o.max_phases && (Q_host[i] - Q_old[i]) > o.threshold))
        running[i] = 0;
      any |= running[i] != 0;
    }
    if (!any) break;
    for (int i = 0; i < G; ++i)
      if (running[i]) Q_old[i] = Q_host[i];
    const auto tl0 = std::chrono::steady_clock::now();
    mgl_status s = level_start();
    if (s) return s;
    s = run_level();
    if (s) return s;
    HIPC(ctx, hipStreamSynchronize(st()));
    const auto tl1 = std::chrono::steady_clock::now();
    std::vector<size_t> nlev(G);
    for (int i = 0; i < G; ++i) nlev[i] = res[i]->levels.size();
    s = level_end(level);
    if (s) return s;
    const auto tl2 = std::chrono::steady_clock::now();
    for (int i = 0; i < G; ++i)
      if (res[i]->levels.size() > nlev[i]) {
        res[i]->levels.back().seconds = std::chrono::duration<double>(tl1 - tl0).count();
        res[i]->levels.back().coarsen_seconds = std::chrono::duration<double>(tl2 - tl1).count();
      }
    ++level;
  }
  // float modes: the reported final modularity is an FP64 reduction over the
  // input graph with the final labels
  std::vector<double> qfin;
  if (lg[0].mode == MODE_F32 || lg[0].mode == MODE_F64) {
    mgl_status s = final_q_fp64(qfin);
    if (s) return s;
  }
  // final labels
  std::vector<int32_t> fin(n0);
  if (n0) {
    HIPC(ctx, hipMemcpyAsync(fin.data(), bfin.p, 4 * n0, hipMemcpyDeviceToHost, st()));
  }
  HIPC(ctx, hipStreamSynchronize(st()));
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const LevelGraph& l0 = lg[0];
  (void)l0;
  for (int i = 0; i < G; ++i) {
    mgl_result* r = out[i];
    const int64_t b = vbase0[i], e = vbase0[i + 1];
    r->final_assign.assign(fin.begin() + b, fin.begin() + e);
    r->final_q = r->levels.empty() ? 0.0 : (qfin.empty() ? r->levels.back().q : qfin[i]);
    r->seconds = secs;
    r->precision = prec;
  }
  return MGL_OK;
}

// ---------------------------------------------------------------------------
// Host helpers
#include <thread>

template <class F>
static void parallel_for(int64_t n, F f) {
  int nt = (int)std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;
  if (nt > 16) nt = 16;
  if (n < 1 << 16) nt = 1;
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([=]() {
      const int64_t b = n * t / nt, e = n * (t + 1) / nt;
      f(b, e, t);
    });
  for (auto& x : th) x.join();
}

// Input weights: doubles (w64, the reference's weight_t) when given, else
// floats, else unit.
static inline bool has_w(const mgl_csr& g) { return g.w64 != nullptr || g.w != nullptr; }
static inline double wt(const mgl_csr& g, int64_t e) {
  return g.w64 ? g.w64[e] : (g.w ? (double)g.w[e] : 1.0);
}

// Smallest S >= 0 such that every weight * 2^S is an integer, and the sum of
// the weights in *total; -1 if a weight is not finite and positive.
static int dyadic_scale(const mgl_csr& g, long double* total) {
  int nt = 16;
  std::vector<int> sc(nt, 0);
  std::vector<int> bad(nt, 0);
  std::vector<long double> tot(nt, 0);
  parallel_for(g.nnz, [&](int64_t b, int64_t e, int t) {
    int S = 0;
    long double s = 0;
    for (int64_t i = b; i < e; ++i) {
      const double x = wt(g, i);
      if (!(x > 0.0) || !std::isfinite(x)) { bad[t] = 1; return; }
      int ex;
      const double m = std::frexp(x, &ex);
      const uint64_t mi = (uint64_t)std::ldexp(m, 53);
      const int need = 53 - __builtin_ctzll(mi) - ex;
      if (need > S) S = need;
      s += x;
    }
    sc[t] = S;
    tot[t] = s;
  });
  int S = 0;
  long double T = 0;
  for (int t = 0; t < nt; ++t) {
    if (bad[t]) return -1;
    S = std::max(S, sc[t]);
    T += tot[t];
  }
  *total = T;
  return S;
}

__global__ void k_add_i64(int64_t* x, int64_t n, int64_t add) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] += add;
}
__global__ void k_add_i32(int32_t* x, int64_t n, int32_t add) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] += add;
}
__global__ void k_fill_one(int to_mode, void* out, int64_t nnz) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (to_mode == MODE_F64) ((double*)out)[i] = 1.0;
    else ((float*)out)[i] = 1.0f;
  }
}
static unsigned grid_n(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  return (unsigned)g;
}

static mgl_status validate_csr(mgl_ctx* ctx, const mgl_csr* g, int level) {
  if (g->n < 0 || g->nnz < 0) FAIL(ctx, MGL_INVALID_ARG, "negative size");
  if (g->n > 0x7fffffffLL) FAIL(ctx, MGL_UNSUPPORTED, "more than 2^31-1 vertices");
  if (g->n > 0 && (!g->offsets || (g->nnz > 0 && !g->cols)))
    FAIL(ctx, MGL_INVALID_ARG, "null CSR array");
  if (g->n == 0) {
    if (g->nnz != 0) FAIL(ctx, MGL_INVALID_GRAPH, "nnz > 0 with no vertices");
    return MGL_OK;
  }
  if (g->offsets[0] != 0 || g->offsets[g->n] != g->nnz)
    FAIL(ctx, MGL_INVALID_GRAPH, "offsets[0] must be 0 and offsets[n] must equal nnz");
  if (level < 1) return MGL_OK;
  std::vector<int> bad(16, 0);
  const int64_t n = g->n;
  parallel_for(n, [&](int64_t b, int64_t e, int t) {
    for (int64_t v = b; v < e; ++v) {
      if (g->offsets[v + 1] < g->offsets[v]) { bad[t] = 1; return; }
    }
  });
  for (int x : bad)
    if (x) FAIL(ctx, MGL_INVALID_GRAPH, "offsets are not non-decreasing");
  parallel_for(g->nnz, [&](int64_t b, int64_t e, int t) {
    for (int64_t i = b; i < e; ++i) {
      const int32_t x = g->cols[i];
      if (x < 0 || x >= n) { bad[t] = 2; return; }
      if (has_w(*g)) {
        const double x = wt(*g, i);
        if (!(x >= 0.0 && std::isfinite(x))) { bad[t] = 3; return; }
      }
    }
  });
  for (int x : bad) {
    if (x == 2) FAIL(ctx, MGL_INVALID_GRAPH, "column index out of range");
    if (x == 3) FAIL(ctx, MGL_INVALID_GRAPH, "edge weight negative or not finite");
  }
  if (level >= 2) {
    // symmetry: every (u,v,w) has a matching (v,u,w)
    std::vector<int> asym(16, 0);
    parallel_for(n, [&](int64_t b, int64_t e, int t) {
      for (int64_t u = b; u < e && !asym[t]; ++u)
        for (int64_t i = g->offsets[u]; i < g->offsets[u + 1]; ++i) {
          const int32_t v = g->cols[i];
          bool found = false;
          for (int64_t j = g->offsets[v]; j < g->offsets[v + 1]; ++j)
            if (g->cols[j] == (int32_t)u && wt(*g, j) == wt(*g, i)) { found = true; break; }
          if (!found) { asym[t] = 1; break; }
        }
    });
    for (int x : asym)
      if (x) FAIL(ctx, MGL_INVALID_GRAPH, "graph is not symmetric");
  }
  return MGL_OK;
}

mgl_status Engine::setup_union(const mgl_csr* graphs, int ng) {
  G = ng;
  if (G < 1) FAIL(ctx, MGL_INVALID_ARG, "no graphs");
  for (int i = 0; i < G; ++i) {
    mgl_status s = validate_csr(ctx, &graphs[i], o.validate);
    if (s) return s;
  }
  bool weighted = false;
  for (int i = 0; i < G; ++i) weighted |= has_w(graphs[i]);
  prec = o.precision;
  int S = 0;
  if (weighted && (prec == MGL_PREC_AUTO || prec == MGL_PREC_EXACT)) {
    // one scale 2^S for the batch; every graph's scaled total (unit weights
    // count 2^S each) must stay below 2^62
    std::vector<long double> tot(G, 0.0L);
    for (int i = 0; i < G && S >= 0; ++i) {
      if (!has_w(graphs[i])) {
        tot[i] = (long double)graphs[i].nnz;
        continue;
      }
      const int si = dyadic_scale(graphs[i], &tot[i]);
      S = si < 0 ? -1 : std::max(S, si);
    }
    for (int i = 0; i < G && S >= 0; ++i)
      if (tot[i] * std::ldexp(1.0L, S) > std::ldexp(1.0L, 62)) S = -1;
    if (S < 0) {
      if (prec == MGL_PREC_EXACT)
        FAIL(ctx, MGL_UNSUPPORTED,
             "EXACT precision needs positive weights that are integer multiples of 2^-S "
             "with every graph's total below 2^(62-S)");
      prec = MGL_PREC_FP64;
      S = 0;
    }
  }
  if (prec == MGL_PREC_AUTO) prec = MGL_PREC_EXACT;
  int mode0;
  if (prec == MGL_PREC_EXACT) mode0 = weighted ? MODE_INT : MODE_UNIT;
  else if (prec == MGL_PREC_FP64) mode0 = MODE_F64;
  else if (prec == MGL_PREC_FP32_HASH) mode0 = MODE_F32;
  else FAIL(ctx, MGL_INVALID_ARG, "unknown precision %d", o.precision);
  scale = S;
  // union sizes
  vbase0.assign(G + 1, 0);
  std::vector<int64_t> eb(G + 1, 0);
  for (int i = 0; i < G; ++i) {
    vbase0[i + 1] = vbase0[i] + graphs[i].n;
    eb[i + 1] = eb[i] + graphs[i].nnz;
  }
  n0 = vbase0[G];
  nnz0 = eb[G];
  max_graph_nnz0 = 0;
  for (int i = 0; i < G; ++i) max_graph_nnz0 = std::max<int64_t>(max_graph_nnz0, graphs[i].nnz);
  if (n0 > 0x7fffffffLL) FAIL(ctx, MGL_UNSUPPORTED, "batch has more than 2^31-1 vertices");
  // per-graph 2m (host_graph.cpp:112-114)
  hgs.assign(G, GState{});
  for (int i = 0; i < G; ++i) {
    const mgl_csr& gr = graphs[i];
    GState& s = hgs[i];
    if (prec == MGL_PREC_EXACT) {
      if (!has_w(gr)) {
        s.mass_i = gr.nnz * ((int64_t)1 << S);  // unit weights carry 2^S in a weighted batch
      } else {
        std::vector<int64_t> part(16, 0);
        const int sc = S;
        parallel_for(gr.nnz, [&](int64_t b, int64_t e, int t) {
          int64_t acc = 0;
          for (int64_t j = b; j < e; ++j) acc += (int64_t)std::ldexp(wt(gr, j), sc);
          part[t] = acc;
        });
        for (auto x : part) s.mass_i += x;
      }
      s.mass = (double)s.mass_i;
    } else {
      // FP64 sum in entry order (the reference's thrust::reduce over the
      // weights, host_graph.cpp:112-114)
      double m = 0;
      if (!has_w(gr)) m = (double)gr.nnz;
      else
        for (int64_t j = 0; j < gr.nnz; ++j) m += wt(gr, j);
      s.mass = m;
    }
  }
  // 8-byte records hold community totals in int32: every total is <= 2m
  rec32 = (prec == MGL_PREC_EXACT);
  for (int i = 0; i < G; ++i) rec32 = rec32 && hgs[i].mass_i < ((int64_t)1 << 31);
  // device union graph
  LevelGraph& g = lg[0];
  cur = 0;
  g.mode = mode0;
  g.n = n0;
  g.nnz = nnz0;
  g.voff = vbase0;
  g.eoff = eb;
  HIPC(ctx, g.off.ensure(8 * (n0 + 1)));
  HIPC(ctx, g.col.ensure(4 * std::max<int64_t>(nnz0, 1)));
  for (int i = 0; i < G; ++i) {
    const mgl_csr& gr = graphs[i];
    if (gr.n > 0) {
      HIPC(ctx, hipMemcpyAsync(g.off.as<int64_t>() + vbase0[i], gr.offsets, 8 * gr.n,
                               hipMemcpyHostToDevice, st()));
      if (eb[i]) k_add_i64<<<grid_n(gr.n), 256, 0, st()>>>(g.off.as<int64_t>() + vbase0[i], gr.n, eb[i]);
    }
    if (gr.nnz > 0) {
      HIPC(ctx, hipMemcpyAsync(g.col.as<int32_t>() + eb[i], gr.cols, 4 * gr.nnz,
                               hipMemcpyHostToDevice, st()));
      if (vbase0[i])
        k_add_i32<<<grid_n(gr.nnz), 256, 0, st()>>>(g.col.as<int32_t>() + eb[i], gr.nnz,
                                                   (int32_t)vbase0[i]);
    }
  }
  HIPC(ctx, hipMemcpyAsync(g.off.as<int64_t>() + n0, &nnz0, 8, hipMemcpyHostToDevice, st()));
  if (mode0 != MODE_UNIT) {
    HIPC(ctx, g.w.ensure(wsize(mode0) * std::max<int64_t>(nnz0, 1)));
    // input weights are staged one graph at a time (stream order keeps a
    // graph's cast ahead of the next graph's copy into the same buffer)
    DBuf stage;
    int64_t smax = 1;
    for (int i = 0; i < G; ++i)
      if (has_w(graphs[i])) smax = std::max<int64_t>(smax, graphs[i].nnz);
    if (weighted) HIPC(ctx, stage.ensure(8 * smax));
    for (int i = 0; i < G; ++i) {
      const mgl_csr& gr = graphs[i];
      if (gr.nnz == 0) continue;
      if (gr.w64) {
        HIPC(ctx, hipMemcpyAsync(stage.p, gr.w64, 8 * gr.nnz, hipMemcpyHostToDevice, st()));
        launch_cast_weights_f64(mode0, stage.as<double>(), (char*)g.w.p + wsize(mode0) * eb[i],
                                gr.nnz, S, st());
        continue;
      } else if (gr.w) {
        HIPC(ctx, hipMemcpyAsync(stage.p, gr.w, 4 * gr.nnz, hipMemcpyHostToDevice, st()));
      } else if (mode0 == MODE_INT) {
        // unit weights of an unweighted graph inside a weighted batch
        std::vector<int64_t> one(gr.nnz, (int64_t)1 << S);
        HIPC(ctx, hipMemcpyAsync(g.w.as<int64_t>() + eb[i], one.data(), 8 * gr.nnz,
                                 hipMemcpyHostToDevice, st()));
        HIPC(ctx, hipStreamSynchronize(st()));
        continue;
      } else {
        k_fill_one<<<grid_n(gr.nnz), 256, 0, st()>>>(mode0, (char*)g.w.p + wsize(mode0) * eb[i], gr.nnz);
        continue;
      }
      launch_cast_weights(mode0, stage.as<float>(), (char*)g.w.p + wsize(mode0) * eb[i],
                          gr.nnz, S, st());
    }
    HIPC(ctx, hipStreamSynchronize(st()));
  }
  if (G > 1) {
    HIPC(ctx, g.gid.ensure(4 * std::max<int64_t>(n0, 1)));
    HIPC(ctx, bgid0.ensure(4 * std::max<int64_t>(n0, 1)));
    DBuf dv;
    HIPC(ctx, dv.ensure(8 * (G + 1)));
    HIPC(ctx, hipMemcpyAsync(dv.p, vbase0.data(), 8 * (G + 1), hipMemcpyHostToDevice, st()));
    launch_gid_fill(dv.as<int64_t>(), G, g.gid.as<int32_t>(), st());
    launch_copy_vid(g.gid.as<int32_t>(), bgid0.as<int32_t>(), n0, st());
    HIPC(ctx, hipStreamSynchronize(st()));
  }
  mgl_status s = alloc_state();
  if (s) return s;
  HIPC(ctx, hipMemcpyAsync(bgs.p, hgs.data(), sizeof(GState) * G, hipMemcpyHostToDevice, st()));
  HIPC(ctx, hipStreamSynchronize(st()));
  HIPC(ctx, hipGetLastError());
  return MGL_OK;
}

// ===========================================================================
// C ABI
extern "C" {

void mgl_opts_default(mgl_opts* o) {
  if (!o) return;
  o->version = 2;
  o->threshold = 1.0e-6;
  o->max_phases = 1000;
  o->max_iter = 0;
  o->precision = MGL_PREC_AUTO;
  o->tie_rule = MGL_TIE_REFERENCE;
  o->fuse_modularity = 1;
  o->keep_levels = 0;
  o->keep_graphs = 0;
  o->validate = 1;
  o->tuning_file = nullptr;
  o->use_graph = 1;
  o->record_layout = MGL_LAYOUT_AUTO;
}

int mgl_version_major(void) { return 1; }

static mgl_status load_tuning(mgl_ctx* ctx, const char* path, Tuning& tu);
}  // extern "C"
int mgl_noalloc_setup(int device);  // mgl_modularity.hip
extern "C" {

mgl_status mgl_create(mgl_ctx** out, int device) {
  if (!out) return MGL_INVALID_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return MGL_DEVICE_ERROR;
  }
  if (device < 0 || device >= ndev) return MGL_INVALID_ARG;
  mgl_ctx* ctx = new mgl_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->st) != hipSuccess ||
      hipStreamCreate(&ctx->st2) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->evfork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->evjoin, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    if (ctx->evjoin) (void)hipEventDestroy(ctx->evjoin);
    if (ctx->evfork) (void)hipEventDestroy(ctx->evfork);
    if (ctx->st2) (void)hipStreamDestroy(ctx->st2);
    if (ctx->st) (void)hipStreamDestroy(ctx->st);
    delete ctx;
    return MGL_DEVICE_ERROR;
  }
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) ==
          hipSuccess &&
      lds > 0)
    ctx->lds_max = lds;
  // the allocation-free entry point's working memory on this device
  (void)mgl_noalloc_setup(device);
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess &&
      khz > 0)
    ctx->clock_hz = 1000.0 * khz;
  (void)hipGetLastError();
  move_kernels_allow_lds(ctx->lds_max);
  hub_allow_lds(ctx->lds_max);
  (void)hipGetLastError();
  *out = ctx;
  return MGL_OK;
}

void mgl_destroy(mgl_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->evjoin) (void)hipEventDestroy(ctx->evjoin);
  if (ctx->evfork) (void)hipEventDestroy(ctx->evfork);
  if (ctx->st2) (void)hipStreamDestroy(ctx->st2);
  if (ctx->st) (void)hipStreamDestroy(ctx->st);
  delete ctx;
}

const char* mgl_last_error(const mgl_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }
}  // extern "C"

// context access for the other host units of the library
void mgl_set_error(mgl_ctx* ctx, const std::string& msg) { ctx->err = msg; }
int mgl_ctx_device(const mgl_ctx* ctx) { return ctx->device; }
hipStream_t mgl_ctx_stream(const mgl_ctx* ctx) { return ctx->st; }

extern "C" {

static mgl_status check_opts(mgl_ctx* ctx, const mgl_opts* o) {
  if (o->version < 0 || o->version > 2) FAIL(ctx, MGL_INVALID_ARG, "version must be 0, 1 or 2");
  if (!(o->threshold >= 0) || !std::isfinite(o->threshold))
    FAIL(ctx, MGL_INVALID_ARG, "threshold must be finite and >= 0");
  if (o->max_phases < 0) FAIL(ctx, MGL_INVALID_ARG, "max_phases must be >= 0");
  if (o->tie_rule != MGL_TIE_REFERENCE && o->tie_rule != MGL_TIE_SMALLEST)
    FAIL(ctx, MGL_INVALID_ARG, "unknown tie rule");
  if (o->record_layout < MGL_LAYOUT_AUTO || o->record_layout > MGL_LAYOUT_BINNED)
    FAIL(ctx, MGL_INVALID_ARG, "unknown record layout");
  return MGL_OK;
}

mgl_status mgl_louvain_batched(mgl_ctx* ctx, const mgl_csr* graphs, int32_t n_graphs,
                               const mgl_opts* opts, mgl_result** results) {
  if (!ctx) return MGL_INVALID_ARG;
  if (!graphs || !results || n_graphs < 1) FAIL(ctx, MGL_INVALID_ARG, "null argument");
  HIPC(ctx, hipSetDevice(ctx->device));
  mgl_opts o;
  if (opts) o = *opts;
  else mgl_opts_default(&o);
  mgl_status s = check_opts(ctx, &o);
  if (s) return s;
  for (int i = 0; i < n_graphs; ++i) results[i] = nullptr;
  Engine E(ctx);
  E.o = o;
  if (o.tuning_file) {
    s = load_tuning(ctx, o.tuning_file, E.tune);
    if (s) return s;
  }
  s = E.setup_union(graphs, n_graphs);
  if (s) return s;
  s = E.alloc_coarsen();
  if (s) return s;
  std::vector<mgl_result*> out(n_graphs);
  for (int i = 0; i < n_graphs; ++i) out[i] = new mgl_result();
  E.res = out;
  s = E.run(out);
  if (s) {
    for (auto* r : out) delete r;
    return s;
  }
  for (int i = 0; i < n_graphs; ++i) results[i] = out[i];
  return MGL_OK;
}

mgl_status mgl_louvain(mgl_ctx* ctx, const mgl_csr* g, const mgl_opts* o, mgl_result** out) {
  if (!ctx) return MGL_INVALID_ARG;
  if (!g || !out) FAIL(ctx, MGL_INVALID_ARG, "null argument");
  return mgl_louvain_batched(ctx, g, 1, o, out);
}

int mgl_result_levels(const mgl_result* r) { return r ? (int)r->levels.size() : 0; }
int64_t mgl_result_level_n(const mgl_result* r, int l) {
  return (r && l >= 0 && l < (int)r->levels.size()) ? r->levels[l].n : -1;
}
double mgl_result_level_q(const mgl_result* r, int l) {
  return (r && l >= 0 && l < (int)r->levels.size()) ? r->levels[l].q : NAN;
}
int mgl_result_level_iters(const mgl_result* r, int l) {
  return (r && l >= 0 && l < (int)r->levels.size()) ? r->levels[l].iters : -1;
}
const double* mgl_result_level_qtraj(const mgl_result* r, int l) {
  return (r && l >= 0 && l < (int)r->levels.size()) ? r->levels[l].qtraj.data() : nullptr;
}
double mgl_result_level_q0(const mgl_result* r, int l) {
  return (r && l >= 0 && l < (int)r->levels.size()) ? r->levels[l].q0 : NAN;
}
const double* mgl_result_level_iter_seconds(const mgl_result* r, int l) {
  if (!r || l < 0 || l >= (int)r->levels.size() || r->levels[l].iter_secs.empty()) return nullptr;
  return r->levels[l].iter_secs.data();
}
const int32_t* mgl_result_level_assign(const mgl_result* r, int l) {
  if (!r || l < 0 || l >= (int)r->levels.size() || r->levels[l].assign.size() != (size_t)r->levels[l].n)
    return nullptr;
  return r->levels[l].assign.data();
}
mgl_status mgl_result_level_graph(const mgl_result* r, int l, int64_t* n, int64_t* nnz,
                                  const int64_t** off, const int32_t** col, const double** w) {
  if (!r || l < 1 || l >= (int)r->levels.size() || !r->levels[l].has_graph)
    return MGL_INVALID_ARG;
  const LevelOut& lo = r->levels[l];
  if (n) *n = lo.gn;
  if (nnz) *nnz = lo.gnnz;
  if (off) *off = lo.goff.data();
  if (col) *col = lo.gcol.data();
  if (w) *w = lo.gw.data();
  return MGL_OK;
}
int64_t mgl_result_vertices(const mgl_result* r) {
  return r ? (int64_t)r->final_assign.size() : 0;
}
const int32_t* mgl_result_final_assign(const mgl_result* r) {
  return r ? r->final_assign.data() : nullptr;
}
double mgl_result_final_q(const mgl_result* r) { return r ? r->final_q : NAN; }
double mgl_result_level_seconds(const mgl_result* r, int l, double* coarsen_seconds) {
  if (!r || l < 0 || l >= (int)r->levels.size()) return NAN;
  if (coarsen_seconds) *coarsen_seconds = r->levels[l].coarsen_seconds;
  return r->levels[l].seconds;
}
int mgl_result_precision(const mgl_result* r) { return r ? r->precision : -1; }
double mgl_result_seconds(const mgl_result* r) { return r ? r->seconds : NAN; }
void mgl_result_free(mgl_result* r) { delete r; }

int mgl_bin_bounds(const mgl_ctx* ctx, int64_t* upper, int max_bins) {
  if (!ctx) return 0;
  BinLayout L = layout_for(ctx, ctx->tune, MODE_UNIT, true);
  for (int i = 0; i < L.nb && i < max_bins; ++i) upper[i] = L.b[i].upper;
  return L.nb;
}

int mgl_device_lds_bytes(const mgl_ctx* ctx) { return ctx ? ctx->lds_max : 0; }

mgl_status mgl_memory_stats(int64_t* live_bytes, int64_t* peak_bytes, int reset_peak) {
  if (live_bytes) *live_bytes = g_dev_live.load();
  if (peak_bytes) *peak_bytes = g_dev_peak.load();
  if (reset_peak) g_dev_peak.store(g_dev_live.load());
  return MGL_OK;
}

// Streaming copy of `bytes` between two device buffers with the library's
// copy kernel (k_stream_copy), `iters` times; returns the best GB/s. Used to
// calibrate HBM byte counters against a transfer of known size.
__global__ __launch_bounds__(256) void k_stream_copy(const int4* __restrict__ s, int4* __restrict__ d,
                                                     int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    d[i] = s[i];
}
mgl_status mgl_membench(mgl_ctx* ctx, int64_t bytes, int iters, double* gbps) {
  if (!ctx) return MGL_INVALID_ARG;
  if (bytes < 16 || iters < 1 || !gbps) FAIL(ctx, MGL_INVALID_ARG, "bad membench arguments");
  HIPC(ctx, hipSetDevice(ctx->device));
  DBuf a, b;
  HIPC(ctx, a.ensure(bytes));
  HIPC(ctx, b.ensure(bytes));
  launch_fill_bytes0(a.p, bytes, ctx->st);
  const int64_t n = bytes / 16;
  hipEvent_t e0, e1;
  HIPC(ctx, hipEventCreate(&e0));
  HIPC(ctx, hipEventCreate(&e1));
  double best = 0;
  for (int i = 0; i < iters; ++i) {
    HIPC(ctx, hipEventRecord(e0, ctx->st));
    k_stream_copy<<<2048, 256, 0, ctx->st>>>(a.as<int4>(), b.as<int4>(), n);
    HIPC(ctx, hipEventRecord(e1, ctx->st));
    HIPC(ctx, hipEventSynchronize(e1));
    float ms = 0;
    HIPC(ctx, hipEventElapsedTime(&ms, e0, e1));
    best = std::max(best, 2.0 * (double)(n * 16) / (ms * 1e-3) / 1e9);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *gbps = best;
  return MGL_OK;
}

}  // extern "C"

// ===========================================================================
// Sessions: one graph resident in HBM, driven pass by pass at level 0.
struct mgl_session {
  mgl_ctx* ctx = nullptr;
  Engine* E = nullptr;
  int it = 0;
  bool coarsen_ready = false;
  bool restart = false;  // the last pass rejected: the next pass restarts the level's loop
  mgl_pass_stats st{};
  int64_t bin_e[MGL_MAX_BINS] = {};  // CSR entries of each bin's vertices (statistics)
  // counters of the move of pass count_it, from mgl_session_count
  int count_it = -1;
  unsigned long long cnt[MGL_MAX_BINS + 3] = {};
};

// CSR entries of the vertices perm[0 .. nb): one atomic per wavefront.
__global__ __launch_bounds__(256) void k_bin_entries(const eid* __restrict__ off,
                                                     const vid* __restrict__ perm, int64_t nb,
                                                     unsigned long long* out) {
  int64_t acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nb; i += (int64_t)gridDim.x * 256) {
    const vid v = perm[i];
    acc += off[v + 1] - off[v];
  }
  acc = group_sum<64>(acc);
  if ((threadIdx.x & 63) == 0 && acc) atomicAdd(out, (unsigned long long)acc);
}

// Per-bin entry counts of the session's level (statistics only; computed when
// the bins are made, outside any timed pass).
static mgl_status session_bin_entries(mgl_session* s) {
  Engine& E = *s->E;
  mgl_ctx* ctx = s->ctx;
  DBuf d;
  HIPC(ctx, d.ensure(8 * MGL_MAX_BINS));
  launch_fill_bytes0(d.p, 8 * MGL_MAX_BINS, ctx->st);
  for (int b = 0; b < E.L.nb; ++b)
    if (E.bin_n[b] > 0)
      k_bin_entries<<<grid_n(E.bin_n[b]), 256, 0, ctx->st>>>(
          E.LG().off.as<eid>(), E.bperm.as<vid>() + E.bin_off[b], E.bin_n[b],
          d.as<unsigned long long>() + b);
  HIPC(ctx, hipGetLastError());
  unsigned long long h[MGL_MAX_BINS];
  HIPC(ctx, hipMemcpyAsync(h, d.p, 8 * MGL_MAX_BINS, hipMemcpyDeviceToHost, ctx->st));
  HIPC(ctx, hipStreamSynchronize(ctx->st));
  for (int b = 0; b < MGL_MAX_BINS; ++b) s->bin_e[b] = b < E.L.nb ? (int64_t)h[b] : 0;
  return MGL_OK;
}

__global__ void k_reactivate(GState* gs) {
  GState& s = gs[0];
  s.active = 1;
  s.fix = 0;
  s.sin_i = 0; s.sin_d = 0; s.sa2_lo = 0; s.sa2_hi = 0; s.sa2_d = 0;
}
__global__ void k_zero_stats(GState* gs, unsigned long long* dist) {
  if (threadIdx.x == 0) gs[0].moved = 0;
  dist[threadIdx.x] = 0;
  if (threadIdx.x < 3) dist[2 * MGL_MAX_BINS + threadIdx.x] = 0;  // LDS probe counters
}

static mgl_status session_prime(mgl_session* s) {
  // fused mode needs the first proposal before the first decided pass
  Engine& E = *s->E;
  if (!E.o.fuse_modularity) return MGL_OK;
  const int up = ((s->it + 1) & 1) ? 1 : 0;
  mgl_status r = E.move(E.c, E.p, up, nullptr, false);
  if (r) return r;
  LevelGraph& g = E.LG();
  launch_update(g.mode, g.n, E.c, E.p, E.bk.p, E.ba.p, nullptr, E.bgs.as<GState>(), E.st());
  launch_reduce(g.mode, g.n, nullptr, E.ba.p, nullptr, E.bgs.as<GState>(), E.st());
  launch_decide(1, E.bgs.as<GState>(), E.o.threshold, E.o.max_iter, 1, mode_exact(g.mode),
                nullptr, 0, 1, E.st());
  HIPC(s->ctx, hipGetLastError());
  return MGL_OK;
}

extern "C" {

mgl_status mgl_session_create(mgl_ctx* ctx, const mgl_csr* g, const mgl_opts* opts,
                              mgl_session** out) {
  if (!ctx) return MGL_INVALID_ARG;
  if (!g || !out) FAIL(ctx, MGL_INVALID_ARG, "null argument");
  HIPC(ctx, hipSetDevice(ctx->device));
  *out = nullptr;
  mgl_opts o;
  if (opts) o = *opts;
  else mgl_opts_default(&o);
  mgl_status s = check_opts(ctx, &o);
  if (s) return s;
  mgl_session* ss = new mgl_session();
  ss->ctx = ctx;
  ss->E = new Engine(ctx);
  ss->E->o = o;
  if (o.tuning_file) {
    s = load_tuning(ctx, o.tuning_file, ss->E->tune);
    if (s) { delete ss->E; delete ss; return s; }
  }
  s = ss->E->setup_union(g, 1);
  if (s) { delete ss->E; delete ss; return s; }
  ss->E->running.assign(1, 1);
  ss->E->Q_host.assign(1, 0.0);
  ss->E->Q_old.assign(1, -1.0);
  ss->E->phase.assign(1, 0);
  s = ss->E->level_start();
  if (!s) s = session_bin_entries(ss);
  if (s) { delete ss->E; delete ss; return s; }
  HIPC(ctx, hipStreamSynchronize(ctx->st));
  *out = ss;
  return MGL_OK;
}

void mgl_session_destroy(mgl_session* s) {
  if (!s) return;
  (void)hipSetDevice(s->ctx->device);
  delete s->E;
  delete s;
}

mgl_status mgl_session_reset(mgl_session* s, double* q0) {
  if (!s) return MGL_INVALID_ARG;
  mgl_ctx* ctx = s->ctx;
  HIPC(ctx, hipSetDevice(ctx->device));
  Engine& E = *s->E;
  mgl_status r = E.level_start();
  if (r) return r;
  r = session_bin_entries(s);
  if (r) return r;
  s->it = 0;
  s->restart = false;
  r = E.read_gs();
  if (r) return r;
  if (q0) *q0 = E.hgs[0].q_cur;
  r = session_prime(s);
  if (r) return r;
  HIPC(ctx, hipStreamSynchronize(ctx->st));
  return MGL_OK;
}

static_assert(MGL_MAX_BINS <= 16, "mgl_pass_stats.bin_ms holds 16 bins");

mgl_status mgl_session_pass(mgl_session* s, double* q, int* accepted) {
  if (!s) return MGL_INVALID_ARG;
  mgl_ctx* ctx = s->ctx;
  HIPC(ctx, hipSetDevice(ctx->device));
  Engine& E = *s->E;
  mgl_status r;
  if (s->restart) {
    // the previous pass rejected its proposal, which ends a level: the loop
    // restarts from the committed assignment with rebuilt community totals
    k_reactivate<<<1, 1, 0, ctx->st>>>(E.bgs.as<GState>());
    r = E.recompute_a(E.c);
    if (r) return r;
    r = session_prime(s);
    if (r) return r;
    s->restart = false;
  }
  unsigned long long* dist = E.bcounts.as<unsigned long long>() + 2 * MGL_MAX_BINS;
  k_zero_stats<<<1, MGL_MAX_BINS, 0, ctx->st>>>(E.bgs.as<GState>(), dist);
  E.timing = true;
  s->it += 1;
  r = E.iteration(s->it);
  E.timing = false;
  if (r) return r;
  r = E.read_gs(true);
  if (r) return r;
  const GState& gs = E.hgs[0];
  const bool acc = E.o.fuse_modularity ? (gs.fix == 0 || gs.fix == 2) && gs.fix != 1
                                       : (gs.fix == 3);
  if (accepted) *accepted = acc ? 1 : 0;
  if (q) *q = gs.q_cur;
  // statistics of the pass
  mgl_pass_stats& st = s->st;
  memset(&st, 0, sizeof(st));
  st.n = E.LG().n;
  st.nnz = E.LG().nnz;
  float ms = 0;
  hipEvent_t* ev = E.ev;
  const int B = 2 * MGL_MAX_BINS;
  HIPC(ctx, hipEventElapsedTime(&ms, ev[B + 0], ev[B + 1]));
  st.move_ms = ms;
  HIPC(ctx, hipEventElapsedTime(&ms, ev[B + 1], ev[B + 2]));
  st.update_ms = ms;
  HIPC(ctx, hipEventElapsedTime(&ms, ev[B + 2], ev[B + 3]));
  st.modularity_ms = ms;
  HIPC(ctx, hipEventElapsedTime(&ms, ev[B + 0], ev[B + 4]));
  st.pass_ms = ms;
  st.dominant_ms = -1;
  // counters of this pass's move, if mgl_session_count ran on the same state
  unsigned long long dh[MGL_MAX_BINS + 3] = {};
  if (s->count_it == s->it) memcpy(dh, s->cnt, sizeof(dh));
  // the dominant launch among bins served by one kernel that ran alone on the
  // device (a bin beside the hub path shares the CUs, so its span is not its
  // own time); the hub path apart
  int hub_bin = -1;
  st.n_bins = E.L.nb;
  for (int b = E.L.nb - 1; b >= 1; --b) {
    if (E.bin_n[b] == 0) continue;
    HIPC(ctx, hipEventElapsedTime(&ms, ev[2 * b], ev[2 * b + 1]));
    st.bin_ms[b] = ms;
    if (E.L.b[b].shape == SHAPE_GLOBAL) {
      st.hub_ms = ms;
      hub_bin = b;
      continue;
    }
    if (b <= E.shared_from) continue;
    if (ms > st.dominant_ms) {
      st.dominant_ms = ms;
      st.dominant_bin = b;
    }
  }
  if (st.dominant_ms >= 0) {
    const int b = st.dominant_bin;
    st.dominant_vertices = E.bin_n[b];
    st.dominant_entries = s->bin_e[b];
    st.dominant_distinct = (int64_t)dh[b];
    move_kernel_name(E.LG().mode, E.rec32_now() ? 1 : 0, E.L.b[b], st.dominant_kernel,
                     (int)sizeof(st.dominant_kernel));
  }
  if (hub_bin >= 0) {
    st.hub_vertices = E.bin_n[hub_bin];
    st.hub_entries = s->bin_e[hub_bin];
    st.hub_distinct = (int64_t)dh[hub_bin];
  }
  st.distinct = 0;
  for (int b = 0; b < MGL_MAX_BINS; ++b) st.distinct += (int64_t)dh[b];
  st.lds_inserts = (int64_t)dh[MGL_MAX_BINS];
  st.lds_probes = (int64_t)dh[MGL_MAX_BINS + 1];
  st.lds_multi_probe = (int64_t)dh[MGL_MAX_BINS + 2];
  st.moved = gs.moved;
  st.q = gs.q_cur;
  st.hub_beside = hub_bin >= 0 ? E.shared_from : 0;
  // a rejected proposal ends the level; the next pass restarts the loop (a
  // caller that times passes can call mgl_session_reset in between instead)
  s->restart = !gs.active;
  return MGL_OK;
}

mgl_status mgl_session_count(mgl_session* s) {
  if (!s) return MGL_INVALID_ARG;
  mgl_ctx* ctx = s->ctx;
  HIPC(ctx, hipSetDevice(ctx->device));
  Engine& E = *s->E;
  // the move of the next pass: same input, direction and output buffer, so
  // nothing the pass reads is changed
  const int it = s->it + 1;
  const bool fused = E.o.fuse_modularity != 0;
  const int up = fused ? (((it + 1) & 1) ? 1 : 0) : ((it & 1) ? 1 : 0);
  const vid* from = fused ? E.p : E.c;
  vid* to = fused ? E.nx : E.p;
  unsigned long long* dist = E.bcounts.as<unsigned long long>() + 2 * MGL_MAX_BINS;
  launch_fill_bytes0(dist, 8 * MGL_MAX_BINS, ctx->st);
  launch_fill_bytes0(dist + 2 * MGL_MAX_BINS, 8 * 3, ctx->st);
  E.counting = true;
  mgl_status r = E.move(from, to, up, nullptr, false);
  E.counting = false;
  if (r) return r;
  HIPC(ctx, hipMemcpyAsync(s->cnt, dist, 8 * MGL_MAX_BINS, hipMemcpyDeviceToHost, ctx->st));
  HIPC(ctx, hipMemcpyAsync(s->cnt + MGL_MAX_BINS, dist + 2 * MGL_MAX_BINS, 8 * 3,
                           hipMemcpyDeviceToHost, ctx->st));
  HIPC(ctx, hipStreamSynchronize(ctx->st));
  s->count_it = it;
  return MGL_OK;
}

mgl_status mgl_session_stats(const mgl_session* s, mgl_pass_stats* st) {
  if (!s || !st) return MGL_INVALID_ARG;
  *st = s->st;
  return MGL_OK;
}

mgl_status mgl_session_bins(mgl_session* s, int64_t* sizes, int* nb) {
  if (!s || !sizes || !nb) return MGL_INVALID_ARG;
  Engine& E = *s->E;
  const int cap = *nb;
  *nb = E.L.nb;
  for (int i = 0; i < E.L.nb && i < cap; ++i) sizes[i] = E.bin_n[i];
  return MGL_OK;
}

mgl_status mgl_session_set_assign(mgl_session* s, const int32_t* c) {
  if (!s || !c) return MGL_INVALID_ARG;
  mgl_ctx* ctx = s->ctx;
  HIPC(ctx, hipSetDevice(ctx->device));
  Engine& E = *s->E;
  const int64_t n = E.LG().n;
  for (int64_t v = 0; v < n; ++v)
    if (c[v] < 0 || c[v] >= n) FAIL(ctx, MGL_INVALID_ARG, "community id out of range");
  HIPC(ctx, hipMemcpyAsync(E.c, c, 4 * n, hipMemcpyHostToDevice, ctx->st));
  launch_copy_vid(E.c, E.p, n, ctx->st);
  launch_copy_vid(E.c, E.nx, n, ctx->st);
  k_reactivate<<<1, 1, 0, ctx->st>>>(E.bgs.as<GState>());
  mgl_status r = E.recompute_a(E.c);
  if (r) return r;
  HIPC(ctx, hipStreamSynchronize(ctx->st));
  return MGL_OK;
}

mgl_status mgl_session_move(mgl_session* s, int up, int32_t* c_new) {
  if (!s || !c_new) return MGL_INVALID_ARG;
  mgl_ctx* ctx = s->ctx;
  HIPC(ctx, hipSetDevice(ctx->device));
  Engine& E = *s->E;
  const int64_t n = E.LG().n;
  k_reactivate<<<1, 1, 0, ctx->st>>>(E.bgs.as<GState>());
  launch_copy_vid(E.c, E.p, n, ctx->st);  // isolated vertices keep their community
  mgl_status r = E.move(E.c, E.p, up ? 1 : 0, nullptr, false);
  if (r) return r;
  r = E.check_err();
  if (r) return r;
  HIPC(ctx, hipMemcpyAsync(c_new, E.p, 4 * n, hipMemcpyDeviceToHost, ctx->st));
  HIPC(ctx, hipStreamSynchronize(ctx->st));
  return MGL_OK;
}

mgl_status mgl_session_modularity(mgl_session* s, double* q) {
  if (!s || !q) return MGL_INVALID_ARG;
  mgl_ctx* ctx = s->ctx;
  HIPC(ctx, hipSetDevice(ctx->device));
  Engine& E = *s->E;
  LevelGraph& g = E.LG();
  k_reactivate<<<1, 1, 0, ctx->st>>>(E.bgs.as<GState>());
  launch_fill_bytes0(E.bvin.p, 8 * (size_t)std::max<int64_t>(g.n, 1), ctx->st);
  mgl_status r = E.rows(OP_INTRA, E.c, nullptr, E.bvin.p, true);
  if (r) return r;
  launch_reduce(g.mode, g.n, E.bvin.p, E.ba.p, nullptr, E.bgs.as<GState>(), ctx->st);
  r = E.read_gs();
  if (r) return r;
  const GState& st = E.hgs[0];
  const double M = st.mass;
  if (M == 0) *q = 0;
  else if (mode_exact(g.mode))
    *q = (double)st.sin_i / M - u128_to_double(st.sa2_hi, st.sa2_lo) / (M * M);
  else
    *q = st.sin_d / M - st.sa2_d / (M * M);
  k_reactivate<<<1, 1, 0, ctx->st>>>(E.bgs.as<GState>());
  HIPC(ctx, hipStreamSynchronize(ctx->st));
  return MGL_OK;
}

mgl_status mgl_session_get_weights(mgl_session* s, double* a, double* k) {
  if (!s) return MGL_INVALID_ARG;
  mgl_ctx* ctx = s->ctx;
  HIPC(ctx, hipSetDevice(ctx->device));
  Engine& E = *s->E;
  const int64_t n = E.LG().n;
  const bool ex = mode_exact(E.LG().mode);
  std::vector<int64_t> bi(n);
  std::vector<double> bd(n);
  for (int which = 0; which < 2; ++which) {
    double* dst = which == 0 ? a : k;
    if (!dst) continue;
    const void* src = which == 0 ? E.ba.p : E.bk.p;
    if (ex) {
      HIPC(ctx, hipMemcpy(bi.data(), src, 8 * n, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < n; ++i) dst[i] = std::ldexp((double)bi[i], -E.scale);
    } else {
      HIPC(ctx, hipMemcpy(dst, src, 8 * n, hipMemcpyDeviceToHost));
    }
  }
  return MGL_OK;
}

mgl_status mgl_session_apply(mgl_session* s, const int32_t* c_new) {
  if (!s || !c_new) return MGL_INVALID_ARG;
  mgl_ctx* ctx = s->ctx;
  HIPC(ctx, hipSetDevice(ctx->device));
  Engine& E = *s->E;
  LevelGraph& g = E.LG();
  const int64_t n = g.n;
  for (int64_t v = 0; v < n; ++v)
    if (c_new[v] < 0 || c_new[v] >= n) FAIL(ctx, MGL_INVALID_ARG, "community id out of range");
  k_reactivate<<<1, 1, 0, ctx->st>>>(E.bgs.as<GState>());
  HIPC(ctx, hipMemcpyAsync(E.p, c_new, 4 * n, hipMemcpyHostToDevice, ctx->st));
  launch_update(g.mode, n, E.c, E.p, E.bk.p, E.ba.p, nullptr, E.bgs.as<GState>(), ctx->st);
  launch_copy_vid(E.p, E.c, n, ctx->st);
  launch_copy_vid(E.p, E.nx, n, ctx->st);
  HIPC(ctx, hipStreamSynchronize(ctx->st));
  HIPC(ctx, hipGetLastError());
  return MGL_OK;
}

mgl_status mgl_session_coarsen(mgl_session* s, int64_t* n2, int64_t* nnz2, int64_t* off2,
                               int32_t* col2, double* w2, int32_t* rank) {
  if (!s || !n2 || !nnz2) return MGL_INVALID_ARG;
  mgl_ctx* ctx = s->ctx;
  HIPC(ctx, hipSetDevice(ctx->device));
  Engine& E = *s->E;
  if (!s->coarsen_ready) {
    mgl_status r = E.alloc_coarsen();
    if (r) return r;
    s->coarsen_ready = true;
  }
  LevelGraph& g = E.LG();
  const int out_mode = coarse_mode(g.mode, E.rec32);
  DBuf doff, dcol, dw;
  HIPC(ctx, doff.ensure(8 * (g.n + 2)));
  HIPC(ctx, dcol.ensure(4 * std::max<int64_t>(g.nnz, 1)));
  HIPC(ctx, dw.ensure(8 * std::max<int64_t>(g.nnz, 1)));
  if (g.mode != MODE_UNIT) {
    HIPC(ctx, E.cv0.ensure(8 * std::max<int64_t>(g.nnz, 1)));
    HIPC(ctx, E.cv1.ensure(8 * std::max<int64_t>(g.nnz, 1)));
    E.CB.v0 = E.cv0.p;
    E.CB.v1 = E.cv1.p;
  }
  int32_t one = 1;
  HIPC(ctx, hipMemcpyAsync(E.bkeep.p, &one, 4, hipMemcpyHostToDevice, ctx->st));
  int64_t C = 0, E2 = 0;
  const int rc = coarsen_level(g.mode, out_mode, g.n, g.off.as<eid>(), g.col.as<vid>(), g.w.p,
                               E.c, nullptr, E.bkeep.as<int32_t>(), nullptr, g.nnz, E.CB,
                               doff.as<eid>(), dcol.as<vid>(), dw.p, &C, &E2, ctx->st);
  if (rc) FAIL(ctx, MGL_DEVICE_ERROR, "coarsening failed (%d)", rc);
  HIPC(ctx, hipStreamSynchronize(ctx->st));
  *n2 = C;
  *nnz2 = E2;
  if (off2) HIPC(ctx, hipMemcpy(off2, doff.p, 8 * (C + 1), hipMemcpyDeviceToHost));
  if (col2 && E2) HIPC(ctx, hipMemcpy(col2, dcol.p, 4 * E2, hipMemcpyDeviceToHost));
  if (w2 && E2) {
    if (out_mode == MODE_I32) {
      std::vector<int32_t> t(E2);
      HIPC(ctx, hipMemcpy(t.data(), dw.p, 4 * E2, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < E2; ++i) w2[i] = std::ldexp((double)t[i], -E.scale);
    } else if (out_mode == MODE_INT) {
      std::vector<int64_t> t(E2);
      HIPC(ctx, hipMemcpy(t.data(), dw.p, 8 * E2, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < E2; ++i) w2[i] = std::ldexp((double)t[i], -E.scale);
    } else if (out_mode == MODE_F64) {
      HIPC(ctx, hipMemcpy(w2, dw.p, 8 * E2, hipMemcpyDeviceToHost));
    } else {
      std::vector<float> t(E2);
      HIPC(ctx, hipMemcpy(t.data(), dw.p, 4 * E2, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < E2; ++i) w2[i] = t[i];
    }
  }
  if (rank && g.n) {
    std::vector<int64_t> rk(g.n + 1);
    std::vector<int32_t> fl(g.n);
    HIPC(ctx, hipMemcpy(rk.data(), E.CB.rank, 8 * (g.n + 1), hipMemcpyDeviceToHost));
    HIPC(ctx, hipMemcpy(fl.data(), E.CB.flag, 4 * g.n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < g.n; ++i) rank[i] = fl[i] ? (int32_t)rk[i] : -1;
  }
  return MGL_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Tuning file (text, written by mgl_autotune): lines starting with '#' and
// empty lines are skipped; "hub_bucket N", "hub_bucket_unclamped N",
// "hub_segmented 0|1", "hub_agg_threads 256|512|1024", "hub_count_threads
// 256|512|1024" and "hub_agg_per_cu 1..8"
// set the hub path; "table_check 0" accepts tables smaller than their bin's rows
// (a move that overflows one then fails with MGL_DEVICE_ERROR);
// "hub_stream 0" keeps the hub path on the main stream; "hub_fork_upper N"
// starts it once the bins of rows longer than N entries are launched (default
// 512, 0 = after every bin);
// "coarsen_group_entries N" bounds the kept entries a batch is coarsened in at
// once (graphs are grouped up to N, a larger graph alone); every other line
// is one bin, "upper shape s size", in increasing order of upper (-1:
// unbounded), whose shape, lane or thread count and table size must name a
// compiled kernel; a file without bin lines keeps the built-in layout.
// Anything else is an error.
static mgl_status load_tuning(mgl_ctx* ctx, const char* path, Tuning& tu) {
  FILE* f = fopen(path, "r");
  if (!f) FAIL(ctx, MGL_INVALID_ARG, "cannot open tuning file %s", path);
  BinLayout L;
  L.nb = 0;
  char line[256];
  const char* bad = nullptr;  // what is wrong, if anything
  int lineno = 0;
  while (!bad && fgets(line, sizeof(line), f)) {
    ++lineno;
    const char* p = line;
    while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
    if (*p == '#' || *p == '\n' || *p == 0) continue;
    long long hb;
    int hs;
    char extra;
    if (sscanf(p, "hub_bucket %lld %c", &hb, &extra) == 1) {
      if (hb < 1) bad = "hub_bucket";
      tu.hub_bucket = hb;
    } else if (sscanf(p, "hub_bucket_unclamped %lld %c", &hb, &extra) == 1) {
      if (hb < 1) bad = "hub_bucket_unclamped";
      tu.hub_bucket = hb;
      tu.hub_bucket_clamp = false;
    } else if (sscanf(p, "hub_segmented %d %c", &hs, &extra) == 1) {
      tu.hub_segmented = hs != 0;
    } else if (sscanf(p, "table_check %d %c", &hs, &extra) == 1) {
      tu.table_check = hs != 0;
    } else if (sscanf(p, "hub_stream %d %c", &hs, &extra) == 1) {
      tu.hub_stream = hs != 0;
    } else if (sscanf(p, "hub_fork_upper %lld %c", &hb, &extra) == 1) {
      if (hb < 0) bad = "hub_fork_upper";
      tu.hub_fork_upper = hb;
    } else if (sscanf(p, "coarsen_group_entries %lld %c", &hb, &extra) == 1) {
      if (hb < 1) bad = "coarsen_group_entries";
      tu.coarsen_group = hb;
    } else if (sscanf(p, "hub_agg_threads %d %c", &hs, &extra) == 1) {
      if (hs != 256 && hs != 512 && hs != 1024) bad = "hub_agg_threads";
      tu.hub_agg_threads = hs;
    } else if (sscanf(p, "hub_count_threads %d %c", &hs, &extra) == 1) {
      if (hs != 256 && hs != 512 && hs != 1024) bad = "hub_count_threads";
      tu.hub_count_threads = hs;
    } else if (sscanf(p, "hub_agg_per_cu %d %c", &hs, &extra) == 1) {
      if (hs < 1 || hs > 8) bad = "hub_agg_per_cu";
      tu.hub_agg_per_cu = hs;
    } else {
      long long up;
      int sh, sv, sz;
      if (sscanf(p, "%lld %d %d %d %c", &up, &sh, &sv, &sz, &extra) != 4) bad = "unknown line";
      else if (L.nb == MGL_MAX_BINS) bad = "more bins than MGL_MAX_BINS";
      else L.b[L.nb++] = {up < 0 ? INT64_MAX : (int64_t)up, sh, sv, sz};
    }
  }
  fclose(f);
  if (bad) FAIL(ctx, MGL_INVALID_ARG, "bad tuning file %s, line %d: %s", path, lineno, bad);
  if (L.nb == 0) return MGL_OK;  // no bin lines: the built-in layout
  if (L.nb < 2 || L.b[0].upper != 0 || L.b[0].shape != SHAPE_NONE)
    FAIL(ctx, MGL_INVALID_ARG, "bad tuning file %s: bin 0 must be degree 0", path);
  for (int i = 1; i < L.nb; ++i) {
    const BinSpec& b = L.b[i];
    if (b.upper <= L.b[i - 1].upper) FAIL(ctx, MGL_INVALID_ARG, "bad tuning file %s: bounds", path);
    if (b.shape == SHAPE_GROUP) {
      if (b.s != 4 && b.s != 8 && b.s != 16 && b.s != 32 && b.s != 64)
        FAIL(ctx, MGL_INVALID_ARG, "bad tuning file %s: bin %d lanes per vertex %d", path, i, b.s);
      // a table of 2^size slots must hold every distinct community of a row
      if (b.size < 1 || b.size > 15 || (tu.table_check && b.upper > ((int64_t)1 << b.size)))
        FAIL(ctx, MGL_INVALID_ARG, "bad tuning file %s: bin %d table too small", path, i);
    } else if (b.shape == SHAPE_BLOCK) {
      if (b.s != 256 && b.s != 512 && b.s != 1024)
        FAIL(ctx, MGL_INVALID_ARG, "bad tuning file %s: bin %d threads %d", path, i, b.s);
      if (b.size < 1 || (tu.table_check && b.upper > b.size))
        FAIL(ctx, MGL_INVALID_ARG, "bad tuning file %s: bin %d table too small", path, i);
    } else if (b.shape == SHAPE_LANE) {
      if (b.upper > 16) FAIL(ctx, MGL_INVALID_ARG, "bad tuning file %s: lane bin above 16", path);
    } else if (b.shape != SHAPE_GLOBAL) {
      FAIL(ctx, MGL_INVALID_ARG, "bad tuning file %s: bin %d shape", path, i);
    }
    if (b.shape == SHAPE_GROUP || b.shape == SHAPE_BLOCK || b.shape == SHAPE_LANE) {
      // every mode instantiates the same shapes
      char kn[128];
      if (b.upper == INT64_MAX || move_kernel_name(MODE_UNIT, 1, b, kn, (int)sizeof(kn)) != 0)
        FAIL(ctx, MGL_INVALID_ARG, "bad tuning file %s: bin %d has no compiled kernel", path, i);
    }
  }
  if (L.b[L.nb - 1].shape != SHAPE_GLOBAL)
    FAIL(ctx, MGL_INVALID_ARG, "bad tuning file %s: the last bin must be the hub path (shape %d)",
         path, (int)SHAPE_GLOBAL);
  tu.bins = L;
  tu.have_bins = true;
  return MGL_OK;
}

// ---------------------------------------------------------------------------
// Autotuner: times the level-0 move pass of g under candidate bin layouts
// (degree bounds, lanes per vertex, LDS table sizes, workgroup sizes) and
// writes the fastest as a tuning file.
static std::vector<BinLayout> tuning_candidates(const mgl_ctx* ctx) {
  std::vector<BinLayout> out;
  const BinLayout D = default_bins(ctx->lds_max, 4 + slot_bytes(MODE_UNIT));
  out.push_back(D);
  // variants of the built-in layout: the shapes are kept, the degree bounds
  // between them move (two entries per lane, earlier workgroups, earlier hubs)
  auto with = [&](std::initializer_list<BinSpec> head) {
    BinLayout L;
    L.nb = 0;
    for (const auto& b : head) L.b[L.nb++] = b;
    // append the built-in bins above the last listed bound
    const int64_t top = L.b[L.nb - 1].upper;
    for (int i = 0; i < D.nb; ++i)
      if (D.b[i].upper > top) L.b[L.nb++] = D.b[i];
    return L;
  };
  // the built-in layout without its half-LDS bin (the largest table from
  // degree 4097)
  {
    BinLayout L;
    L.nb = 0;
    for (int i = 0; i < D.nb; ++i)
      if (!(D.b[i].shape == SHAPE_BLOCK && D.b[i].upper == 8192)) L.b[L.nb++] = D.b[i];
    if (L.nb != D.nb) out.push_back(L);
  }
  // workgroups of 256 from degree 257
  out.push_back(with({{0, SHAPE_NONE, 0, 0}, {4, SHAPE_LANE, 1, 0}, {8, SHAPE_LANE, 1, 0},
                      {16, SHAPE_GROUP, 4, 5}, {32, SHAPE_GROUP, 8, 6},
                      {64, SHAPE_GROUP, 16, 7}, {128, SHAPE_GROUP, 32, 8},
                      {256, SHAPE_GROUP, 64, 9}, {1024, SHAPE_BLOCK, 256, 2048},
                      {4096, SHAPE_BLOCK, 1024, 8192}}));
  // one lane per vertex up to 16
  out.push_back(with({{0, SHAPE_NONE, 0, 0}, {4, SHAPE_LANE, 1, 0}, {8, SHAPE_LANE, 1, 0},
                      {16, SHAPE_LANE, 1, 0}}));
  // more lanes per vertex (one entry each)
  out.push_back(with({{0, SHAPE_NONE, 0, 0}, {4, SHAPE_LANE, 1, 0}, {8, SHAPE_LANE, 1, 0},
                      {16, SHAPE_GROUP, 16, 5}, {32, SHAPE_GROUP, 32, 6},
                      {64, SHAPE_GROUP, 64, 7}, {128, SHAPE_GROUP, 64, 8},
                      {256, SHAPE_GROUP, 64, 9}}));
  // fewer lanes per vertex (four entries each) further up, wavefront up to 512
  out.push_back(with({{0, SHAPE_NONE, 0, 0}, {4, SHAPE_LANE, 1, 0}, {8, SHAPE_LANE, 1, 0},
                      {16, SHAPE_GROUP, 4, 5}, {32, SHAPE_GROUP, 8, 6},
                      {64, SHAPE_GROUP, 16, 7}, {128, SHAPE_GROUP, 32, 8},
                      {256, SHAPE_GROUP, 64, 9}, {512, SHAPE_GROUP, 64, 10},
                      {2048, SHAPE_BLOCK, 512, 4096}}));
  // workgroups from degree 129
  out.push_back(with({{0, SHAPE_NONE, 0, 0}, {4, SHAPE_LANE, 1, 0}, {8, SHAPE_LANE, 1, 0},
                      {16, SHAPE_GROUP, 4, 5}, {32, SHAPE_GROUP, 8, 6},
                      {64, SHAPE_GROUP, 16, 7}, {128, SHAPE_GROUP, 32, 8},
                      {512, SHAPE_BLOCK, 256, 1024}, {4096, SHAPE_BLOCK, 1024, 8192}}));
  return out;
}

// Median move time (ms) of passes 2-4 from singletons under the context's
// current tuning.
static mgl_status tune_time(mgl_ctx* ctx, const mgl_csr* g, double& med) {
  mgl_opts o;
  mgl_opts_default(&o);
  mgl_session* s = nullptr;
  mgl_status st = mgl_session_create(ctx, g, &o, &s);
  if (st) return st;
  double q;
  int acc;
  std::vector<double> t;
  st = mgl_session_reset(s, &q);
  for (int i = 0; i < 4 && st == MGL_OK; ++i) {
    st = mgl_session_pass(s, &q, &acc);
    if (st == MGL_OK && i > 0) t.push_back(s->st.move_ms);
  }
  mgl_session_destroy(s);
  if (st) return st;
  std::sort(t.begin(), t.end());
  med = t.empty() ? 1e300 : t[t.size() / 2];
  return MGL_OK;
}

extern "C" mgl_status mgl_autotune(mgl_ctx* ctx, const mgl_csr* g, const char* path) {
  if (!ctx) return MGL_INVALID_ARG;
  if (!g || !path) FAIL(ctx, MGL_INVALID_ARG, "null argument");
  const Tuning saved = ctx->tune;
  double best_ms = 1e300;
  BinLayout best = default_bins(ctx->lds_max, 4 + slot_bytes(MODE_UNIT));
  for (const BinLayout& cand : tuning_candidates(ctx)) {
    ctx->tune.bins = cand;
    ctx->tune.have_bins = true;
    double med = 0;
    const mgl_status st = tune_time(ctx, g, med);
    if (st) { ctx->tune = saved; return st; }
    if (med < best_ms) { best_ms = med; best = cand; }
  }
  // then the hub path's aggregate shape on the best layout (threads per
  // workgroup, workgroups per CU, candidates per bucket); without hub rows
  // every candidate times the same and the first one stays
  struct HubShape { int threads, per_cu; int64_t bucket; };
  const HubShape hubs[] = {{512, 4, 3072}, {1024, 2, 4096}, {512, 4, 4096}, {512, 2, 4096},
                           {1024, 2, 3072}, {256, 4, 2048}};
  ctx->tune.bins = best;
  ctx->tune.have_bins = true;
  HubShape hbest = hubs[0];
  double hbest_ms = 1e300;
  for (const HubShape& h : hubs) {
    ctx->tune.hub_agg_threads = h.threads;
    ctx->tune.hub_agg_per_cu = h.per_cu;
    ctx->tune.hub_bucket = h.bucket;
    double med = 0;
    const mgl_status st = tune_time(ctx, g, med);
    if (st) { ctx->tune = saved; return st; }
    if (med < hbest_ms * 0.99) { hbest_ms = med; hbest = h; }
  }
  best_ms = std::min(best_ms, hbest_ms);
  ctx->tune = saved;
  FILE* f = fopen(path, "w");
  if (!f) FAIL(ctx, MGL_INVALID_ARG, "cannot write %s", path);
  fprintf(f,
          "# mgl bin layout: upper shape lanes_or_threads size (GROUP: log2 slots, BLOCK: "
          "slots) -- move median %.4f ms\n",
          best_ms);
  fprintf(f, "hub_agg_threads %d\nhub_agg_per_cu %d\nhub_bucket %lld\n", hbest.threads,
          hbest.per_cu, (long long)hbest.bucket);
  for (int i = 0; i < best.nb; ++i)
    fprintf(f, "%lld %d %d %d\n", best.b[i].upper == INT64_MAX ? -1LL : (long long)best.b[i].upper,
            best.b[i].shape, best.b[i].s, best.b[i].size);
  fclose(f);
  return MGL_OK;
}
