// CPU oracle for the Louvain pass of MG-alphaGCD (C++17 / OpenMP).
//
// TEST INFRASTRUCTURE ONLY. This file is the checker: tests/, the smoke check
// and the CPU-baseline leg of bench.py load it; the product library never
// links, loads or calls it.
//
// It restates, in plain sequential-semantics C++, the reference's
//   * move rule       src/louvain/louvain.cu:454-585 (tile), :633-805 (block),
//                     :855-1026 (grid); v1 filter src/louvain/louvain_gl.cu:504,578
//   * weight update   src/louvain/louvain.cu:1328-1336,1377-1381 (v1/v2, delta buffer)
//                     src/louvain/louvain_baseline.cu:1339-1361 (v0, direct atomics)
//   * modularity      src/louvain/louvain.cu:146-165 (sum a^2), :187-206 (sum_in), :234 (Q)
//   * vertex weights  src/louvain/louvain.cu:30-74 (row sums, self-loops included)
//   * driver          src/louvain/louvain.cu:1470-1693 (phase loop, inner loop,
//                     commit-or-reject, up/down flip)
//   * coarsening      src/coarsen_graph/coarsen_graph_mg.cu:211-226,267-286 (dense-rank
//                     renumbering), :816-917 (aggregation incl. self-loop), :2224-2228
//   * 2m              src/graph/host_graph.cpp:112-114
//
// Two arithmetic modes:
//   exact : weights are integers (unit graphs) or integer multiples of 2^-S
//           (dyadic weights, scaled by 2^S here). All sums are int64 / __int128,
//           hence order independent; FP64 appears only in the gain expression and
//           the final Q division, exactly as in the reference expression order.
//   fp64  : weights and sums in double, sequential order.
//
// Pinning (what fixes each function independently of this restatement) is
// listed in tests/test_oracle.py: hand-derived known answers, brute-force
// exact-rational dQ on tiny graphs, closed-form modularities, networkx.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <omp.h>

// ORC_PROGRESS=1 in the environment: one stderr line per inner iteration and
// per coarsening of the full driver (long runs show where they are).
static bool progress() {
  static const int on = [] {
    const char* e = std::getenv("ORC_PROGRESS");
    return (e && *e && *e != '0') ? 1 : 0;
  }();
  return on != 0;
}

typedef int64_t i64;
typedef int32_t i32;
typedef unsigned __int128 u128;

// u128 -> double, correctly rounded (round-to-nearest-even), by folding the
// bits below the top 64 into a sticky bit. The product's device code uses the
// same construction; both are checked against long-double/fractions in tests.
static double u128_to_double(u128 x) {
  uint64_t hi = (uint64_t)(x >> 64), lo = (uint64_t)x;
  if (hi == 0) return (double)lo;
  int s = __builtin_clzll(hi);
  uint64_t top = s ? ((hi << s) | (lo >> (64 - s))) : hi;
  uint64_t rest = lo << s;
  top |= (rest != 0);
  return std::ldexp((double)top, 64 - s);
}

extern "C" double orc_u128_to_double(uint64_t hi, uint64_t lo) {
  return u128_to_double(((u128)hi << 64) | lo);
}

// ---------------------------------------------------------------------------
// Graph view. w == nullptr means unit weights.
template <class W>
struct G {
  i64 n;
  const i64* off;
  const i32* col;
  const W* w;
  W wt(i64 e) const { return w ? w[e] : (W)1; }
};

// k[v] = sum of row weights, self-loops included (louvain.cu:30-74).
template <class W>
static void vertex_weights(const G<W>& g, W* k) {
#pragma omp parallel for schedule(dynamic, 4096)
  for (i64 v = 0; v < g.n; ++v) {
    W s = 0;
    for (i64 e = g.off[v]; e < g.off[v + 1]; ++e) s += g.wt(e);
    k[v] = s;
  }
}

// ---------------------------------------------------------------------------
// Move rule, one synchronous pass (louvain.cu:454-585).
//   version 0/2: candidates = every neighbour community != own; argmax; clamp.
//   version 1  : candidates filtered by direction before the argmax; no clamp.
//   tie 0      : equal gains go to the smaller id on up passes and to the larger
//                id on down passes (louvain.cu:544-554);
//   tie 1      : equal gains always go to the smaller id (the deterministic
//                smallest-id rule; no reference counterpart). The filter and
//                the clamp keep the pass direction in both rules.
template <class W>
static void move_pass(const G<W>& g, const W* k, const i32* c, const W* a, W mass,
                      int up, int version, int tie, i32* c_new) {
  const bool small = tie ? true : (up != 0);
  const double M = (double)mass / 2.0;  // louvain.cu:454 "mass = mass / 2.0"
#pragma omp parallel
  {
    std::vector<W> acc(g.n > 0 ? g.n : 1, (W)0);
    std::vector<char> seen(g.n > 0 ? g.n : 1, 0);
    std::vector<i32> touched;
#pragma omp for schedule(dynamic, 1024)
    for (i64 v = 0; v < g.n; ++v) {
      const i32 ci = c[v];
      const W ki = k[v];
      const W aci = a[ci] - ki;  // louvain.cu:483-494
      W eici = 0;
      touched.clear();
      for (i64 e = g.off[v]; e < g.off[v + 1]; ++e) {
        const i32 u = g.col[e];
        const i32 cj = c[u];
        const W wv = g.wt(e);
        bool cand;
        if (version == 1)
          cand = up ? (ci > cj) : (ci < cj);  // louvain_gl.cu:504
        else
          cand = (ci != cj);  // louvain.cu:506
        if (cand) {
          if (!seen[cj]) { seen[cj] = 1; touched.push_back(cj); }
          acc[cj] += wv;
        } else if (ci == cj && u != (i32)v) {
          eici += wv;  // self-loop excluded, louvain.cu:516-518
        }
      }
      double best = 0.0;
      i32 dst = ci;
      const double deici = (double)eici, dki = (double)ki, daci = (double)aci;
      for (i32 cj : touched) {
        // exact FP64 expression order of louvain.cu:540-543
        double wij = (double)acc[cj];
        const double acj = (double)a[cj];
        wij -= (deici - dki * (daci - acj) / (2. * M));
        wij /= M;
        bool better = small ? ((wij > best) || (wij == best && cj < dst))
                            : ((wij > best) || (wij == best && cj > dst));  // :544-554
        if (better) { dst = cj; best = wij; }
        acc[cj] = 0;
        seen[cj] = 0;
      }
      i32 out;
      if (version == 1) out = dst;  // louvain_gl.cu:578
      else out = up ? std::min(dst, ci) : std::max(dst, ci);  // louvain.cu:579-585
      c_new[v] = out;
    }
  }
}

// a[c_old] -= k, a[c_new] += k for every moved vertex (louvain.cu:1328-1336 +
// 1377-1381 through a zeroed delta buffer; louvain_baseline.cu:1339-1361
// directly). In exact mode both orders give the same integers.
template <class W>
static void update_weights(i64 n, const W* k, const i32* c, const i32* c_new, W* a) {
  for (i64 v = 0; v < n; ++v) {
    if (c[v] != c_new[v]) {
      a[c[v]] -= k[v];
      a[c_new[v]] += k[v];
    }
  }
}

// Q = sum_in / mass - sum_c a_c^2 / mass^2 (louvain.cu:146-165, 187-206, 234).
static double modularity_exact(const G<i64>& g, const i32* c, const i64* a, i64 mass) {
  i64 sin = 0;
#pragma omp parallel for reduction(+ : sin) schedule(dynamic, 4096)
  for (i64 v = 0; v < g.n; ++v) {
    const i32 cv = c[v];
    for (i64 e = g.off[v]; e < g.off[v + 1]; ++e)
      if (c[g.col[e]] == cv) sin += g.wt(e);
  }
  // sum a^2 in 128-bit: per-thread partials then a fixed-order combine
  int nt = omp_get_max_threads();
  std::vector<u128> part(nt, 0);
#pragma omp parallel
  {
    u128 s = 0;
#pragma omp for schedule(static)
    for (i64 i = 0; i < g.n; ++i) {
      const i64 x = a[i];
      const u128 ax = (u128)(x < 0 ? -x : x);
      s += ax * ax;
    }
    part[omp_get_thread_num()] = s;
  }
  u128 sa2 = 0;
  for (auto p : part) sa2 += p;
  const double dm = (double)mass;
  return (double)sin / dm - u128_to_double(sa2) / (dm * dm);
}

static double modularity_fp64(const G<double>& g, const i32* c, const double* a, double mass) {
  double sa2 = 0;
  for (i64 i = 0; i < g.n; ++i) sa2 += a[i] * a[i];
  double sin = 0;
  for (i64 v = 0; v < g.n; ++v) {
    const i32 cv = c[v];
    for (i64 e = g.off[v]; e < g.off[v + 1]; ++e)
      if (c[g.col[e]] == cv) sin += g.wt(e);
  }
  return sin / mass - sa2 / (mass * mass);
}

// ---------------------------------------------------------------------------
// Coarsening: dense rank of the used community ids in ascending order
// (coarsen_graph_mg.cu:211-226, 267-286), then every fine entry (u,v,w) adds w
// to coarse entry (rank c[u], rank c[v]); intra-community entries form the
// self-loop (:816-917). Rows are emitted sorted by column (canonical form; the
// reference emits hash-slot order, which no later step depends on in exact mode).
template <class W>
struct CSR {
  i64 n = 0;
  std::vector<i64> off;
  std::vector<i32> col;
  std::vector<W> w;
};

template <class W>
static void coarsen(const G<W>& g, const i32* c, CSR<W>& out, std::vector<i32>& rank) {
  std::vector<i32> flag(g.n, 0);
  for (i64 v = 0; v < g.n; ++v) flag[c[v]] = 1;
  rank.assign(g.n, -1);
  i32 C = 0;
  for (i64 v = 0; v < g.n; ++v)
    if (flag[v]) rank[v] = C++;
  // bucket fine rows by coarse row
  std::vector<i64> cnt(C + 1, 0);
  for (i64 v = 0; v < g.n; ++v) cnt[rank[c[v]] + 1] += g.off[v + 1] - g.off[v];
  for (i32 r = 0; r < C; ++r) cnt[r + 1] += cnt[r];
  const i64 E = g.off[g.n];
  std::vector<std::pair<i32, W>> ent(E);
  std::vector<i64> pos(cnt.begin(), cnt.end() - 1);
  for (i64 v = 0; v < g.n; ++v) {
    const i32 r = rank[c[v]];
    for (i64 e = g.off[v]; e < g.off[v + 1]; ++e)
      ent[pos[r]++] = {rank[c[g.col[e]]], g.wt(e)};
  }
  std::vector<i64> rowlen(C, 0);
  // sort each coarse row by column; merge equal columns in sorted (stable) order
#pragma omp parallel for schedule(dynamic, 64)
  for (i32 r = 0; r < C; ++r) {
    auto b = ent.begin() + cnt[r], e = ent.begin() + cnt[r + 1];
    std::stable_sort(b, e, [](const std::pair<i32, W>& x, const std::pair<i32, W>& y) {
      return x.first < y.first;
    });
    i64 o = cnt[r];
    for (i64 i = cnt[r]; i < cnt[r + 1]; ++i) {
      if (i > cnt[r] && ent[i].first == ent[o - 1].first) ent[o - 1].second += ent[i].second;
      else ent[o++] = ent[i];
    }
    rowlen[r] = o - cnt[r];
  }
  out.n = C;
  out.off.assign(C + 1, 0);
  for (i32 r = 0; r < C; ++r) out.off[r + 1] = out.off[r] + rowlen[r];
  out.col.resize(out.off[C]);
  out.w.resize(out.off[C]);
  for (i32 r = 0; r < C; ++r)
    for (i64 i = 0; i < rowlen[r]; ++i) {
      out.col[out.off[r] + i] = ent[cnt[r] + i].first;
      out.w[out.off[r] + i] = ent[cnt[r] + i].second;
    }
}

// ---------------------------------------------------------------------------
// Full driver (louvain.cu:1470-1693).
struct Level {
  i64 n = 0;
  double q = 0;                 // Q_host after the phase
  std::vector<double> qtraj;    // new_Q after every inner iteration (after reject: cur_Q)
  std::vector<i32> assign;      // committed community ids (level vertex ids)
  int iters = 0;
};

template <class W>
struct Result {
  std::vector<Level> levels;
  std::vector<CSR<W>> graphs;   // graphs[l] = input graph of level l (graphs[0] = copy of input)
  std::vector<i32> final_assign; // dense labels for level-0 vertices
  double final_q = 0;
};

struct ResultAny {
  int exact = 1;
  Result<i64> ri;
  Result<double> rd;
};

template <class W>
static double modq(const G<W>& g, const i32* c, const W* a, W mass);
// A graph without edges has mass 0; its Q is defined as 0 here (the reference
// would divide 0 by 0).
template <>
double modq<i64>(const G<i64>& g, const i32* c, const i64* a, i64 mass) {
  return mass == 0 ? 0.0 : modularity_exact(g, c, a, mass);
}
template <>
double modq<double>(const G<double>& g, const i32* c, const double* a, double mass) {
  return mass == 0 ? 0.0 : modularity_fp64(g, c, a, mass);
}

template <class W>
static void louvain(CSR<W> in, double thr, int max_phases, int max_iter, int version, int tie,
                    Result<W>& R) {
  W mass = 0;
  for (i64 e = 0; e < (i64)in.col.size(); ++e) mass += in.w.empty() ? (W)1 : in.w[e];
  if (in.w.empty()) in.w.assign(in.col.size(), (W)1);
  R.graphs.push_back(in);
  double Q_host = 0, Q_old = -1;
  int phase = 0;
  std::vector<std::vector<i32>> ranks;
  while (phase < max_phases && (Q_host - Q_old) > thr) {
    Q_old = Q_host;
    CSR<W>& cur = R.graphs.back();
    G<W> g{cur.n, cur.off.data(), cur.col.data(), cur.w.data()};
    Level L;
    L.n = cur.n;
    std::vector<W> k(cur.n), a(cur.n);
    std::vector<i32> c(cur.n), cn(cur.n);
    for (i64 v = 0; v < cur.n; ++v) c[v] = (i32)v;
    vertex_weights(g, k.data());
    a = k;
    double new_Q = modq(g, c.data(), a.data(), mass);
    double cur_Q = new_Q - 1;
    bool up = true;
    int it = 0;
    while ((new_Q - cur_Q) > thr && (max_iter <= 0 || it < max_iter)) {
      cur_Q = new_Q;
      move_pass(g, k.data(), c.data(), a.data(), mass, up ? 1 : 0, version, tie, cn.data());
      up = !up;
      update_weights(cur.n, k.data(), c.data(), cn.data(), a.data());
      new_Q = modq(g, cn.data(), a.data(), mass);
      if ((new_Q - cur_Q) > thr) c = cn;
      else new_Q = cur_Q;
      L.qtraj.push_back(new_Q);
      ++it;
      if (progress())
        fprintf(stderr, "oracle: level %d (%lld vertices) iteration %d Q=%.12f\n", phase,
                (long long)cur.n, it, new_Q);
    }
    L.iters = it;
    L.q = new_Q;
    L.assign = c;
    Q_host = new_Q;
    R.levels.push_back(L);
    if ((Q_host - Q_old) <= thr) break;
    CSR<W> nxt;
    std::vector<i32> rank;
    coarsen(g, c.data(), nxt, rank);
    if (progress())
      fprintf(stderr, "oracle: level %d coarsened to %lld vertices, %lld entries\n", phase,
              (long long)nxt.n, (long long)nxt.col.size());
    R.graphs.push_back(std::move(nxt));
    ++phase;
  }
  // compose: final label of level-0 vertex = dense rank of its last-level community
  const i64 n0 = R.graphs[0].n;
  std::vector<i32> lab(n0);
  for (i64 v = 0; v < n0; ++v) lab[v] = (i32)v;
  for (size_t l = 0; l < R.levels.size(); ++l) {
    const auto& as = R.levels[l].assign;
    // dense rank of used ids at this level (the coarse id)
    std::vector<i32> flag(R.levels[l].n, 0), rk(R.levels[l].n, -1);
    for (auto x : as) flag[x] = 1;
    i32 C = 0;
    for (i64 i = 0; i < R.levels[l].n; ++i)
      if (flag[i]) rk[i] = C++;
    for (i64 v = 0; v < n0; ++v) lab[v] = rk[as[lab[v]]];
  }
  R.final_assign = lab;
  R.final_q = R.levels.empty() ? 0.0 : R.levels.back().q;
}

// ---------------------------------------------------------------------------
// C interface (ctypes).

// Choose S so that every weight * 2^S is an integer; returns -1 if a weight is
// not finite/positive or if the scaled total would not fit comfortably in int64.
extern "C" int orc_dyadic_scale(i64 nnz, const double* w) {
  int S = 0;
  long double tot = 0;
  for (i64 e = 0; e < nnz; ++e) {
    double x = w[e];
    if (!(x > 0) || !std::isfinite(x)) return -1;
    int ex;
    double m = std::frexp(x, &ex);  // x = m * 2^ex, m in [0.5,1)
    // m * 2^53 is an integer; strip trailing zero bits
    uint64_t mi = (uint64_t)std::ldexp(m, 53);
    int tz = __builtin_ctzll(mi);
    int need = 53 - tz - ex;  // x = (mi>>tz) * 2^(ex-53+tz)
    if (need > S) S = need;
    tot += x;
  }
  if (tot * std::ldexp(1.0L, S) > std::ldexp(1.0L, 62)) return -1;
  return S;
}

template <class W>
static void copy_in(CSR<W>& g, i64 n, const i64* off, const i32* col) {
  g.n = n;
  g.off.assign(off, off + n + 1);
  g.col.assign(col, col + off[n]);
}

// mode 0: exact (w may be NULL = unit, or dyadic doubles); mode 1: fp64.
// tie: 0 = reference rule (alternating with the pass direction), 1 = smallest id.
extern "C" void* orc_louvain(i64 n, const i64* off, const i32* col, const double* w, int mode,
                             double thr, int max_phases, int max_iter, int version, int tie) {
  ResultAny* R = new ResultAny();
  if (mode == 0) {
    R->exact = 1;
    CSR<i64> g;
    copy_in(g, n, off, col);
    if (w) {
      int S = orc_dyadic_scale(off[n], w);
      if (S < 0) { delete R; return nullptr; }
      g.w.resize(off[n]);
      for (i64 e = 0; e < off[n]; ++e) g.w[e] = (i64)std::ldexp(w[e], S);
    }
    louvain(g, thr, max_phases, max_iter, version, tie, R->ri);
  } else {
    R->exact = 0;
    CSR<double> g;
    copy_in(g, n, off, col);
    if (w) g.w.assign(w, w + off[n]);
    louvain(g, thr, max_phases, max_iter, version, tie, R->rd);
  }
  return R;
}

#define RSEL(R, expr) ((R)->exact ? (R)->ri.expr : (R)->rd.expr)
extern "C" int orc_res_levels(void* p) { auto R = (ResultAny*)p; return (int)RSEL(R, levels.size()); }
extern "C" i64 orc_res_n(void* p, int l) { auto R = (ResultAny*)p; return RSEL(R, levels[l].n); }
extern "C" double orc_res_q(void* p, int l) { auto R = (ResultAny*)p; return RSEL(R, levels[l].q); }
extern "C" int orc_res_iters(void* p, int l) { auto R = (ResultAny*)p; return RSEL(R, levels[l].iters); }
extern "C" const double* orc_res_qtraj(void* p, int l) { auto R = (ResultAny*)p; return RSEL(R, levels[l].qtraj.data()); }
extern "C" const i32* orc_res_assign(void* p, int l) { auto R = (ResultAny*)p; return RSEL(R, levels[l].assign.data()); }
extern "C" const i32* orc_res_final(void* p) { auto R = (ResultAny*)p; return RSEL(R, final_assign.data()); }
extern "C" int orc_res_ngraphs(void* p) { auto R = (ResultAny*)p; return (int)RSEL(R, graphs.size()); }
extern "C" i64 orc_res_graph_n(void* p, int l) { auto R = (ResultAny*)p; return RSEL(R, graphs[l].n); }
extern "C" const i64* orc_res_graph_off(void* p, int l) { auto R = (ResultAny*)p; return RSEL(R, graphs[l].off.data()); }
extern "C" const i32* orc_res_graph_col(void* p, int l) { auto R = (ResultAny*)p; return RSEL(R, graphs[l].col.data()); }
// weights of graph l as doubles (exact mode: scaled integers converted; caller divides by 2^S)
extern "C" void orc_res_graph_w(void* p, int l, double* out) {
  auto R = (ResultAny*)p;
  if (R->exact) {
    auto& g = R->ri.graphs[l];
    for (size_t i = 0; i < g.w.size(); ++i) out[i] = (double)g.w[i];
  } else {
    auto& g = R->rd.graphs[l];
    std::copy(g.w.begin(), g.w.end(), out);
  }
}
extern "C" void orc_res_free(void* p) { delete (ResultAny*)p; }

// --- single-step entry points (exact mode; w NULL = unit, else scaled int64) ---
extern "C" void orc_vertex_weights_i64(i64 n, const i64* off, const i64* w, i64* k) {
  G<i64> g{n, off, nullptr, w};
  vertex_weights(g, k);
}
extern "C" void orc_move_i64(i64 n, const i64* off, const i32* col, const i64* w, const i64* k,
                             const i32* c, const i64* a, i64 mass, int up, int version,
                             int tie, i32* c_new) {
  G<i64> g{n, off, col, w};
  move_pass(g, k, c, a, mass, up, version, tie, c_new);
}
extern "C" void orc_update_weights_i64(i64 n, const i64* k, const i32* c, const i32* c_new,
                                       i64* a) {
  update_weights(n, k, c, c_new, a);
}
extern "C" double orc_modularity_i64(i64 n, const i64* off, const i32* col, const i64* w,
                                     const i32* c, const i64* a, i64 mass) {
  G<i64> g{n, off, col, w};
  return modularity_exact(g, c, a, mass);
}
// fp64 step versions
extern "C" void orc_move_f64(i64 n, const i64* off, const i32* col, const double* w,
                             const double* k, const i32* c, const double* a, double mass, int up,
                             int version, int tie, i32* c_new) {
  G<double> g{n, off, col, w};
  move_pass(g, k, c, a, mass, up, version, tie, c_new);
}
extern "C" double orc_modularity_f64(i64 n, const i64* off, const i32* col, const double* w,
                                     const i32* c, const double* a, double mass) {
  G<double> g{n, off, col, w};
  return modularity_fp64(g, c, a, mass);
}

// Coarsen (exact mode). Returns coarse size; fills caller buffers sized for the
// worst case (n+1 offsets, nnz entries) and rank[n].
extern "C" i64 orc_coarsen_i64(i64 n, const i64* off, const i32* col, const i64* w,
                               const i32* c, i64* off2, i32* col2, i64* w2, i32* rank) {
  G<i64> g{n, off, col, w};
  CSR<i64> out;
  std::vector<i32> rk;
  coarsen(g, c, out, rk);
  std::copy(out.off.begin(), out.off.end(), off2);
  std::copy(out.col.begin(), out.col.end(), col2);
  std::copy(out.w.begin(), out.w.end(), w2);
  std::copy(rk.begin(), rk.end(), rank);
  return out.n;
}

// One timed level-0 pass for the CPU baseline: move + weight update +
// modularity + commit, exactly as the driver's inner-loop body. Returns Q.
extern "C" double orc_pass_i64(i64 n, const i64* off, const i32* col, const i64* w, const i64* k,
                               i32* c, i64* a, i64 mass, int up, int version, i32* c_new,
                               double cur_q, double thr) {
  G<i64> g{n, off, col, w};
  move_pass(g, k, c, a, mass, up, version, 0, c_new);
  update_weights(n, k, c, c_new, a);
  double q = modularity_exact(g, c_new, a, mass);
  if (q - cur_q > thr) std::memcpy(c, c_new, sizeof(i32) * n);
  return q;
}

// The same pass in FP64 (weights, k, a as doubles; the reference's weight_t),
// the CPU baseline of weighted graphs whose weights are not dyadic.
extern "C" void orc_vertex_weights_f64(i64 n, const i64* off, const double* w, double* k) {
  G<double> g{n, off, nullptr, w};
  vertex_weights(g, k);
}
extern "C" double orc_pass_f64(i64 n, const i64* off, const i32* col, const double* w,
                               const double* k, i32* c, double* a, double mass, int up,
                               int version, i32* c_new, double cur_q, double thr) {
  G<double> g{n, off, col, w};
  move_pass(g, k, c, a, mass, up, version, 0, c_new);
  update_weights(n, k, c, c_new, a);
  double q = modularity_fp64(g, c_new, a, mass);
  if (q - cur_q > thr) std::memcpy(c, c_new, sizeof(i32) * n);
  return q;
}

extern "C" int orc_num_threads() { return omp_get_max_threads(); }
