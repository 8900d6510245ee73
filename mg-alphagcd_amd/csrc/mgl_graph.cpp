// Host graph tools of the MI355X Louvain engine: seeded R-MAT and planted
// partition generators, canonicalisation (symmetrise, drop self-loops, dedup,
// sort rows), the Matrix Market reader and result writers.
//
// Random numbers are counter based (SplitMix64 of seed and sample index), so a
// graph is a pure function of its parameters and does not depend on the
// number of threads.
#include "../../include/mgl.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t rng(uint64_t seed, uint64_t stream, uint64_t i) {
  return mix64(mix64(seed * 0xD1B54A32D192ED03ull + stream) ^ (i * 0x8CB92BA72F3D8DD7ull));
}
inline double u01(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

int nthreads() {
  int nt = (int)std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;
  if (nt > 16) nt = 16;
  return nt;
}

template <class F>
void pfor(int64_t n, F f) {
  const int nt = (n < (1 << 15)) ? 1 : nthreads();
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([=]() { f(n * t / nt, n * (t + 1) / nt, t); });
  for (auto& x : th) x.join();
}

// Parallel stable LSD radix sort of 64-bit keys on their low `bits` bits.
void radix_sort_u64(std::vector<uint64_t>& a, int bits) {
  const int64_t n = (int64_t)a.size();
  if (n < 2 || bits <= 0) return;
  std::vector<uint64_t> b(n);
  const int D = 11, R = 1 << D;
  const int nt = (n < (1 << 16)) ? 1 : nthreads();
  std::vector<int64_t> hist((size_t)nt * R);
  for (int sh = 0; sh < bits; sh += D) {
    std::fill(hist.begin(), hist.end(), 0);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
      th.emplace_back([&, t]() {
        int64_t* h = &hist[(size_t)t * R];
        const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        for (int64_t i = lo; i < hi; ++i) h[(a[i] >> sh) & (R - 1)]++;
      });
    for (auto& x : th) x.join();
    th.clear();
    int64_t run = 0;
    for (int d = 0; d < R; ++d)
      for (int t = 0; t < nt; ++t) {
        int64_t c = hist[(size_t)t * R + d];
        hist[(size_t)t * R + d] = run;
        run += c;
      }
    for (int t = 0; t < nt; ++t)
      th.emplace_back([&, t]() {
        int64_t* h = &hist[(size_t)t * R];
        const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        for (int64_t i = lo; i < hi; ++i) b[h[(a[i] >> sh) & (R - 1)]++] = a[i];
      });
    for (auto& x : th) x.join();
    a.swap(b);
  }
}

int bits_for(int64_t n) {
  int b = 1;
  while ((1ll << b) < n) ++b;
  return b;
}

// Symmetric weight of the undirected pair {u, v}.
float pair_weight(int kind, uint64_t seed, uint32_t u, uint32_t v) {
  const uint32_t lo = u < v ? u : v, hi = u < v ? v : u;
  const uint64_t h = rng(seed, 0x5745494748ull, ((uint64_t)hi << 32) | lo);
  if (kind == 1) {
    // 1.5 - U, U uniform on [0,1) with 2^-24 resolution: a float in (0.5, 1.5]
    const double U = (double)(h >> 40) * (1.0 / 16777216.0);
    return (float)(1.5 - U);
  }
  // lognormal(0, 1) by Box-Muller
  const uint64_t h2 = mix64(h ^ 0xA5A5A5A5ull);
  double u1 = u01(h);
  if (u1 < 1e-300) u1 = 1e-300;
  const double u2 = u01(h2);
  const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  return (float)std::exp(z);
}

// Build a canonical CSR from directed keys (u<<32|v): sort, dedup, offsets.
mgl_status build_csr(std::vector<uint64_t>& keys, int64_t n, int weights, uint64_t seed,
                     mgl_host_csr* out) {
  // repack as (u << b | v) so the sort covers 2*b bits only
  const int b = bits_for(n > 1 ? n : 2);
  const uint64_t vmask = (1ull << b) - 1;
  pfor((int64_t)keys.size(), [&](int64_t lo, int64_t hi, int) {
    for (int64_t i = lo; i < hi; ++i) keys[i] = ((keys[i] >> 32) << b) | (keys[i] & 0xFFFFFFFFull);
  });
  radix_sort_u64(keys, 2 * b);
  // dedup (keys are sorted): per-thread compaction, then concatenation
  const int64_t K = (int64_t)keys.size();
  const int nt = (K < (1 << 16)) ? 1 : nthreads();
  std::vector<int64_t> cnt(nt + 1, 0);
  {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
      th.emplace_back([&, t]() {
        const int64_t lo = K * t / nt, hi = K * (t + 1) / nt;
        int64_t o = lo;
        for (int64_t i = lo; i < hi; ++i)
          if (i == 0 || keys[i] != keys[i - 1]) keys[o++] = keys[i];
        cnt[t + 1] = o - lo;
      });
    for (auto& x : th) x.join();
  }
  for (int t = 0; t < nt; ++t) {
    const int64_t lo = K * t / nt;
    if (cnt[t] != lo || t == 0) {
      if (t > 0) std::memmove(&keys[cnt[t]], &keys[lo], sizeof(uint64_t) * cnt[t + 1]);
    }
    cnt[t + 1] += cnt[t];
  }
  const int64_t m = cnt[nt];
  keys.resize(m);
  out->n = n;
  out->nnz = m;
  out->offsets = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
  out->cols = (int32_t*)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
  out->w = weights ? (float*)malloc(sizeof(float) * (m > 0 ? m : 1)) : nullptr;
  out->w64 = nullptr;
  if (!out->offsets || !out->cols || (weights && !out->w)) {
    mgl_csr_free(out);
    return MGL_OOM;
  }
  // offsets: the first entry of each row, found by binary search per row
  pfor(n + 1, [&](int64_t lo, int64_t hi, int) {
    for (int64_t v = lo; v < hi; ++v)
      out->offsets[v] = std::lower_bound(keys.begin(), keys.end(), (uint64_t)v << b) - keys.begin();
  });
  pfor(m, [&](int64_t lo, int64_t hi, int) {
    for (int64_t i = lo; i < hi; ++i) {
      const uint32_t u = (uint32_t)(keys[i] >> b), v = (uint32_t)(keys[i] & vmask);
      out->cols[i] = (int32_t)v;
      if (weights) out->w[i] = pair_weight(weights, seed, u, v);
    }
  });
  return MGL_OK;
}

}  // namespace

extern "C" {

void mgl_csr_free(mgl_host_csr* g) {
  if (!g) return;
  free(g->offsets);
  free(g->cols);
  free(g->w);
  free(g->w64);
  g->offsets = nullptr;
  g->cols = nullptr;
  g->w = nullptr;
  g->w64 = nullptr;
  g->n = g->nnz = 0;
}

// R-MAT (Chakrabarti, Zhan, Faloutsos 2004): each of ef * 2^scale samples
// descends `scale` levels choosing quadrant a, b, c or d. Both directions of
// every non-loop sample are emitted; duplicates and self-loops are dropped.
mgl_status mgl_gen_rmat(int scale, double ef, double a, double b, double c, uint64_t seed,
                        int weights, int64_t n_limit, mgl_host_csr* out) {
  if (!out || scale < 1 || scale > 30 || ef <= 0 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
    return MGL_INVALID_ARG;
  if (weights < 0 || weights > 2) return MGL_INVALID_ARG;
  memset(out, 0, sizeof(*out));
  const int64_t N = (int64_t)1 << scale;
  const int64_t n = (n_limit > 0 && n_limit < N) ? n_limit : N;
  const int64_t m = (int64_t)std::llround(ef * (double)N);
  const double ab = a + b, abc = a + b + c;
  const int nt = (m < (1 << 15)) ? 1 : nthreads();
  std::vector<std::vector<uint64_t>> part(nt);
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t]() {
      const int64_t lo = m * t / nt, hi = m * (t + 1) / nt;
      auto& P = part[t];
      P.reserve((size_t)(hi - lo) * 2);
      const uint64_t sk = mix64(seed * 0xD1B54A32D192ED03ull + 0x524D4154ull);
      for (int64_t i = lo; i < hi; ++i) {
        uint64_t u = 0, v = 0;
        // one 64-bit draw serves two levels (32-bit uniforms)
        const uint64_t base = mix64(sk ^ ((uint64_t)i * 0x8CB92BA72F3D8DD7ull));
        uint64_t x = 0;
        for (int l = 0; l < scale; ++l) {
          if ((l & 1) == 0) x = mix64(base + (uint64_t)(l >> 1) * 0x9E3779B97F4A7C15ull);
          const double r = (double)((l & 1) ? (x & 0xFFFFFFFFull) : (x >> 32)) * (1.0 / 4294967296.0);
          u <<= 1;
          v <<= 1;
          if (r < a) {
          } else if (r < ab) {
            v |= 1;
          } else if (r < abc) {
            u |= 1;
          } else {
            u |= 1;
            v |= 1;
          }
        }
        if (u == v || (int64_t)u >= n || (int64_t)v >= n) continue;
        P.push_back((u << 32) | v);
        P.push_back((v << 32) | u);
      }
    });
  for (auto& x : th) x.join();
  size_t tot = 0;
  for (auto& P : part) tot += P.size();
  std::vector<uint64_t> keys;
  keys.reserve(tot);
  for (auto& P : part) {
    keys.insert(keys.end(), P.begin(), P.end());
    std::vector<uint64_t>().swap(P);
  }
  return build_csr(keys, n, weights, seed, out);
}

// Planted partition: n_comm communities with power-law sizes (P(s) ~ s^-exponent,
// s >= min_size), rescaled to sum n; s*deg_in/2 uniform intra-community samples
// per community and n*deg_out/2 uniform inter-community samples; vertex ids
// randomly permuted.
mgl_status mgl_gen_sbm(int64_t n, int64_t n_comm, double exponent, int64_t min_size,
                       double deg_in, double deg_out, uint64_t seed, int weights,
                       mgl_host_csr* out) {
  if (!out || n < 1 || n_comm < 1 || n_comm > n || exponent <= 1.0 || min_size < 1 ||
      deg_in < 0 || deg_out < 0 || weights < 0 || weights > 2 || n > 0x7fffffffLL)
    return MGL_INVALID_ARG;
  memset(out, 0, sizeof(*out));
  // community sizes
  std::vector<double> raw(n_comm);
  double tot = 0;
  for (int64_t i = 0; i < n_comm; ++i) {
    const double u = u01(rng(seed, 1, (uint64_t)i));
    double s = (double)min_size * std::pow(1.0 - u, -1.0 / (exponent - 1.0));
    s = std::min(s, (double)n / 4.0);
    raw[i] = s;
    tot += s;
  }
  std::vector<int64_t> size(n_comm);
  int64_t sum = 0;
  const int64_t floor_sz = std::min<int64_t>(min_size, n / n_comm);
  for (int64_t i = 0; i < n_comm; ++i) {
    size[i] = std::max<int64_t>(floor_sz, (int64_t)std::floor(raw[i] * (double)n / tot));
    sum += size[i];
  }
  // fix the total on the largest communities
  std::vector<int64_t> order(n_comm);
  for (int64_t i = 0; i < n_comm; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
    return size[x] != size[y] ? size[x] > size[y] : x < y;
  });
  int64_t k = 0;
  while (sum != n) {
    const int64_t j = order[k % n_comm];
    if (sum < n) { size[j]++; sum++; }
    else if (size[j] > 1) { size[j]--; sum--; }
    ++k;
  }
  std::vector<int64_t> cstart(n_comm + 1, 0);
  for (int64_t i = 0; i < n_comm; ++i) cstart[i + 1] = cstart[i] + size[i];
  std::vector<int32_t> comm_of(n);
  for (int64_t i = 0; i < n_comm; ++i)
    for (int64_t v = cstart[i]; v < cstart[i + 1]; ++v) comm_of[v] = (int32_t)i;
  // random relabelling: sort vertices by a random key
  std::vector<uint64_t> pk(n);
  for (int64_t v = 0; v < n; ++v) pk[v] = (rng(seed, 2, (uint64_t)v) & ~0xFFFFFFFFull) | (uint64_t)v;
  radix_sort_u64(pk, 64);
  std::vector<int32_t> label(n);
  for (int64_t i = 0; i < n; ++i) label[(uint32_t)pk[i]] = (int32_t)i;
  std::vector<uint64_t>().swap(pk);
  // samples
  std::vector<int64_t> intra_start(n_comm + 1, 0);
  for (int64_t i = 0; i < n_comm; ++i)
    intra_start[i + 1] = intra_start[i] + (int64_t)std::llround((double)size[i] * deg_in / 2.0);
  const int64_t m_in = intra_start[n_comm];
  const int64_t m_out = (int64_t)std::llround((double)n * deg_out / 2.0);
  const int nt = ((m_in + m_out) < (1 << 15)) ? 1 : nthreads();
  std::vector<std::vector<uint64_t>> part(nt);
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t]() {
      auto& P = part[t];
      // intra: community-major sample index
      const int64_t lo = m_in * t / nt, hi = m_in * (t + 1) / nt;
      int64_t ci = std::upper_bound(intra_start.begin(), intra_start.end(), lo) - intra_start.begin() - 1;
      for (int64_t i = lo; i < hi; ++i) {
        while (i >= intra_start[ci + 1]) ++ci;
        const int64_t s = size[ci];
        if (s < 2) continue;
        const uint64_t r = rng(seed, 3, (uint64_t)i);
        const int64_t x = cstart[ci] + (int64_t)((r >> 32) % (uint64_t)s);
        const int64_t y = cstart[ci] + (int64_t)((r & 0xFFFFFFFFull) % (uint64_t)s);
        if (x == y) continue;
        const uint64_t u = (uint64_t)label[x], v = (uint64_t)label[y];
        P.push_back((u << 32) | v);
        P.push_back((v << 32) | u);
      }
      const int64_t lo2 = m_out * t / nt, hi2 = m_out * (t + 1) / nt;
      for (int64_t i = lo2; i < hi2; ++i) {
        for (int tries = 0; tries < 64; ++tries) {
          const uint64_t r = rng(seed, 4 + (uint64_t)tries, (uint64_t)i);
          const int64_t x = (int64_t)((r >> 32) % (uint64_t)n);
          const int64_t y = (int64_t)((r & 0xFFFFFFFFull) % (uint64_t)n);
          if (comm_of[x] == comm_of[y]) continue;
          const uint64_t u = (uint64_t)label[x], v = (uint64_t)label[y];
          P.push_back((u << 32) | v);
          P.push_back((v << 32) | u);
          break;
        }
      }
    });
  for (auto& x : th) x.join();
  size_t total = 0;
  for (auto& P : part) total += P.size();
  std::vector<uint64_t> keys;
  keys.reserve(total);
  for (auto& P : part) {
    keys.insert(keys.end(), P.begin(), P.end());
    std::vector<uint64_t>().swap(P);
  }
  return build_csr(keys, n, weights, seed, out);
}

// Matrix Market reader with the reference loader's semantics
// (src/mmio/mmio_wrapper.cu:76-295).
mgl_status mgl_read_mtx(const char* path, mgl_host_csr* out) {
  if (!path || !out) return MGL_INVALID_ARG;
  memset(out, 0, sizeof(*out));
  FILE* f = fopen(path, "r");
  if (!f) return MGL_INVALID_ARG;
  std::string line;
  char buf[4096];
  if (!fgets(buf, sizeof(buf), f)) { fclose(f); return MGL_INVALID_GRAPH; }
  bool symmetric = false;
  // banner: "general" keeps lower-triangle entries with weights; "symmetric"
  // uses unit weights (mmio_wrapper.cu:97-102, :141-142)
  if (strstr(buf, "general")) symmetric = false;
  else if (strstr(buf, "symmetric")) symmetric = true;
  do {
    if (!fgets(buf, sizeof(buf), f)) { fclose(f); return MGL_INVALID_GRAPH; }
  } while (buf[0] == '%');
  long long nr = 0, nc = 0, nz = 0;
  if (sscanf(buf, "%lld %lld %lld", &nr, &nc, &nz) != 3 || nr < 0 || nz < 0) {
    fclose(f);
    return MGL_INVALID_GRAPH;
  }
  if (nr > 0x7fffffffLL || nc > 0x7fffffffLL) { fclose(f); return MGL_UNSUPPORTED; }
  const int64_t n = nr;
  std::vector<uint64_t> keys;
  std::vector<double> wv;
  keys.reserve((size_t)nz * 2);
  if (!symmetric) wv.reserve((size_t)nz * 2);
  while (fgets(buf, sizeof(buf), f)) {
    char* p = buf;
    char* q;
    const long long r = strtoll(p, &q, 10);
    if (q == p) continue;
    p = q;
    const long long c = strtoll(p, &q, 10);
    if (q == p) continue;
    p = q;
    if (r < 1 || c < 1 || r > n || c > n) { fclose(f); return MGL_INVALID_GRAPH; }
    const uint64_t u = (uint64_t)(r - 1), v = (uint64_t)(c - 1);
    if (symmetric) {
      // every entry kept, off-diagonal mirrored (mmio_wrapper.cu:148-170)
      keys.push_back((u << 32) | v);
      if (u != v) keys.push_back((v << 32) | u);
    } else {
      // only col < row kept, weight parsed or 1.0, mirrored (:233-263)
      if (v >= u) continue;
      double w = strtod(p, &q);
      if (q == p) w = 1.0;
      keys.push_back((u << 32) | v);
      keys.push_back((v << 32) | u);
      wv.push_back(w);
      wv.push_back(w);
    }
  }
  fclose(f);
  const int64_t m = (int64_t)keys.size();
  out->n = n;
  out->nnz = m;
  out->offsets = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  out->cols = (int32_t*)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
  // weights stay doubles, as the reference parses them into weight_t (:243-250)
  out->w64 = symmetric ? nullptr : (double*)malloc(sizeof(double) * (m > 0 ? m : 1));
  if (!out->offsets || !out->cols || (!symmetric && !out->w64)) {
    mgl_csr_free(out);
    return MGL_OOM;
  }
  // rows sorted by column (stable for equal (row, col), which keeps duplicates)
  std::vector<int64_t> idx(m);
  for (int64_t i = 0; i < m; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) { return keys[x] < keys[y]; });
  for (int64_t i = 0; i < m; ++i) {
    const uint64_t k = keys[idx[i]];
    out->offsets[(k >> 32) + 1]++;
    out->cols[i] = (int32_t)(uint32_t)k;
    if (!symmetric) out->w64[i] = wv[idx[i]];
  }
  for (int64_t v = 0; v < n; ++v) out->offsets[v + 1] += out->offsets[v];
  return MGL_OK;
}

mgl_status mgl_write_assign(const char* path, const int32_t* assign, int64_t n) {
  if (!path || (n > 0 && !assign)) return MGL_INVALID_ARG;
  FILE* f = fopen(path, "w");
  if (!f) return MGL_INVALID_ARG;
  for (int64_t i = 0; i < n; ++i) fprintf(f, "%d\n", assign[i]);
  fclose(f);
  return MGL_OK;
}

static mgl_status write_assign_bin(const char* path, const int32_t* a, int64_t n) {
  FILE* f = fopen(path, "wb");
  if (!f) return MGL_INVALID_ARG;
  const char tag[8] = {'M', 'G', 'L', 'A', 'S', 'G', '1', 0};
  bool ok = fwrite(tag, 1, 8, f) == 8 && fwrite(&n, 8, 1, f) == 1 &&
            (n == 0 || fwrite(a, 4, (size_t)n, f) == (size_t)n);
  ok = (fclose(f) == 0) && ok;
  return ok ? MGL_OK : MGL_INVALID_ARG;
}

static mgl_status write_graph(const char* path, int binary, int64_t n, int64_t nnz,
                              const int64_t* off, const int32_t* col, const double* w) {
  FILE* f = fopen(path, binary ? "wb" : "w");
  if (!f) return MGL_INVALID_ARG;
  bool ok = true;
  if (binary) {
    const char tag[8] = {'M', 'G', 'L', 'C', 'S', 'R', '1', 0};
    ok = fwrite(tag, 1, 8, f) == 8 && fwrite(&n, 8, 1, f) == 1 && fwrite(&nnz, 8, 1, f) == 1 &&
         fwrite(off, 8, (size_t)n + 1, f) == (size_t)n + 1 &&
         (nnz == 0 || (fwrite(col, 4, (size_t)nnz, f) == (size_t)nnz &&
                       fwrite(w, 8, (size_t)nnz, f) == (size_t)nnz));
  } else {
    ok = fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%lld %lld %lld\n",
                 (long long)n, (long long)n, (long long)nnz) > 0;
    for (int64_t v = 0; ok && v < n; ++v)
      for (int64_t e = off[v]; ok && e < off[v + 1]; ++e)
        ok = fprintf(f, "%lld %d %.17g\n", (long long)(v + 1), col[e] + 1, w[e]) > 0;
  }
  ok = (fclose(f) == 0) && ok;
  return ok ? MGL_OK : MGL_INVALID_ARG;
}

mgl_status mgl_write_result(const mgl_result* r, const char* prefix, int binary) {
  if (!r || !prefix || binary < 0 || binary > 1) return MGL_INVALID_ARG;
  const std::string p = prefix;
  const char* ax = binary ? ".bin" : ".txt";
  const int L = mgl_result_levels(r);
  FILE* f = fopen((p + ".levels.txt").c_str(), "w");
  if (!f) return MGL_INVALID_ARG;
  fprintf(f, "# level vertices iterations Q trajectory...\n");
  for (int l = 0; l < L; ++l) {
    const int it = mgl_result_level_iters(r, l);
    fprintf(f, "%d %lld %d %.17g", l, (long long)mgl_result_level_n(r, l), it,
            mgl_result_level_q(r, l));
    const double* qt = mgl_result_level_qtraj(r, l);
    for (int i = 0; qt && i < it; ++i) fprintf(f, " %.17g", qt[i]);
    fprintf(f, "\n");
  }
  if (fclose(f) != 0) return MGL_INVALID_ARG;
  const int64_t n0 = L > 0 ? mgl_result_level_n(r, 0) : 0;
  const int32_t* fin = mgl_result_final_assign(r);
  mgl_status s = binary ? write_assign_bin((p + ".final" + ax).c_str(), fin, fin ? n0 : 0)
                        : mgl_write_assign((p + ".final" + ax).c_str(), fin, fin ? n0 : 0);
  if (s) return s;
  for (int l = 0; l < L; ++l) {
    const std::string base = p + ".level" + std::to_string(l);
    const int32_t* a = mgl_result_level_assign(r, l);
    if (a) {
      const int64_t n = mgl_result_level_n(r, l);
      s = binary ? write_assign_bin((base + ax).c_str(), a, n)
                 : mgl_write_assign((base + ax).c_str(), a, n);
      if (s) return s;
    }
    int64_t n = 0, nnz = 0;
    const int64_t* off = nullptr;
    const int32_t* col = nullptr;
    const double* w = nullptr;
    if (l >= 1 && mgl_result_level_graph(r, l, &n, &nnz, &off, &col, &w) == MGL_OK) {
      s = write_graph((base + (binary ? ".csr" : ".mtx")).c_str(), binary, n, nnz, off, col, w);
      if (s) return s;
    }
  }
  return MGL_OK;
}

}  // extern "C"
