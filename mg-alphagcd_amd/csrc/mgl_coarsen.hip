// Graph coarsening by sort + segmented reduce (gfx950).
//
// Reference behaviour (src/coarsen_graph/coarsen_graph_mg.cu):
//   renumbering : flag every used community id, inclusive scan, id <- scan-1,
//                 i.e. the dense rank in ascending old id (:211-226, :267-286,
//                 :1330-1363)
//   aggregation : coarse entry (rank c[u], rank c[v]) carries the sum of the
//                 weights of all fine entries between the two communities; the
//                 intra-community entries become the self-loop (:816-917,
//                 :919-1020, :1022-1125); install :2224-2228
// The reference builds each coarse row with a per-community hash and emits it
// in slot order; here every fine entry becomes a 64-bit key
// (row << b | col), the keys are radix-sorted (own LSD sort, 8-bit digits,
// stable) and equal keys are reduced, which yields rows sorted by column.
#include "mgl_kernels.h"
#include <type_traits>

static unsigned gridc(int64_t n, int per, unsigned cap) {
  int64_t g = (n + per - 1) / per;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// ---------------------------------------------------------------------------
// Exclusive scan (int32/int64 in -> int64 out, n+1 outputs: out[n] = total).
#define SCAN_T 256
#define SCAN_I 8
#define SCAN_TILE (SCAN_T * SCAN_I)

__device__ __forceinline__ int64_t wave_incl_scan(int64_t x) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int lo = __shfl_up((int)(uint32_t)x, d, 64);
    int hi = __shfl_up((int)(uint32_t)((uint64_t)x >> 32), d, 64);
    int64_t y = (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
    if (l >= d) x += y;
  }
  return x;
}

// value of lane src of the wavefront
__device__ __forceinline__ int64_t bcast64(int64_t x, int src) {
  const int lo = __shfl((int)(uint32_t)x, src, 64);
  const int hi = __shfl((int)(uint32_t)((uint64_t)x >> 32), src, 64);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// block-wide exclusive scan of one value per thread; returns the block total
__device__ __forceinline__ int64_t block_excl_scan(int64_t x, int64_t& total) {
  __shared__ int64_t ws[SCAN_T / 64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  int64_t inc = wave_incl_scan(x);
  if (l == 63) ws[w] = inc;
  __syncthreads();
  int64_t pre = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < SCAN_T / 64; ++i) {
    if (i < w) pre += ws[i];
    tot += ws[i];
  }
  __syncthreads();
  total = tot;
  return pre + inc - x;
}

template <class T>
__global__ __launch_bounds__(SCAN_T) void k_tile_sums(const T* __restrict__ in, int64_t n,
                                                      int64_t* __restrict__ sums) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_I; ++i) {
    const int64_t j = base + (int64_t)i * SCAN_T + threadIdx.x;
    if (j < n) s += (int64_t)in[j];
  }
  int64_t tot;
  (void)block_excl_scan(s, tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <class T>
__global__ __launch_bounds__(SCAN_T) void k_tile_scan(const T* __restrict__ in, int64_t n,
                                                      const int64_t* __restrict__ pref,
                                                      int64_t* __restrict__ out) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_I;
  int64_t v[SCAN_I];
  int64_t s = 0;
#pragma unroll
  for (int i = 0; i < SCAN_I; ++i) {
    const int64_t j = base + i;
    v[i] = j < n ? (int64_t)in[j] : 0;
    s += v[i];
  }
  int64_t tot;
  int64_t run = block_excl_scan(s, tot) + (pref ? pref[blockIdx.x] : 0);
#pragma unroll
  for (int i = 0; i < SCAN_I; ++i) {
    const int64_t j = base + i;
    if (j < n) out[j] = run;
    run += v[i];
  }
  // the thread that owns element n-1 writes the total after it
  if (base <= n - 1 && n - 1 < base + SCAN_I) out[n] = run;
}

size_t DevScan::temp_bytes(int64_t n) {
  size_t b = 0;
  int64_t m = n;
  while (m > SCAN_TILE) {
    const int64_t nb = (m + SCAN_TILE - 1) / SCAN_TILE;
    b += (size_t)(nb + 1) * 8 * 2;
    m = nb;
  }
  return b + 256;
}

template <class T>
static void scan_rec(const T* in, int64_t* out, int64_t n, int64_t* temp, hipStream_t st) {
  if (n <= 0) {
    launch_fill_bytes0(out, 8, st);
    return;
  }
  const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (nb == 1) {
    k_tile_scan<T><<<1, SCAN_T, 0, st>>>(in, n, nullptr, out);
    return;
  }
  int64_t* sums = temp;
  int64_t* pref = temp + (nb + 1);
  k_tile_sums<T><<<(unsigned)nb, SCAN_T, 0, st>>>(in, n, sums);
  scan_rec<int64_t>(sums, pref, nb, temp + 2 * (nb + 1), st);
  k_tile_scan<T><<<(unsigned)nb, SCAN_T, 0, st>>>(in, n, pref, out);
}

void scan_exclusive_i32(const int32_t* in, int64_t* out, int64_t n, void* temp, hipStream_t st) {
  scan_rec<int32_t>(in, out, n, (int64_t*)temp, st);
}
void scan_exclusive_i64(const int64_t* in, int64_t* out, int64_t n, void* temp, hipStream_t st) {
  scan_rec<int64_t>(in, out, n, (int64_t*)temp, st);
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (u64 key, value) pairs, 8-bit digits.
#define RS_T 256
#define RS_I 16
#define RS_TILE (RS_T * RS_I)

__global__ __launch_bounds__(RS_T) void k_rs_hist(const uint64_t* __restrict__ keys, int64_t n,
                                                  int shift, int64_t nblk,
                                                  int64_t* __restrict__ hist) {
  __shared__ unsigned int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
  for (int i = 0; i < RS_I; ++i) {
    const int64_t j = base + (int64_t)i * RS_T + threadIdx.x;
    if (j < n) atomicAdd(&h[(keys[j] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// One tile per workgroup: every key gets its stable rank within its digit
// (wavefront ballots per round of 256 keys, running counts over the rounds),
// the tile is staged in LDS in digit order, and then written out so that
// consecutive threads store consecutive positions of one digit's run.
template <class VT, bool HAS_V>
__global__ __launch_bounds__(RS_T) void k_rs_scatter(const uint64_t* __restrict__ kin,
                                                     const VT* __restrict__ vin, int64_t n,
                                                     int shift, int64_t nblk,
                                                     const int64_t* __restrict__ hist,
                                                     const int64_t* __restrict__ offs,
                                                     uint64_t* __restrict__ kout,
                                                     VT* __restrict__ vout) {
  __shared__ uint64_t sk[RS_TILE];
  __shared__ VT sv[HAS_V ? RS_TILE : 1];
  __shared__ int64_t gbase[256];
  __shared__ unsigned int tstart[256];
  __shared__ unsigned int run[256];
  __shared__ unsigned int wc[RS_T / 64][256];
  __shared__ unsigned int wsum[RS_T / 64];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  {
    // digit offsets inside the tile: exclusive scan of the tile's histogram
    const unsigned c = (unsigned)hist[(int64_t)threadIdx.x * nblk + blockIdx.x];
    unsigned x = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const unsigned y = __shfl_up(x, d, 64);
      if (l >= d) x += y;
    }
    if (l == 63) wsum[w] = x;
    gbase[threadIdx.x] = offs[(int64_t)threadIdx.x * nblk + blockIdx.x];
    run[threadIdx.x] = 0;
    __syncthreads();
    unsigned pre = 0;
#pragma unroll
    for (int i = 0; i < RS_T / 64; ++i) pre += (i < w) ? wsum[i] : 0u;
    tstart[threadIdx.x] = pre + x - c;
  }
  const uint64_t lt = (l == 0) ? 0ull : (~0ull >> (64 - l));
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  const int tn = (int)((n - base < RS_TILE) ? n - base : RS_TILE);
  // the tile's keys (and values) in registers, all loads in flight at once
  uint64_t kr[RS_I];
  VT vr[RS_I];
#pragma unroll
  for (int r = 0; r < RS_I; ++r) {
    const int jt = r * RS_T + threadIdx.x;
    kr[r] = jt < tn ? kin[base + jt] : 0;
    if (HAS_V) vr[r] = jt < tn ? vin[base + jt] : VT{};
  }
#pragma unroll
  for (int r = 0; r < RS_I; ++r) {
#pragma unroll
    for (int i = 0; i < RS_T / 64; ++i) wc[i][threadIdx.x] = 0;
    __syncthreads();
    const int jt = r * RS_T + threadIdx.x;
    const bool ok = jt < tn;
    const uint64_t key = kr[r];
    const VT val = HAS_V ? vr[r] : VT{};
    const unsigned d = (unsigned)((key >> shift) & 255u);
    // lanes of this wave holding the same digit
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const uint64_t bmask = __ballot((d >> bit) & 1u);
      peers &= ((d >> bit) & 1u) ? bmask : ~bmask;
    }
    const unsigned rank_w = (unsigned)__popcll(peers & lt);
    if (ok && rank_w == 0) wc[w][d] = (unsigned)__popcll(peers);
    __syncthreads();
    {
      // thread t owns digit t: prefix over waves, then advance the running count
      const unsigned t = threadIdx.x;
      unsigned acc = run[t];
#pragma unroll
      for (int i = 0; i < RS_T / 64; ++i) {
        const unsigned c = wc[i][t];
        wc[i][t] = acc;
        acc += c;
      }
      run[t] = acc;
    }
    __syncthreads();
    if (ok) {
      const unsigned p = tstart[d] + wc[w][d] + rank_w;
      sk[p] = key;
      if (HAS_V) sv[p] = val;
    }
    __syncthreads();  // wc is read above and cleared by the next round
  }
  for (int p = threadIdx.x; p < tn; p += RS_T) {
    const uint64_t key = sk[p];
    const unsigned d = (unsigned)((key >> shift) & 255u);
    const int64_t pos = gbase[d] + (p - (int)tstart[d]);
    kout[pos] = key;
    if (HAS_V) vout[pos] = sv[p];
  }
}

template <class VT, bool HAS_V>
static uint64_t* radix_sort(uint64_t* k0, uint64_t* k1, VT* v0, VT* v1, int64_t n, int bits,
                            int64_t* hist, void* scan_tmp, VT** vres, hipStream_t st) {
  uint64_t* ki = k0;
  uint64_t* ko = k1;
  VT* vi = v0;
  VT* vo = v1;
  const int64_t nblk = (n + RS_TILE - 1) / RS_TILE;
  for (int shift = 0; shift < bits; shift += 8) {
    k_rs_hist<<<(unsigned)nblk, RS_T, 0, st>>>(ki, n, shift, nblk, hist);
    int64_t* offs = hist + 256 * nblk + 1;
    scan_exclusive_i64(hist, offs, 256 * nblk, scan_tmp, st);
    k_rs_scatter<VT, HAS_V><<<(unsigned)nblk, RS_T, 0, st>>>(ki, vi, n, shift, nblk, hist, offs,
                                                                ko, vo);
    uint64_t* tk = ki; ki = ko; ko = tk;
    VT* tv = vi; vi = vo; vo = tv;
  }
  *vres = vi;
  return ki;
}

// ---------------------------------------------------------------------------
__global__ void k_flag(int64_t n, const vid* __restrict__ c, const int32_t* __restrict__ gid,
                       const int32_t* __restrict__ keep, int32_t* __restrict__ flag) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    if (keep && !keep[gid ? gid[v] : 0]) continue;
    flag[c[v]] = 1;
  }
}

__global__ void k_newc(int64_t n, const vid* __restrict__ c, const int32_t* __restrict__ gid,
                       const int32_t* __restrict__ keep, const int64_t* __restrict__ rank,
                       vid* __restrict__ newc) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x) {
    if (keep && !keep[gid ? gid[v] : 0]) { newc[v] = -1; continue; }
    newc[v] = (vid)rank[c[v]];
  }
}

// Keys of every entry of a kept row of the group (rows [v_lo, v_hi), entries
// [e_lo, e_hi)), written at e + delta[g] - kbase (compaction of the kept
// graphs' entry ranges; kbase = the group's first compacted position). The
// key holds the coarse row and column relative to the group's first coarse
// row r_lo (every entry of a graph stays inside its graph, hence inside its
// group). Entry-parallel: a wavefront takes 256 consecutive entries, finds the
// rows spanning them once, and each lane finds its entry's row by binary
// search over that short range of offsets.
__device__ __forceinline__ int64_t row_of(const eid* __restrict__ off, int64_t lo, int64_t hi,
                                          int64_t e) {
  // largest r in [lo, hi] with off[r] <= e
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (off[mid] <= e) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <class M, class VT>
__global__ __launch_bounds__(256) void k_keys(int64_t v_lo, int64_t v_hi, int64_t e_lo,
                                              int64_t e_hi, const eid* __restrict__ off,
                                              const vid* __restrict__ col, const void* w,
                                              const vid* __restrict__ newc,
                                              const int32_t* __restrict__ gid,
                                              const int64_t* __restrict__ delta, int64_t kbase,
                                              int64_t r_lo, int bits, uint64_t* __restrict__ keys,
                                              VT* __restrict__ vals) {
  const int64_t wave = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t e0 = e_lo + wave * 256;
  if (e0 >= e_hi) return;
  const int64_t e1 = (e0 + 256 < e_hi) ? e0 + 256 : e_hi;
  int64_t r = 0;
  if (lane < 2) r = row_of(off, v_lo, v_hi - 1, lane == 0 ? e0 : e1 - 1);
  const int64_t rlo = bcast64(r, 0), rhi = bcast64(r, 1);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t e = e0 + lane + 64 * j;
    if (e >= e1) continue;
    const int64_t v = row_of(off, rlo, rhi, e);
    const vid rr = newc[v];
    if (rr < 0) continue;
    const int64_t dl = (delta ? delta[gid ? gid[v] : 0] : 0) - kbase;
    const uint64_t cc = (uint64_t)((int64_t)newc[col[e]] - r_lo);
    keys[e + dl] = ((uint64_t)((int64_t)rr - r_lo) << bits) | cc;
    if (M::HAS_W) vals[e + dl] = (VT)((const typename M::W*)w)[e];
  }
}

__global__ void k_fill_off(eid* __restrict__ off, int64_t n, int64_t val) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    off[i] = val;
}

__global__ void k_heads(const uint64_t* __restrict__ k, int64_t n, int32_t* __restrict__ head) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    head[i] = (i == 0 || k[i] != k[i - 1]) ? 1 : 0;
}

// off2[lo..hi] = val, lo..hi of each lane; short ranges by the lane itself,
// long ones (runs of empty coarse rows) by the whole wavefront
__device__ __forceinline__ void fill_rows(eid* __restrict__ off2, int64_t lo, int64_t hi,
                                          int64_t val) {
  const int lane = threadIdx.x & 63;
  const int64_t len = hi - lo + 1;
  if (len > 0 && len <= 16)
    for (int64_t r = lo; r <= hi; ++r) off2[r] = val;
  uint64_t big = __ballot(len > 16);
  while (big) {
    const int src = __ffsll((unsigned long long)big) - 1;
    const int64_t blo = bcast64(lo, src), bhi = bcast64(hi, src), bv = bcast64(val, src);
    for (int64_t r = blo + lane; r <= bhi; r += 64) off2[r] = bv;
    big &= big - 1;
  }
}

// Coarse entries from the sorted keys. Equal keys are contiguous: each
// wavefront sums the runs it holds with a segmented scan; a run that starts
// and ends inside the wavefront is stored by its last lane, a run crossing a
// wavefront boundary is added atomically (one atomic per wavefront it
// touches). Row offsets come from the heads that change row: the head writes
// the offsets of its row and of the empty rows before it; the last entry
// writes those of the empty rows after it.
template <class T>
__device__ __forceinline__ T shfl_up_t(T v, int d) {
  if constexpr (sizeof(T) == 8) {
    uint64_t x;
    __builtin_memcpy(&x, &v, 8);
    int lo = __shfl_up((int)(uint32_t)x, d, 64);
    int hi = __shfl_up((int)(uint32_t)(x >> 32), d, 64);
    uint64_t y = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
    T r;
    __builtin_memcpy(&r, &y, 8);
    return r;
  } else {
    return __shfl_up(v, d, 64);
  }
}

// Group form: rows are relative to the group's first coarse row r_lo and the
// group's last row is r_hi (off2[r_hi] is written too); coarse entries go to
// positions U0 + u.
template <class VT, class OT, bool HAS_V>
__global__ __launch_bounds__(256) void k_emit(const uint64_t* __restrict__ k,
                                              const VT* __restrict__ v, int64_t n,
                                              const int32_t* __restrict__ head,
                                              const int64_t* __restrict__ uidx, int bits,
                                              vid* __restrict__ col2, OT* __restrict__ w2,
                                              eid* __restrict__ off2, int64_t r_lo, int64_t r_hi,
                                              int64_t U0, int64_t U) {
  const uint64_t mask = (bits >= 64) ? ~0ull : ((1ull << bits) - 1ull);
  const int lane = threadIdx.x & 63;
  for (int64_t base = (int64_t)blockIdx.x * 256 + (threadIdx.x & ~63); base < n;
       base += (int64_t)gridDim.x * 256) {
    const int64_t i = base + lane;
    const bool ok = i < n;
    const uint64_t key = ok ? k[i] : ~0ull;
    OT x = ok ? (HAS_V ? (OT)v[i] : (OT)1) : (OT)0;
    int64_t u = ok ? uidx[i + 1] - 1 : -1;
    int64_t glo = 0, ghi = -1, tlo = 0, thi = -1;
    if (ok && head[i]) {
      col2[U0 + u] = (vid)((int64_t)(key & mask) + r_lo);
      glo = r_lo + ((i == 0) ? 0 : (int64_t)(k[i - 1] >> bits) + 1);
      ghi = r_lo + (int64_t)(key >> bits);
    }
    if (ok && i == n - 1) {
      tlo = r_lo + (int64_t)(key >> bits) + 1;
      thi = r_hi;
    }
    fill_rows(off2, glo, ghi, U0 + u);
    fill_rows(off2, tlo, thi, U0 + U);
    // segmented inclusive scan over equal u within the wavefront
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const OT y = shfl_up_t(x, d);
      const int64_t uy = shfl_up_t(u, d);
      if (lane >= d && uy == u) x += y;
    }
    int64_t u_next, u_first;
    {
      int lo = __shfl_down((int)(uint32_t)(uint64_t)u, 1, 64);
      int hi = __shfl_down((int)(uint32_t)((uint64_t)u >> 32), 1, 64);
      u_next = (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
      lo = __shfl((int)(uint32_t)(uint64_t)u, 0, 64);
      hi = __shfl((int)(uint32_t)((uint64_t)u >> 32), 0, 64);
      u_first = (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
    }
    const bool at_end = (i + 1 >= n);
    const bool last = ok && (lane == 63 || u_next != u || at_end);
    if (last) {
      // a run wholly inside this wavefront is stored; a run crossing a
      // wavefront boundary is accumulated atomically (w2 starts at zero)
      const bool started = (u != u_first) || head[base];
      const bool ended = at_end || (lane < 63 ? u_next != u : head[i + 1] != 0);
      if (started && ended) {
        w2[U0 + u] = x;
      } else if constexpr (std::is_same_v<OT, int64_t>) {
        atomicAdd((unsigned long long*)&w2[U0 + u], (unsigned long long)x);
      } else {
        atomicAdd(&w2[U0 + u], x);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Renumbering of the whole level (louvain-level dense rank of the used ids of
// the kept graphs): rank[0..n] (rank[n] = C) and newc[v] (-1 in dropped graphs).
static int coarsen_rank(int64_t n, const vid* c, const int32_t* gid, const int32_t* keep,
                        CoarsenBufs& B, int64_t* C_out, hipStream_t st) {
  launch_fill_bytes0(B.flag, sizeof(int32_t) * (size_t)(n > 0 ? n : 1), st);
  k_flag<<<gridc(n, 256, 8192), 256, 0, st>>>(n, c, gid, keep, B.flag);
  scan_exclusive_i32(B.flag, B.rank, n, B.scan_tmp, st);
  int64_t C = 0;
  if (hipMemcpyAsync(&C, B.rank + n, 8, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
  if (hipStreamSynchronize(st) != hipSuccess) return -1;
  k_newc<<<gridc(n, 256, 8192), 256, 0, st>>>(n, c, gid, keep, B.rank, B.newc);
  *C_out = C;
  return 0;
}

// One group of consecutive graphs: fine rows [v_lo, v_hi) and entries
// [e_lo, e_hi) of which e_keep are kept (compacted from kbase on), coarse rows
// [r_lo, r_hi); the coarse entries are written from U0 on and their number
// returned in *U_out.
template <class M, class VT, bool HAS_V, class OT>
static int coarsen_group_t(int64_t v_lo, int64_t v_hi, int64_t e_lo, int64_t e_hi,
                           const eid* off, const vid* col, const void* w, const int32_t* gid,
                           const int64_t* delta, int64_t kbase, int64_t e_keep, int64_t r_lo,
                           int64_t r_hi, int64_t U0, CoarsenBufs& B, eid* off2, vid* col2,
                           void* w2, int64_t* U_out, hipStream_t st) {
  const int64_t Cg = r_hi - r_lo;
  *U_out = 0;
  if (Cg <= 0) return 0;
  if (e_keep <= 0) {
    // coarse vertices without entries (kept graphs of isolated vertices)
    k_fill_off<<<gridc(Cg + 1, 256, 4096), 256, 0, st>>>(off2 + r_lo, Cg + 1, U0);
    return 0;
  }
  if (e_keep > B.cap) return -2;
  int bits = 1;
  while ((1ll << bits) < Cg) ++bits;
  const int kbits = 2 * bits;
  const int64_t waves = (e_hi - e_lo + 255) / 256;
  k_keys<M, VT><<<(unsigned)((waves + 3) / 4), 256, 0, st>>>(
      v_lo, v_hi, e_lo, e_hi, off, col, w, B.newc, gid, delta, kbase, r_lo, bits, B.k0,
      (VT*)B.v0);
  const int64_t need = 256 * ((e_keep + RS_TILE - 1) / RS_TILE) * 2 + 2;
  if (need > B.hist_cap) return -2;
  VT* vs = nullptr;
  uint64_t* ks = radix_sort<VT, HAS_V>(B.k0, B.k1, (VT*)B.v0, (VT*)B.v1, e_keep, kbits, B.hist,
                                       B.scan_tmp, &vs, st);
  k_heads<<<gridc(e_keep, 256, 8192), 256, 0, st>>>(ks, e_keep, B.head);
  // uidx[0] = 0, uidx[i+1] = inclusive count
  scan_exclusive_i32(B.head, B.uidx, e_keep, B.scan_tmp, st);
  int64_t U = 0;
  if (hipMemcpyAsync(&U, B.uidx + e_keep, 8, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
  if (hipStreamSynchronize(st) != hipSuccess) return -1;
  // exclusive scan gives uidx[i] = count before i; emit needs the inclusive
  // count at i, i.e. uidx[i+1]; out[e_keep] = total makes that valid for all i
  launch_fill_bytes0((OT*)w2 + U0, sizeof(OT) * (size_t)(U > 0 ? U : 1), st);
  k_emit<VT, OT, HAS_V><<<gridc(e_keep, 256, 16384), 256, 0, st>>>(
      ks, vs, e_keep, B.head, B.uidx, bits, col2, (OT*)w2, off2, r_lo, r_hi, U0, U);
  *U_out = U;
  return 0;
}

int coarsen_renumber(int64_t n, const vid* c, const int32_t* gid, const int32_t* keep,
                     CoarsenBufs& B, int64_t* C_out, hipStream_t st) {
  return coarsen_rank(n, c, gid, keep, B, C_out, st);
}

int coarsen_group(int mode_in, int mode_out, int64_t v_lo, int64_t v_hi, int64_t e_lo,
                  int64_t e_hi, const eid* off, const vid* col, const void* w, const int32_t* gid,
                  const int64_t* delta, int64_t kbase, int64_t e_keep, int64_t r_lo,
                  int64_t r_hi, int64_t U0, CoarsenBufs& B, eid* off2, vid* col2, void* w2,
                  int64_t* U_out, hipStream_t st) {
#define C_(M_, VT_, HV_, OT_)                                                                 \
  return coarsen_group_t<M_, VT_, HV_, OT_>(v_lo, v_hi, e_lo, e_hi, off, col, w, gid, delta, \
                                            kbase, e_keep, r_lo, r_hi, U0, B, off2, col2, w2, \
                                            U_out, st)
  switch (mode_in) {
    case MODE_UNIT:
      if (mode_out == MODE_I32) C_(ModeUnit, int32_t, false, int32_t);
      C_(ModeUnit, int64_t, false, int64_t);
    case MODE_I32: C_(ModeI32, int32_t, true, int32_t);
    case MODE_INT: C_(ModeInt, int64_t, true, int64_t);
    case MODE_F64: C_(ModeF64, double, true, double);
    case MODE_F32: C_(ModeF32, float, true, float);
  }
#undef C_
  return -1;
}

int coarsen_level(int mode_in, int mode_out, int64_t n, const eid* off, const vid* col,
                  const void* w, const vid* c, const int32_t* gid, const int32_t* keep,
                  const int64_t* delta, int64_t e_keep, CoarsenBufs& B, eid* off2, vid* col2,
                  void* w2, int64_t* C_out, int64_t* E2_out, hipStream_t st) {
  int64_t C = 0;
  int rc = coarsen_rank(n, c, gid, keep, B, &C, st);
  if (rc) return rc;
  int64_t nnz = 0;
  if (n > 0) {
    if (hipMemcpyAsync(&nnz, off + n, 8, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
    if (hipStreamSynchronize(st) != hipSuccess) return -1;
  }
  int64_t U = 0;
  rc = coarsen_group(mode_in, mode_out, 0, n, 0, nnz, off, col, w, gid, delta, 0, e_keep, 0, C,
                     0, B, off2, col2, w2, &U, st);
  if (rc) return rc;
  if (C == 0) launch_fill_bytes0(off2, sizeof(eid), st);
  *C_out = C;
  *E2_out = U;
  return 0;
}
