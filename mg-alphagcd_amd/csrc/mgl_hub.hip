// Hub path of the move pass: rows too long for one workgroup's LDS table.
//
// The reference gives each such vertex one thread block and a hash table of
// 2*max_degree slots in global memory, filled with atomicCAS/atomicAdd per
// edge (src/louvain/louvain.cu:810-1030, table allocated at :1262-1263).
// Here a hub's row is split into chunks of HUB_CHUNK entries, and its
// candidates into P buckets (P ~ degree / bucket size) by a hash of the
// community id. No step issues a global atomic per edge. Segmented form (the
// default while no hub of the level has more than HUB_SEG_MAX chunks):
//   count    : one workgroup per chunk gathers the neighbour record {c, a}
//              once per entry, sums the entries of the own community (eici,
//              self-loop) and writes the chunk's candidates (ids and totals
//              apart) into the chunk's region grouped by bucket (counting
//              sort in LDS), with the chunk's bucket offsets
//   aggregate: one workgroup per bucket reads the bucket's segment of every
//              chunk, sums equal ids in an LDS table (an HBM table if a bucket
//              is unusually large), evaluates the gain of each distinct id
//              and keeps the bucket's best
//   finalize : per hub, the best over its buckets, the bi-probing clamp
// Scatter form (longer hubs): count (bucket histogram, candidates kept in
// the chunk's region), scan of the bucket counts, scatter into contiguous
// buckets, aggregate, finalize.
// Sums are exact in the integer modes and the argmax is a total order, so the
// result does not depend on the bucket layout or the order of the scatter.
#pragma clang fp contract(off)

#include "mgl_kernels.h"
#include "mgl_tables.h"

#define HUB_CHUNK 8192   // entries of one count workgroup (8 to 32 per thread)
#define HUB_T 256
#define HUB_ITEMS (HUB_CHUNK / HUB_T)
#define HUB_MAXP 1024
#define HUB_SEG_MAX 512   // chunks per hub in the segmented form
// Aggregate workgroups: BS threads (HubPlan::agg_bs: 256, 512 or 1024),
// 6,144 / BS candidates per thread on the owner path.
#define HUB_AGG_CAP 4096

__device__ __forceinline__ uint32_t hub_bucket(vid key, uint32_t P) {
  return (uint32_t)(((uint64_t)((uint32_t)key * 2654435761u) * P) >> 32);
}
__device__ __forceinline__ uint32_t hub_slot(vid key, uint32_t H) {
  uint32_t x = (uint32_t)key * 0x85EBCA6Bu;
  x ^= x >> 15;
  x *= 0xC2B2AE35u;
  return (uint32_t)(((uint64_t)x * H) >> 32);
}

__device__ __forceinline__ void hatom_add(int32_t* p, int32_t x) { atomicAdd(p, x); }
__device__ __forceinline__ void hatom_add(int64_t* p, int64_t x) {
  atomicAdd((unsigned long long*)p, (unsigned long long)x);
}
__device__ __forceinline__ void hatom_add(float* p, float x) { atomicAdd(p, x); }
__device__ __forceinline__ void hatom_add(double* p, double x) { atomicAdd(p, x); }

// table of H slots (any H >= 1), linear probing; the creator of a key stores
// its community total
template <class ST, class AT, bool GLOBAL>
__device__ __forceinline__ void tab_insert(vid* K, ST* V, AT* Av, vid key, ST w, AT ak,
                                           uint32_t H) {
  uint32_t h = hub_slot(key, H);
  while (true) {
    vid cur = GLOBAL ? __hip_atomic_load(&K[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                     : __hip_atomic_load(&K[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cur == MGL_EMPTY) {
      cur = atomicCAS(&K[h], MGL_EMPTY, key);
      if (cur == MGL_EMPTY) Av[h] = ak;
    }
    if (cur == MGL_EMPTY || cur == key) { hatom_add(&V[h], w); return; }
    h = (h + 1u == H) ? 0u : h + 1u;
  }
}

// Sum over a workgroup of NT threads (red holds NT / 64 entries).
template <int NT, class T>
__device__ __forceinline__ T bsum_t(T v, T* red) {
  v = group_sum<64>(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  T s = 0;
#pragma unroll
  for (int i = 0; i < NT / 64; ++i) s += red[i];
  return s;
}

// ---- count ------------------------------------------------------------------
template <class M, class AT>
__global__ __launch_bounds__(HUB_T) void k_hub_count(MoveArgs A, HubPlan H) {
  typedef typename M::S ST;
  typedef typename M::K KT;
  __shared__ unsigned int hist[HUB_MAXP];
  __shared__ KT red[HUB_T / 64];
  const HubChunk hc = H.chunks[blockIdx.x];
  const int64_t i = hc.hub, e0 = hc.e0, e1 = hc.e1, pos0 = hc.pos0;
  const vid v = hc.v;
  if (!A.gs[hc.g].active) return;
  const uint32_t P = (uint32_t)hc.P;
  for (uint32_t b = threadIdx.x; b < P; b += HUB_T) hist[b] = 0;
  __syncthreads();
  const Rec<AT>* R = (const Rec<AT>*)A.rec;
  const vid vp = A.rpos ? A.rpos[v] : v;
  const vid ci = H.gin[i].ci;
  const bool v1 = A.version == 1;
  KT eici = 0, selfw = 0;
  for (int64_t e = e0 + threadIdx.x; e < e1; e += HUB_T) {
    const vid u = colload(A.col, e);
    const Rec<AT> r = R[u];
    const vid cj = r.c;
    const ST wv = wload<M>(A.w, e);
    const int64_t pos = pos0 + (e - e0);
    const bool cand = v1 ? (A.up ? cj < ci : cj > ci) : (cj != ci);
    if (cand) {
      H.tmp_c[pos] = cj;
      ((AT*)H.tmp_a)[pos] = r.a;
      if (M::HAS_W) ((ST*)H.tmp_w)[pos] = wv;
      atomicAdd(&hist[hub_bucket(cj, P)], 1u);
    } else {
      H.tmp_c[pos] = MGL_EMPTY;
      if (cj == ci) {
        if (u != vp) eici += (KT)wv;
        else selfw += (KT)wv;
      }
    }
  }
  eici = bsum_t<HUB_T>(eici, red);
  selfw = bsum_t<HUB_T>(selfw, red);
  if (threadIdx.x == 0) {
    if (eici != (KT)0) hatom_add(&((KT*)H.hub_eici)[i], eici);
    if (selfw != (KT)0) hatom_add(&((KT*)H.hub_self)[i], selfw);
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < P; b += HUB_T)
    if (hist[b]) atomicAdd((unsigned long long*)&H.bcount[H.hub_bk[i] + b], (unsigned long long)hist[b]);
}

// ---- segmented form: head and count --------------------------------------------
// Per pass: zero the hubs' own-community sums and the HBM-table pool, and
// record each hub's gain inputs (community, total, community total without it).
template <class M>
__global__ void k_hub_head(MoveArgs A, HubPlan H) {
  typedef typename M::K KT;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *H.galloc = 0;
  if (i < H.m) {
    ((int64_t*)H.hub_eici)[i] = 0;  // 8-byte sums (int64 or double zero)
    ((int64_t*)H.hub_self)[i] = 0;
    const vid v = H.hub_v[i];
    const vid ci = A.c[v];
    const KT ki = ((const KT*)A.k)[A.rpos ? A.rpos[v] : v];  // k by record position
    const KT acik = ((const KT*)A.a)[ci] - ki;
    HubGain hg;
    __builtin_memcpy(&hg.ki, &ki, 8);
    __builtin_memcpy(&hg.acik, &acik, 8);
    hg.ci = ci;
    hg.pad = 0;
    H.gin[i] = hg;
  }
}
template <class KT>
__device__ __forceinline__ KT gin_bits(int64_t x) {
  KT r;
  __builtin_memcpy(&r, &x, 8);
  return r;
}

// Gathers c[u] once per entry of a chunk, sums the own community (eici,
// self-loop), and writes the chunk's candidates into its region of tmp grouped
// by bucket (counting sort in LDS); cbo receives the chunk's bucket offsets.
template <class M, class AT, int CT>
__global__ __launch_bounds__(CT) void k_hub_count_seg(MoveArgs A, HubPlan H) {
  typedef typename M::S ST;
  typedef typename M::K KT;
  constexpr int CI = HUB_CHUNK / CT;  // entries per thread
  __shared__ unsigned int hist[HUB_MAXP + 1];
  __shared__ unsigned int wsum[CT / 64];
  __shared__ KT red[CT / 64];
  // the chunk record is the only load the rest waits on: the columns, the
  // graph's state and the hub's community are requested together after it
  const HubChunk hc = H.chunks[blockIdx.x];
  const int64_t i = hc.hub, e0 = hc.e0, e1 = hc.e1, pos0 = hc.pos0;
  const vid v = hc.v;
  const uint32_t P = (uint32_t)hc.P;
  vid key[CI];
  Rec<AT> r[CI];
#pragma unroll
  for (int j = 0; j < CI; ++j) {
    const int64_t e = e0 + (int64_t)j * CT + threadIdx.x;
    key[j] = e < e1 ? colload(A.col, e) : MGL_EMPTY;  // neighbour position first
  }
  const vid ci = H.gin[i].ci;
  const vid vp = A.rpos ? A.rpos[v] : v;
  if (!A.gs[hc.g].active) return;
  for (uint32_t b = threadIdx.x; b <= P; b += CT) hist[b] = 0;
  const Rec<AT>* R = (const Rec<AT>*)A.rec;
  const bool v1 = A.version == 1;
#pragma unroll
  for (int j = 0; j < CI; ++j)
    if (key[j] != MGL_EMPTY) r[j] = R[key[j]];
  __syncthreads();  // hist cleared
  KT eici = 0, selfw = 0;
  uint32_t slot[CI];  // bucket << 16 | rank within the chunk's bucket
#pragma unroll
  for (int j = 0; j < CI; ++j) {
    if (key[j] == MGL_EMPTY) continue;
    const vid u = key[j];
    const vid cj = r[j].c;
    const bool cand = v1 ? (A.up ? cj < ci : cj > ci) : (cj != ci);
    if (cand) {
      const uint32_t bk = hub_bucket(cj, P);
      slot[j] = (bk << 16) | atomicAdd(&hist[bk], 1u);
      key[j] = cj;
    } else {
      key[j] = MGL_EMPTY;
      if (cj == ci) {
        const ST wv = wload<M>(A.w, e0 + (int64_t)j * CT + threadIdx.x);
        if (u != vp) eici += (KT)wv;
        else selfw += (KT)wv;
      }
    }
  }
  __syncthreads();
  // exclusive scan of the P <= 1024 bucket counts: four buckets per thread
  uint32_t c4[4], s4 = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t b = 4 * threadIdx.x + k;
    c4[k] = b < P ? hist[b] : 0u;
    s4 += c4[k];
  }
  uint32_t inc = s4;
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(inc, d, 64);
    if (l >= d) inc += y;
  }
  if (l == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t off = inc - s4;
  for (int k = 0; k < w; ++k) off += wsum[k];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t b = 4 * threadIdx.x + k;
    if (b < P) hist[b] = off;
    off += c4[k];
  }
  if (threadIdx.x == CT - 1) hist[P] = off;  // candidates of the chunk
  __syncthreads();
  const int64_t cq = hc.cq;
  for (uint32_t b = threadIdx.x; b <= P; b += CT) H.cbo[cq + b] = (int32_t)hist[b];
#pragma unroll
  for (int j = 0; j < CI; ++j) {
    if (key[j] == MGL_EMPTY) continue;
    const int64_t dst = pos0 + hist[slot[j] >> 16] + (slot[j] & 0xFFFFu);
    H.tmp_c[dst] = key[j];
    ((AT*)H.tmp_a)[dst] = r[j].a;
    if (M::HAS_W)
      ((ST*)H.tmp_w)[dst] = wload<M>(A.w, e0 + (int64_t)j * CT + threadIdx.x);
  }
  eici = bsum_t<CT>(eici, red);
  selfw = bsum_t<CT>(selfw, red);
  if (threadIdx.x == 0) {
    if (eici != (KT)0) hatom_add(&((KT*)H.hub_eici)[i], eici);
    if (selfw != (KT)0) hatom_add(&((KT*)H.hub_self)[i], selfw);
  }
}

// ---- scatter ----------------------------------------------------------------
template <class M, class AT>
__global__ __launch_bounds__(HUB_T) void k_hub_scatter(MoveArgs A, HubPlan H) {
  typedef typename M::S ST;
  __shared__ unsigned int cnt[HUB_MAXP];
  __shared__ int64_t base[HUB_MAXP];
  const HubChunk hc = H.chunks[blockIdx.x];
  const int64_t i = hc.hub, pos0 = hc.pos0;
  if (!A.gs[hc.g].active) return;
  const int64_t len = hc.e1 - hc.e0;
  const uint32_t P = (uint32_t)hc.P;
  for (uint32_t b = threadIdx.x; b < P; b += HUB_T) cnt[b] = 0;
  __syncthreads();
  vid key[HUB_ITEMS];
  unsigned int rk[HUB_ITEMS];
  uint32_t bk[HUB_ITEMS];
#pragma unroll
  for (int j = 0; j < HUB_ITEMS; ++j) {
    const int64_t r = (int64_t)j * HUB_T + threadIdx.x;
    key[j] = (r < len) ? H.tmp_c[pos0 + r] : MGL_EMPTY;
    if (key[j] != MGL_EMPTY) {
      bk[j] = hub_bucket(key[j], P);
      rk[j] = atomicAdd(&cnt[bk[j]], 1u);
    }
  }
  __syncthreads();
  const int64_t bb = H.hub_bk[i];
  for (uint32_t b = threadIdx.x; b < P; b += HUB_T)
    if (cnt[b]) base[b] = (int64_t)atomicAdd((unsigned long long*)&H.bcur[bb + b], (unsigned long long)cnt[b]);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < HUB_ITEMS; ++j) {
    if (key[j] == MGL_EMPTY) continue;
    const int64_t dst = base[bk[j]] + rk[j];
    const int64_t src = pos0 + (int64_t)j * HUB_T + threadIdx.x;
    H.bkt_c[dst] = key[j];
    ((AT*)H.bkt_a)[dst] = ((const AT*)H.tmp_a)[src];
    if (M::HAS_W) ((ST*)H.bkt_w)[dst] = ((const ST*)H.tmp_w)[src];
  }
}

// ---- aggregate ----------------------------------------------------------------
// LDS / HBM table layout: keys, values, totals in 16-byte aligned parts
__host__ __device__ inline size_t hub_tab_bytes(uint32_t H, size_t sv, size_t sa) {
  return ((4 * (size_t)H + 15) / 16) * 16 + ((sv * H + 15) / 16) * 16 + sa * H;
}

// Prime table sizes for a bucket: the smallest listed prime at or above 3/2
// of the bucket's entries (load <= 2/3), capped by the largest LDS table.
__device__ __forceinline__ uint32_t hub_table_size(int64_t cnt, uint32_t Hmax) {
  const uint32_t primes[] = {509, 1021, 1531, 2039, 3067, 4093, 6143, 8191, 10223, 12281, 16381,
                             20479};
  const uint64_t want = (uint64_t)cnt + (uint64_t)cnt / 2;
#pragma unroll
  for (int i = 0; i < 12; ++i)
    if (primes[i] >= want && primes[i] <= Hmax) return primes[i];
  return Hmax;
}

// The largest listed prime at or below `bound` (bound >= 509).
__device__ __forceinline__ uint32_t hub_prime_below(uint32_t bound) {
  const uint32_t primes[] = {509, 1021, 1531, 2039, 3067, 4093, 6143, 8191, 10223, 12281, 16381,
                             20479};
  uint32_t r = 509;
#pragma unroll
  for (int i = 0; i < 12; ++i)
    if (primes[i] <= bound) r = primes[i];
  return r;
}

// Packed (key, 4-byte sum) insert into a prime table of H slots: the slot if
// the key was created, -1 if it existed, -2 if every slot holds another key.
template <class ST>
__device__ __forceinline__ int hub_ins_packed(uint64_t* T, vid key, ST w, uint32_t H,
                                              ProbeCount& pc) {
  uint32_t h = any_slot(key, H);
  const uint32_t st = probe_step(key, H, false, 0);
  uint32_t wb;
  __builtin_memcpy(&wb, &w, 4);
  const uint64_t empty = (uint64_t)(uint32_t)MGL_EMPTY << 32;
  const uint64_t want = ((uint64_t)(uint32_t)key << 32) | wb;
  for (uint32_t i = 0; i < H; ++i) {
    const uint64_t old = atomicCAS((unsigned long long*)&T[h], (unsigned long long)empty,
                                   (unsigned long long)want);
    if (old == empty) { pc.add(i + 1); return (int)h; }
    if ((uint32_t)(old >> 32) == (uint32_t)key) {
      atom_add((ST*)&T[h], w);
      pc.add(i + 1);
      return -1;
    }
    h = probe_next(h, st, H);
  }
  pc.add(H);
  return -2;
}

// Where the candidates of one bucket are. Contiguous: after the scatter,
// positions s0 .. s0+cnt of the bucketed arrays. Segmented: one segment in the
// region of every chunk of the hub (LDS prefix of the segment lengths).
struct BucketContig {
  const vid* c;
  const void* a;
  const void* w;
  int64_t s0;
  __device__ __forceinline__ int64_t at(int64_t x) const { return s0 + x; }
};
struct BucketSegs {
  const vid* c;
  const void* a;
  const void* w;
  int64_t base;          // tmp position of the hub's first chunk
  const int32_t* pre;    // LDS [nch+1]: segment prefix
  const int32_t* src;    // LDS [nch]: segment start relative to base
  int nch;
  __device__ __forceinline__ int64_t at(int64_t x) const {
    int lo = 0, hi = nch;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if ((int64_t)pre[mid] <= x) lo = mid;
      else hi = mid;
    }
    return base + (int64_t)src[lo] + (x - (int64_t)pre[lo]);
  }
};

// Aggregation of one bucket of hub i by a workgroup of BS threads. Buckets
// of at most IT entries per thread (and at most 7/8 of the largest LDS table)
// use an owner-based LDS table as the move kernels do: the creator of a key
// evaluates it with the total it staged, so the table is neither scanned nor
// sized beyond the bucket. Larger buckets (many entries of few communities)
// first try an LDS table that is scanned at the end (4-byte sums); if their
// distinct communities overfill it, or the sums take 8 bytes, an HBM table of
// 2*cnt slots from the pass's pool (H.galloc) is used.
// The gain inputs of a bucket's hub, loaded by the caller at its start.
struct HubIn {
  double mass;     // 2m of the hub's graph
  int64_t mass_i;  // 2m in integer units (exact modes)
  int64_t ki, acik, eici;  // bits of the mode's total type
  vid ci;
};

template <class M, class AT, int BS, int IT, bool CNT, class SRC>
__device__ __forceinline__ void hub_aggregate_body(const MoveArgs& A, const HubPlan& H,
                                                   uint32_t Hmax, int64_t b, const HubIn& in,
                                                   const SRC& src, int64_t cnt,
                                                   unsigned char* smem) {
  typedef typename M::S ST;
  typedef typename M::K KT;
  constexpr bool PACK = sizeof(ST) == 4;
  typedef OwnTab<ST, PACK> TB;
  __shared__ double rg[BS / 64];
  __shared__ vid rd[BS / 64];
  __shared__ int rc[BS / 64];
  __shared__ int ovf;
  __shared__ unsigned int ncre;
  __shared__ unsigned long long gbase;
  const vid ci = in.ci;
  const bool small = A.tie_small ? true : (A.up != 0);
  double best = 0.0;
  vid dst = ci;
  int occ = 0;
  ProbeCount pc;
  if (cnt > 0) {
    const GainIn G = gain_in<M, AT>(in.mass, in.mass_i, gin_bits<KT>(in.ki),
                                    gin_bits<KT>(in.acik), gin_bits<KT>(in.eici));
    const AT* ba = (const AT*)src.a;
    if (cnt <= (int64_t)IT * BS && (uint64_t)cnt * 8 <= (uint64_t)Hmax * 7) {
      const uint32_t Ht = hub_table_size(cnt, Hmax);
      TB tb;
      tb.init(smem, Ht);
      for (uint32_t s = threadIdx.x; s < Ht; s += BS) tb.clear(s);
      int own[IT];
      AT oa[IT];
      vid key[IT];
      ST wv[IT];
#pragma unroll
      for (int j = 0; j < IT; ++j) {
        const int64_t x = threadIdx.x + (int64_t)j * BS;
        key[j] = MGL_EMPTY;
        if (x < cnt) {
          const int64_t q = src.at(x);
          key[j] = src.c[q];
          oa[j] = ba[q];
          wv[j] = M::HAS_W ? ((const ST*)src.w)[q] : (ST)1;
        }
      }
      lds_barrier();  // the table is cleared; the candidate loads stay in flight
#pragma unroll
      for (int j = 0; j < IT; ++j)
        own[j] = key[j] != MGL_EMPTY ? tb.ins(key[j], wv[j], Ht, false, 0, pc) : -1;
      __syncthreads();
      own_gains<M, AT, IT, TB>(tb, own, oa, G, small, best, dst, occ);
    } else {
      bool done = false;
      if constexpr (PACK) {
        // LDS table of Hb packed slots followed by the creators' totals
        const uint32_t Hb = hub_prime_below((uint32_t)(8ull * Hmax / (8 + sizeof(AT))));
        uint64_t* T = (uint64_t*)smem;
        AT* Avs = (AT*)(smem + 8 * (size_t)Hb);
        const uint64_t empty = (uint64_t)(uint32_t)MGL_EMPTY << 32;
        for (uint32_t s = threadIdx.x; s < Hb; s += BS) T[s] = empty;
        if (threadIdx.x == 0) {
          ovf = 0;
          ncre = 0;
        }
        __syncthreads();
        const unsigned lim = Hb / 8 * 7;
        for (int64_t x = threadIdx.x; x < cnt; x += BS) {
          if (__hip_atomic_load(&ovf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
          const int64_t q = src.at(x);
          const ST w1 = M::HAS_W ? ((const ST*)src.w)[q] : (ST)1;
          const int h = hub_ins_packed<ST>(T, src.c[q], w1, Hb, pc);
          if (h >= 0) {
            Avs[h] = ba[q];
            if (atomicAdd(&ncre, 1u) >= lim)
              __hip_atomic_store(&ovf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          } else if (h == -2) {
            __hip_atomic_store(&ovf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        }
        __syncthreads();
        if (!ovf) {
          Argmax<M, AT> am(ci);
          for (uint32_t s = threadIdx.x; s < Hb; s += BS) {
            const uint64_t t = T[s];
            const vid cj = (vid)(uint32_t)(t >> 32);
            if (cj == MGL_EMPTY) continue;
            const uint32_t lo = (uint32_t)t;
            ST ev;
            __builtin_memcpy(&ev, &lo, 4);
            ++occ;
            am.offer(G, ev, Avs[s], cj, small);
          }
          best = am.best;
          dst = am.dst;
          done = true;
        }
        __syncthreads();  // the LDS table is not reused below, but ovf is read above
      }
      if (!done) {
        const uint32_t Ht = (uint32_t)(2 * cnt);
        if (threadIdx.x == 0) gbase = atomicAdd(H.galloc, (unsigned long long)Ht);
        __syncthreads();
        vid* K = H.gk + gbase;
        ST* V = (ST*)H.gv + gbase;
        AT* Av = (AT*)H.ga + gbase;
        for (uint32_t s = threadIdx.x; s < Ht; s += BS) { K[s] = MGL_EMPTY; V[s] = (ST)0; }
        __threadfence();
        __syncthreads();
        for (int64_t x = threadIdx.x; x < cnt; x += BS) {
          const int64_t q = src.at(x);
          const ST w1 = M::HAS_W ? ((const ST*)src.w)[q] : (ST)1;
          tab_insert<ST, AT, true>(K, V, Av, src.c[q], w1, ba[q], Ht);
        }
        __threadfence();
        __syncthreads();
        Argmax<M, AT> am(ci);
        for (uint32_t s = threadIdx.x; s < Ht; s += BS) {
          const vid cj = __hip_atomic_load(&K[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (cj == MGL_EMPTY) continue;
          const ST ev = __hip_atomic_load(&V[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const AT ak = __hip_atomic_load(&Av[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ++occ;
          am.offer(G, ev, ak, cj, small);
        }
        best = am.best;
        dst = am.dst;
      }
    }
  }
  // workgroup argmax (cnt is uniform, every thread reaches this point)
  if constexpr (CNT) flush_probes(A.probes, pc);
  flag_overflow(A.err, pc);
  group_argmax<64>(best, dst, small);
  occ = group_sum<64>(occ);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) { rg[w] = best; rd[w] = dst; rc[w] = occ; }
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int k2 = 0; k2 < BS / 64; ++k2) {
      if (better(rg[k2], rd[k2], best, dst, small)) { best = rg[k2]; dst = rd[k2]; }
      tot += rc[k2];
    }
    H.bgain[b] = best;
    H.bdst[b] = dst;
    if (CNT && A.stats && tot) atomicAdd(A.stats, (unsigned long long)tot);
  }
}

// One workgroup per bucket, after the scatter (contiguous buckets).
// Gain inputs of hub i of graph g.
__device__ __forceinline__ HubIn hub_in(const MoveArgs& A, const HubPlan& H, int64_t i, int g) {
  const HubGain hg = H.gin[i];
  HubIn in;
  in.mass = A.gs[g].mass;
  in.mass_i = A.gs[g].mass_i;
  in.ki = hg.ki;
  in.acik = hg.acik;
  in.eici = ((const int64_t*)H.hub_eici)[i];  // 8-byte sums: the bits of KT
  in.ci = hg.ci;
  return in;
}

template <class M, class AT, int BS, int IT, bool CNT>
__global__ __launch_bounds__(BS, BS >= 512 ? 8 : 1) void k_hub_aggregate(MoveArgs A, HubPlan H, uint32_t Hmax) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int64_t b = blockIdx.x;
  const int64_t i = H.bk_hub[b];
  const vid v = H.hub_v[i];
  const int g = A.gid ? A.gid[v] : 0;
  if (!A.gs[g].active) return;
  const int64_t s0 = H.boff[b], s1 = H.boff[b + 1];
  BucketContig src{H.bkt_c, H.bkt_a, H.bkt_w, s0};
  const HubIn in = hub_in(A, H, i, g);
  hub_aggregate_body<M, AT, BS, IT, CNT>(A, H, Hmax, b, in, src, s1 - s0, smem);
}

// Block-wide exclusive scan of one value per thread (BS threads).
template <int BS>
__device__ __forceinline__ int32_t exscan_block(int32_t x, int32_t* wsum /* [BS / 64] */) {
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  int32_t inc = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int32_t y = __shfl_up(inc, d, 64);
    if (l >= d) inc += y;
  }
  if (l == 63) wsum[w] = inc;
  __syncthreads();
  int32_t off = 0;
  for (int k = 0; k < w; ++k) off += wsum[k];
  __syncthreads();
  return off + inc - x;
}

// One workgroup per bucket in the segmented form: the bucket's segment in each
// chunk of its hub is located from the chunks' bucket offsets.
template <class M, class AT, int BS, int IT, bool CNT>
__global__ __launch_bounds__(BS, BS >= 512 ? 8 : 1) void k_hub_aggregate_seg(MoveArgs A, HubPlan H, uint32_t Hmax) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int32_t pre[HUB_SEG_MAX + 1];
  __shared__ int32_t srcs[HUB_SEG_MAX];
  __shared__ int32_t wsum[BS / 64];
  const int64_t b = blockIdx.x;
  const HubBucket hb = H.bmeta[b];  // one (scalar) load: the bucket's fixed layout
  const int64_t i = hb.hub;
  // everything below depends on the bucket record only: the segment offsets,
  // the hub's gain inputs and the graph's state are requested together
  const HubIn in = hub_in(A, H, i, hb.g);
  const int act = A.gs[hb.g].active;
  const int64_t lb = hb.lb;
  const int64_t P = hb.P;
  const int nch = hb.nch;
  const int64_t cb0 = hb.cb0;
  const int64_t base = hb.base;
  // segment lengths, CPT consecutive chunks per thread (nch <= HUB_SEG_MAX);
  // chunk c of the hub starts at base + c * HUB_CHUNK in tmp
  constexpr int CPT = (HUB_SEG_MAX + BS - 1) / BS;
  int32_t len[CPT];
  int32_t tot = 0;
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int c = CPT * threadIdx.x + k;
    len[k] = 0;
    if (c < nch) {
      const int64_t q = cb0 + (int64_t)c * (P + 1) + lb;
      const int32_t s = H.cbo[q];
      len[k] = H.cbo[q + 1] - s;
      srcs[c] = c * HUB_CHUNK + s;
    }
    tot += len[k];
  }
  if (!act) return;  // uniform: the bucket's graph has finished
  int32_t ex = exscan_block<BS>(tot, wsum);
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int c = CPT * threadIdx.x + k;
    if (c < nch) pre[c] = ex;
    ex += len[k];
  }
  if (threadIdx.x == BS - 1) pre[nch] = ex;
  __syncthreads();
  const int64_t cnt = pre[nch];
  BucketSegs src{H.tmp_c, H.tmp_a, H.tmp_w, base, pre, srcs, nch};
  hub_aggregate_body<M, AT, BS, IT, CNT>(A, H, Hmax, b, in, src, cnt, smem);
}

// ---- finalize -----------------------------------------------------------------
// One wavefront per hub: its lanes take the hub's buckets in turn, then a
// wavefront argmax (a total order, so the result does not depend on how the
// buckets are split over the lanes) and the bi-probing clamp.
template <class M, class AT>
__global__ __launch_bounds__(256) void k_hub_final(MoveArgs A, HubPlan H) {
  typedef typename M::K KT;
  const int64_t i = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;  // uniform per wavefront
  const int lane = threadIdx.x & 63;
  if (i >= H.m) return;
  const vid v = H.hub_v[i];
  const int g = A.gid ? A.gid[v] : 0;
  if (!A.gs[g].active) return;
  const vid ci = A.c[v];
  const bool small = A.tie_small ? true : (A.up != 0);
  double best = 0.0;
  vid dst = ci;
  for (int64_t b = H.hub_bk[i] + lane; b < H.hub_bk[i + 1]; b += 64)
    if (better(H.bgain[b], H.bdst[b], best, dst, small)) { best = H.bgain[b]; dst = H.bdst[b]; }
  group_argmax<64>(best, dst, small);
  if (lane != 0) return;
  A.c_new[v] = (A.version == 1) ? dst : (A.up ? (dst < ci ? dst : ci) : (dst > ci ? dst : ci));
  if (A.vin) ((KT*)A.vin)[v] = ((const KT*)H.hub_eici)[i] + ((const KT*)H.hub_self)[i];
}

__global__ void k_hub_copy(const int64_t* s, int64_t* d, int64_t n) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x)
    d[j] = s[j];
}

template <class M, class AT>
static int hub_t(const MoveArgs& A, const HubPlan& H, uint32_t Hlds, hipStream_t st) {
  typedef typename M::S ST;
  if (H.m <= 0) return 0;
  const bool cnt = A.stats || A.probes;
  const size_t lds = (sizeof(ST) == 4) ? 8 * (size_t)Hlds
                                       : ((4 * (size_t)Hlds + 15) / 16) * 16 + sizeof(ST) * (size_t)Hlds;
  k_hub_head<M><<<(unsigned)((H.m + 255) / 256), 256, 0, st>>>(A, H);
  const unsigned nbk = (unsigned)H.nbuckets;
#define AGG_(K_, BS_)                                                                    \
  if (H.agg_bs == BS_) {                                                                 \
    if (cnt) K_<M, AT, BS_, HUB_AGG_CAP / BS_, true><<<nbk, BS_, lds, st>>>(A, H, Hlds);  \
    else K_<M, AT, BS_, HUB_AGG_CAP / BS_, false><<<nbk, BS_, lds, st>>>(A, H, Hlds);     \
  }
  if (H.segmented) {
    if (H.count_bs == 512) k_hub_count_seg<M, AT, 512><<<(unsigned)H.nchunks, 512, 0, st>>>(A, H);
    else if (H.count_bs == 1024)
      k_hub_count_seg<M, AT, 1024><<<(unsigned)H.nchunks, 1024, 0, st>>>(A, H);
    else k_hub_count_seg<M, AT, 256><<<(unsigned)H.nchunks, 256, 0, st>>>(A, H);
    AGG_(k_hub_aggregate_seg, 256) AGG_(k_hub_aggregate_seg, 512) AGG_(k_hub_aggregate_seg, 1024)
  } else {
    launch_fill_bytes0(H.bcount, 8 * (size_t)H.nbuckets, st);
    k_hub_count<M, AT><<<(unsigned)H.nchunks, HUB_T, 0, st>>>(A, H);
    scan_exclusive_i64(H.bcount, H.boff, H.nbuckets, H.scan_tmp, st);
    {
      int64_t gr = (H.nbuckets + 255) / 256;
      if (gr > 4096) gr = 4096;
      k_hub_copy<<<(unsigned)gr, 256, 0, st>>>(H.boff, H.bcur, H.nbuckets);
    }
    k_hub_scatter<M, AT><<<(unsigned)H.nchunks, HUB_T, 0, st>>>(A, H);
    AGG_(k_hub_aggregate, 256) AGG_(k_hub_aggregate, 512) AGG_(k_hub_aggregate, 1024)
  }
#undef AGG_
  k_hub_final<M, AT><<<(unsigned)((H.m * 64 + 255) / 256), 256, 0, st>>>(A, H);
  return 0;
}

int launch_hub(int mode, int rec32, const MoveArgs& A, const HubPlan& H, hipStream_t st) {
  switch (mode) {
    case MODE_UNIT:
      return rec32 ? hub_t<ModeUnit, int32_t>(A, H, H.lds_slots, st)
                   : hub_t<ModeUnit, int64_t>(A, H, H.lds_slots, st);
    case MODE_INT: return hub_t<ModeInt, int64_t>(A, H, H.lds_slots, st);
    case MODE_I32:
      return rec32 ? hub_t<ModeI32, int32_t>(A, H, H.lds_slots, st)
                   : hub_t<ModeI32, int64_t>(A, H, H.lds_slots, st);
    case MODE_F64: return hub_t<ModeF64, double>(A, H, H.lds_slots, st);
    case MODE_F32: return hub_t<ModeF32, double>(A, H, H.lds_slots, st);
  }
  return -1;
}

void hub_allow_lds(int bytes) {
#define B_(K_, M_, AT_, BS_, C_)                                                       \
  (void)hipFuncSetAttribute((const void*)K_<M_, AT_, BS_, HUB_AGG_CAP / BS_, C_>,      \
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
#define S_(M_, AT_, BS_)                                                               \
  B_(k_hub_aggregate, M_, AT_, BS_, false) B_(k_hub_aggregate, M_, AT_, BS_, true)     \
  B_(k_hub_aggregate_seg, M_, AT_, BS_, false) B_(k_hub_aggregate_seg, M_, AT_, BS_, true)
#define A_(M_, AT_) S_(M_, AT_, 256) S_(M_, AT_, 512) S_(M_, AT_, 1024)
  A_(ModeUnit, int32_t) A_(ModeUnit, int64_t) A_(ModeInt, int64_t) A_(ModeF64, double)
  A_(ModeF32, double) A_(ModeI32, int32_t) A_(ModeI32, int64_t)
#undef A_
#undef S_
#undef B_
}

int hub_chunk_edges() { return HUB_CHUNK; }
int hub_max_buckets() { return HUB_MAXP; }
int hub_seg_max() { return HUB_SEG_MAX; }
int hub_agg_cap() { return HUB_AGG_CAP; }
