// LDS hash tables shared by the move kernels and the hub aggregate: slot
// atomics, the FP64 gain, owner-based tables and their evaluation.
#pragma once
#pragma clang fp contract(off)

#include "mgl_kernels.h"

// atomics on any slot type
__device__ __forceinline__ void atom_add(int32_t* p, int32_t x) { atomicAdd(p, x); }
__device__ __forceinline__ void atom_add(int64_t* p, int64_t x) {
  atomicAdd((unsigned long long*)p, (unsigned long long)x);
}
__device__ __forceinline__ void atom_add(float* p, float x) { atomicAdd(p, x); }
__device__ __forceinline__ void atom_add(double* p, double x) { atomicAdd(p, x); }

// Per-vertex gain of moving into cj (louvain.cu:540-543, same operation order):
//   wij = e_{v,cj}; wij -= (eici - ki*(aci - acj)/(2*M)); wij /= M;  M = mass/2
__host__ __device__ __forceinline__ double gain(double e, double eici, double ki, double aci,
                                       double acj, double Mh) {
  double wij = e;
  wij -= (eici - ki * (aci - acj) / (2. * Mh));
  wij /= Mh;
  return wij;
}

// Per-thread statistics of LDS hash inserts: inserts, compare-and-swap probes
// and inserts that needed more than one probe. Summed per wavefront and added
// to A.probes[0..2] when statistics are requested.
// `fail` records an insert that found every slot taken by other keys (a table
// smaller than the row's distinct communities); flag_overflow reports it.
struct ProbeCount {
  uint32_t ins = 0, probes = 0, multi = 0, fail = 0;
  __device__ __forceinline__ void add(uint32_t np) {
    ++ins;
    probes += np;
    multi += np > 1u ? 1u : 0u;
  }
  __device__ __forceinline__ void over(uint32_t np) {
    add(np);
    fail = 1;
  }
};
// A dropped entry makes the row's result wrong: the launch sets *err and the
// host turns it into MGL_DEVICE_ERROR. The built-in layouts and every tuning
// file that passes validation size the tables for their bins' rows, so this
// does not happen unless that validation is switched off.
__device__ __forceinline__ void flag_overflow(int* err, const ProbeCount& pc) {
  if (pc.fail && err) atomicOr(err, 1);
}
// Every lane of the wavefront must call this (shuffle reduction).
__device__ __forceinline__ void flush_probes(unsigned long long* out, const ProbeCount& pc) {
  if (!out) return;
  const uint32_t a = group_sum<64>(pc.ins), b = group_sum<64>(pc.probes),
                 c = group_sum<64>(pc.multi);
  if ((threadIdx.x & 63) == 0 && a) {
    atomicAdd(out, (unsigned long long)a);
    atomicAdd(out + 1, (unsigned long long)b);
    atomicAdd(out + 2, (unsigned long long)c);
  }
}

// Owner-based variants (rows of at most IT entries per lane): each lane keeps
// in registers the entries it gathered and the slots whose key it created.
// The table holds key and partial sum only; after the barrier every creator
// evaluates its own keys with the total it gathered itself, so neither a
// table scan nor a second gather is needed.
// Tables whose sums take 4 bytes keep (sum, key) in one 64-bit word per slot
// (low half = sum, high half = key): a new key is one 64-bit CAS that also
// stores its first weight; a known key is one 32-bit add on the low half.
template <class ST>
__device__ __forceinline__ int own_packed(uint64_t* T, vid key, ST w, uint32_t H, bool pow2,
                                          int lg, ProbeCount& pc) {
  uint32_t h = pow2 ? slot_hash(key, lg) : any_slot(key, H);
  const uint32_t st = probe_step(key, H, pow2, lg);
  uint32_t wb;
  __builtin_memcpy(&wb, &w, 4);
  const uint64_t empty = (uint64_t)(uint32_t)MGL_EMPTY << 32;
  const uint64_t want = ((uint64_t)(uint32_t)key << 32) | wb;
  for (uint32_t i = 0; i < H; ++i) {
    const uint64_t old = atomicCAS((unsigned long long*)&T[h], (unsigned long long)empty,
                                   (unsigned long long)want);
    if (old == empty) { pc.add(i + 1); return (int)h; }
    if ((uint32_t)(old >> 32) == (uint32_t)key) {
      atom_add((ST*)&T[h], w);  // the low half
      pc.add(i + 1);
      return -1;
    }
    h = probe_next(h, st, H);
  }
  pc.over(H);
  return -1;
}

// Separate key / sum arrays (8-byte sums): the CAS on the key decides the
// creator; the sum is added by everyone.
template <class ST>
__device__ __forceinline__ int own_sep(vid* K, ST* V, vid key, ST w, uint32_t H, bool pow2,
                                       int lg, ProbeCount& pc) {
  uint32_t h = pow2 ? slot_hash(key, lg) : any_slot(key, H);
  const uint32_t st = probe_step(key, H, pow2, lg);
  for (uint32_t i = 0; i < H; ++i) {
    const vid cur = atomicCAS(&K[h], MGL_EMPTY, key);
    if (cur == MGL_EMPTY || cur == key) {
      atom_add(&V[h], w);
      pc.add(i + 1);
      return cur == MGL_EMPTY ? (int)h : -1;
    }
    h = probe_next(h, st, H);
  }
  pc.over(H);
  return -1;
}

// Owner-table views: packed (key, 4-byte sum) words, or separate key / sum arrays.
template <class ST, bool PACK>
struct OwnTab;
template <class ST>
struct OwnTab<ST, true> {
  uint64_t* T;
  static constexpr size_t slot = 8;
  __device__ void init(unsigned char* base, uint32_t) { T = (uint64_t*)base; }
  __device__ void clear(uint32_t i) { T[i] = (uint64_t)(uint32_t)MGL_EMPTY << 32; }
  __device__ int ins(vid key, ST w, uint32_t H, bool pow2, int lg, ProbeCount& pc) {
    return own_packed<ST>(T, key, w, H, pow2, lg, pc);
  }
  __device__ vid key(int i) const { return (vid)(uint32_t)(T[i] >> 32); }
  __device__ ST val(int i) const {
    const uint32_t lo = (uint32_t)T[i];
    ST x;
    __builtin_memcpy(&x, &lo, 4);
    return x;
  }
};
template <class ST>
struct OwnTab<ST, false> {
  vid* K;
  ST* V;
  static constexpr size_t slot = 4 + sizeof(ST);
  __device__ void init(unsigned char* base, uint32_t H) {
    K = (vid*)base;
    V = (ST*)(base + ((4 * (size_t)H + 15) / 16) * 16);
  }
  __device__ void clear(uint32_t i) { K[i] = MGL_EMPTY; V[i] = (ST)0; }
  __device__ int ins(vid key, ST w, uint32_t H, bool pow2, int lg, ProbeCount& pc) {
    return own_sep<ST>(K, V, key, w, H, pow2, lg, pc);
  }
  __device__ vid key(int i) const { return K[i]; }
  __device__ ST val(int i) const { return V[i]; }
};

// ---------------------------------------------------------------------------
// Argmax over the candidate communities of one vertex (louvain.cu:526-556),
// starting from (gain 0.0, own community).
//
// Exact prefilter, for the exact modes whose slot sums and community totals
// are 32-bit (then 2m < 2^31): with D = 2m, m = D/2,
//   X(c) = D*(e_c - eici) + ki*(aci - a_c)
// is the exact value of gain(c) * D * m (an integer below 2^63). The FP64
// expression of gain() rounds five times on terms bounded by ki (e_c, eici
// <= ki and |aci - a_c| <= D), so it differs from X/(D*m) by less than
// 8u*ki/m with u = 2^-53, and two exact values more than 2u apart, relative,
// round to distinct doubles. With delta = D*ki/2^48 (= 32u*D*ki) therefore
//   X(c) < X(b) - delta  =>  gain(c) < gain(b)  strictly,
// where b is the current best (X = 0 and gain 0.0 for staying). Such a
// candidate cannot be the argmax and its gain() is not evaluated. A candidate
// with the same (e, a) as the current best has the same gain; only the ids
// are compared. Every other candidate is decided by gain() as before, so the
// result is bit-for-bit that of evaluating every candidate; the prefilter
// only saves the FP64 work (two correctly rounded divisions per candidate).
struct GainIn {
  double deici, dki, daci, Mh;  // FP64 inputs of gain()
  int64_t D, ki, aci, eici;     // exact inputs of the prefilter
  int64_t delta;
};
template <class M, class AT>
__host__ __device__ constexpr bool gain_prefilter() {
  return M::EXACT && sizeof(typename M::S) == 4 && sizeof(AT) == 4;
}

// Inputs of a vertex's gains: 2m as double and (exact modes) as an integer,
// the vertex's total ki, its own community's total without it (aci) and its
// weight to its own community without self-loops (eici).
template <class M, class AT, class KT, class ST>
__host__ __device__ __forceinline__ GainIn gain_in(double mass, int64_t mass_i, KT ki, KT aci_k, ST eici) {
  GainIn g;
  g.Mh = mass / 2.0;
  g.deici = (double)eici;
  g.dki = (double)ki;
  g.daci = (double)aci_k;
  g.D = g.ki = g.aci = g.eici = g.delta = 0;
  if constexpr (gain_prefilter<M, AT>()) {
    g.D = mass_i;
    g.ki = (int64_t)ki;
    g.aci = (int64_t)aci_k;
    g.eici = (int64_t)eici;
    g.delta = ((mass_i * (int64_t)ki) >> 48) + 1;  // mass_i, ki < 2^31
  }
  return g;
}

template <class M, class AT>
struct Argmax {
  typedef typename M::S ST;
  static constexpr bool PRE = gain_prefilter<M, AT>();
  double best = 0.0;
  vid dst;
  int64_t xb = 0;  // X of the current best
  ST eb = 0;
  AT ab = 0;
  bool have = false;  // the current best is a candidate (not the own community)
  __host__ __device__ explicit Argmax(vid ci) : dst(ci) {}
  __host__ __device__ __forceinline__ void offer(const GainIn& G, ST e, AT a, vid cj,
                                                      bool small) {
    if constexpr (PRE) {
      const int64_t x = (int64_t)(int32_t)G.D * (int64_t)((int32_t)e - (int32_t)G.eici) +
                        (int64_t)(int32_t)G.ki * (int64_t)((int32_t)G.aci - (int32_t)a);
      if (x < xb - G.delta) return;
      if (have && e == eb && a == ab) {
        if (small ? cj < dst : cj > dst) dst = cj;
        return;
      }
      const double g = gain((double)e, G.deici, G.dki, G.daci, (double)a, G.Mh);
      if (better(g, cj, best, dst, small)) {
        best = g;
        dst = cj;
        xb = x;
        eb = e;
        ab = a;
        have = true;
      }
    } else {
      const double g = gain((double)e, G.deici, G.dki, G.daci, (double)a, G.Mh);
      if (better(g, cj, best, dst, small)) { best = g; dst = cj; }
    }
  }
};

template <class M, class AT, int IT, class TB>
__device__ __forceinline__ void own_gains(const TB& tb, const int (&own)[IT], const AT (&oa)[IT],
                                          const GainIn& G, bool small, double& best, vid& dst,
                                          int& cnt) {
  Argmax<M, AT> am(dst);
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    if (own[j] < 0) continue;
    ++cnt;
    am.offer(G, tb.val(own[j]), oa[j], tb.key(own[j]), small);
  }
  best = am.best;
  dst = am.dst;
}
