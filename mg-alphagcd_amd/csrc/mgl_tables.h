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
__device__ __forceinline__ double gain(double e, double eici, double ki, double aci,
                                       double acj, double Mh) {
  double wij = e;
  wij -= (eici - ki * (aci - acj) / (2. * Mh));
  wij /= Mh;
  return wij;
}

// Per-thread statistics of LDS hash inserts: inserts, compare-and-swap probes
// and inserts that needed more than one probe. Summed per wavefront and added
// to A.probes[0..2] when statistics are requested.
struct ProbeCount {
  uint32_t ins = 0, probes = 0, multi = 0;
  __device__ __forceinline__ void add(uint32_t np) {
    ++ins;
    probes += np;
    multi += np > 1u ? 1u : 0u;
  }
};
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
  pc.add(H);
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
  pc.add(H);
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

template <class M, class AT, int IT, class TB>
__device__ __forceinline__ void own_gains(const TB& tb, const int (&own)[IT], const AT (&oa)[IT],
                                          double deici, double dki, double daci, double Mh,
                                          bool small, double& best, vid& dst, int& cnt) {
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    if (own[j] < 0) continue;
    ++cnt;
    const vid cj = tb.key(own[j]);
    const double wij = gain((double)tb.val(own[j]), deici, dki, daci, (double)oa[j], Mh);
    if (better(wij, cj, best, dst, small)) { best = wij; dst = cj; }
  }
}

