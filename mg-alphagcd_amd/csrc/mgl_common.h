// Shared types and device helpers of the MI355X Louvain engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef int32_t vid;  // vertex / community id (reference vertex_t, common.h:85)
typedef int64_t eid;  // CSR position (reference edge_t = uint32, widened: >4.29e9 entries)
typedef unsigned __int128 u128;

#define MGL_EMPTY (-1)

// ---------------------------------------------------------------------------
// Arithmetic modes. Each fixes the storage type of edge weights (W), of hash
// slots / per-vertex partial sums (S) and of vertex / community totals (K).
//   Unit : level-0 graph without weights; counts in int32 slots, totals int64.
//   Int  : integer or dyadic fixed-point weights (scaled by 2^S) in int64.
//   F64  : double weights, slots and totals (the reference's weight_t).
//   F32  : float weights and slots; double totals, 2m and Q.
//   I32  : integer weights below 2^31 (coarse levels of a unit graph whose 2m
//          is below 2^31): int32 weights and slots, int64 totals.
enum ModeId { MODE_UNIT = 0, MODE_INT = 1, MODE_F64 = 2, MODE_F32 = 3, MODE_I32 = 4 };
__host__ __device__ inline bool mode_is_int(int m) {
  return m == MODE_UNIT || m == MODE_INT || m == MODE_I32;
}

struct ModeUnit {
  typedef int32_t W; typedef int32_t S; typedef int64_t K;
  static constexpr bool HAS_W = false; static constexpr bool EXACT = true;
};
struct ModeInt {
  typedef int64_t W; typedef int64_t S; typedef int64_t K;
  static constexpr bool HAS_W = true; static constexpr bool EXACT = true;
};
struct ModeI32 {
  typedef int32_t W; typedef int32_t S; typedef int64_t K;
  static constexpr bool HAS_W = true; static constexpr bool EXACT = true;
};
struct ModeF64 {
  typedef double W; typedef double S; typedef double K;
  static constexpr bool HAS_W = true; static constexpr bool EXACT = false;
};
struct ModeF32 {
  typedef float W; typedef float S; typedef double K;
  static constexpr bool HAS_W = true; static constexpr bool EXACT = false;
};

// Column and weight arrays are streamed once per pass: non-temporal loads keep
// them from evicting the randomly gathered per-vertex records from L2.
__device__ __forceinline__ vid colload(const vid* col, eid e) {
  return __builtin_nontemporal_load(col + e);
}

template <class M>
__device__ __forceinline__ typename M::S wload(const void* w, eid e) {
  if constexpr (M::HAS_W)
    return (typename M::S)__builtin_nontemporal_load((const typename M::W*)w + e);
  else return (typename M::S)1;
}

// Per-vertex record gathered once per CSR entry: the vertex's community and
// that community's total weight.
template <class AT>
struct __align__(sizeof(AT) >= 8 ? 16 : 8) Rec {
  vid c;
  AT a;
};

// Per-graph control block (device). One per graph of a batch; a plain run has one.
struct GState {
  double mass;        // 2m as double (reference "mass", host_graph.cpp:112-114)
  double q_cur;       // committed modularity of the level (reference cur_Q / new_Q)
  int64_t mass_i;     // 2m in integer units (exact modes)
  int32_t active;     // inner loop still running at this level
  int32_t iters;      // inner iterations done (reference loop_num)
  int32_t fix;        // set by decide: 0 none, 1 all buffers := c, 2 all buffers := p
  int32_t pad;
  // reduction accumulators of the current iteration
  int64_t sin_i;      // sum_in, exact
  double sin_d;       // sum_in, float modes
  uint64_t sa2_lo, sa2_hi;   // sum a^2 (current), exact 128-bit
  double sa2_d;
  uint64_t sa2p_lo, sa2p_hi; // sum a^2 of the proposal under evaluation (fused mode)
  double sa2p_d;
  int64_t moved;      // statistics
  int64_t distinct;
  double q0;          // Q of the level's singleton start (reference "Loop 0" row)
  int64_t tlast;      // device clock at the level start / the last decision
};

// ---------------------------------------------------------------------------
// u128 -> double, correctly rounded: the top 64 significant bits with the rest
// folded into a sticky bit, then one RNE conversion (exact scaling after).
__host__ __device__ inline double u128_to_double(uint64_t hi, uint64_t lo) {
  if (hi == 0) return (double)lo;
  int s = __builtin_clzll(hi);
  uint64_t top = s ? ((hi << s) | (lo >> (64 - s))) : hi;
  uint64_t rest = lo << s;
  top |= (rest != 0);
  double r = (double)top;
  // multiply by 2^(64-s) exactly
  int ex = 64 - s;
  while (ex > 0) { int st = ex > 60 ? 60 : ex; r *= (double)(1ull << st); ex -= st; }
  return r;
}

// Exact 128-bit atomic accumulation as two 64-bit atomics with carry.
__device__ __forceinline__ void atomic_add_u128(uint64_t* lo, uint64_t* hi, u128 x) {
  uint64_t xl = (uint64_t)x, xh = (uint64_t)(x >> 64);
  uint64_t old = atomicAdd((unsigned long long*)lo, (unsigned long long)xl);
  uint64_t carry = (old + xl < old) ? 1ull : 0ull;
  if (xh + carry) atomicAdd((unsigned long long*)hi, (unsigned long long)(xh + carry));
}

// Multiplicative (Fibonacci) hash into a table of 2^lg slots, lg in [1, 31].
__device__ __forceinline__ uint32_t slot_hash(vid key, int lg) {
  return ((uint32_t)key * 2654435761u) >> (32 - lg);
}

// Slot of a key in a table of H slots (any H): a mixed 32-bit hash scaled by
// a multiply-high, so no power-of-two size is needed.
__device__ __forceinline__ uint32_t any_slot(vid key, uint32_t H) {
  uint32_t x = (uint32_t)key * 0x85EBCA6Bu;
  x ^= x >> 15;
  x *= 0xC2B2AE35u;
  return (uint32_t)(((uint64_t)x * H) >> 32);
}

// Probe step of double hashing, from a second hash of the key. Tables of 2^lg
// slots take an odd step; tables of any other size H are prime-sized and take
// a step in [1, H-1]. Either way the step is coprime with the table size, so a
// probe sequence visits every slot once in H probes.
__device__ __forceinline__ uint32_t key_mix2(vid key) {
  uint32_t x = (uint32_t)key ^ 0x5bd1e995u;
  x *= 0x27d4eb2du;
  x ^= x >> 15;
  x *= 0x165667b1u;
  return x ^ (x >> 13);
}
__device__ __forceinline__ uint32_t probe_step(vid key, uint32_t H, bool pow2, int lg) {
  const uint32_t x = key_mix2(key);
  (void)lg;
  return pow2 ? ((x & (H - 1u)) | 1u) : 1u + (uint32_t)(((uint64_t)x * (H - 1u)) >> 32);
}
__device__ __forceinline__ uint32_t probe_next(uint32_t h, uint32_t step, uint32_t H) {
  h += step;
  return h >= H ? h - H : h;
}

// Total order used by the argmax (louvain.cu:544-554): higher gain first; on
// equal gain the smaller id on up passes and the larger id on down passes
// (tie_small forces "smaller" always).
__host__ __device__ __forceinline__ bool better(double g, vid d, double bg, vid bd,
                                                bool prefer_small) {
  return (g > bg) || (g == bg && (prefer_small ? d < bd : d > bd));
}

template <class T>
__device__ __forceinline__ T shfl_xor_t(T v, int m, int width) {
  return __shfl_xor(v, m, width);
}
__device__ __forceinline__ int64_t shfl_xor_t(int64_t v, int m, int width) {
  int lo = __shfl_xor((int)(uint32_t)v, m, width);
  int hi = __shfl_xor((int)(uint32_t)((uint64_t)v >> 32), m, width);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Sum over a group of S consecutive lanes (S power of two, <= 64).
template <int S, class T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int m = S / 2; m > 0; m >>= 1) v += shfl_xor_t(v, m, S);
  return v;
}

// Argmax over a group of S lanes with the total order above; every lane ends
// with the group's best.
template <int S>
__device__ __forceinline__ void group_argmax(double& bg, vid& bd, bool prefer_small) {
#pragma unroll
  for (int m = S / 2; m > 0; m >>= 1) {
    double og = __shfl_xor(bg, m, S);
    vid od = __shfl_xor(bd, m, S);
    if (better(og, od, bg, bd, prefer_small)) { bg = og; bd = od; }
  }
}

// Workgroup-wide helpers (BS threads, wave64).
template <int BS, class T>
__device__ __forceinline__ T block_sum(T v, T* red /* [BS/64] */) {
  v = group_sum<64>(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  T s = 0;
#pragma unroll
  for (int i = 0; i < BS / 64; ++i) s += red[i];
  return s;
}

// Sums of a and b over a workgroup with one barrier; the caller has passed a
// barrier since any earlier use of ra and rb.
template <int BS, class T>
__device__ __forceinline__ void block_sum2(T& a, T& b, T* ra, T* rb /* [BS/64] each */) {
  a = group_sum<64>(a);
  b = group_sum<64>(b);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    ra[w] = a;
    rb[w] = b;
  }
  __syncthreads();
  T sa = 0, sb = 0;
#pragma unroll
  for (int i = 0; i < BS / 64; ++i) {
    sa += ra[i];
    sb += rb[i];
  }
  a = sa;
  b = sb;
}

// Workgroup barrier that orders LDS accesses only: loads from global memory
// issued before it stay in flight (__syncthreads also waits for those).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// LEAD = false: rg and rd have not been used before in the kernel, so the
// barrier that orders their reuse is left out.
template <int BS, bool LEAD = true>
__device__ __forceinline__ void block_argmax(double& bg, vid& bd, bool small, double* rg,
                                             vid* rd) {
  group_argmax<64>(bg, bd, small);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if constexpr (LEAD) __syncthreads();
  if (l == 0) { rg[w] = bg; rd[w] = bd; }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < BS / 64; ++i)
    if (better(rg[i], rd[i], bg, bd, small)) { bg = rg[i]; bd = rd[i]; }
}
