// Move kernels of the Louvain pass (louvain.cu:404-1030; v1 filter
// louvain_gl.cu:504,578) as templates, instantiated per arithmetic mode in
// mgl_move_*.hip so that the modes compile in parallel.
//
// CNT = true instantiates the counting variant (distinct (vertex, community)
// pairs and LDS hash-probe statistics through global atomics); timed passes
// launch CNT = false, which carries no counting code.
//
// The gain is evaluated in FP64 in the reference's exact expression order;
// every file including this one is compiled with -ffp-contract=off and the
// pragma below so that no multiply-add is fused.
#pragma once
#pragma clang fp contract(off)

#include "mgl_kernels.h"
#include "mgl_tables.h"
#include <cstdio>

// ---------------------------------------------------------------------------
template <class M>
__device__ __forceinline__ bool vertex_active(const MoveArgs& A, vid v, int& g) {
  g = A.gid ? A.gid[v] : 0;
  return A.gs[g].active != 0;
}

__device__ __forceinline__ vid clamp_dst(vid dst, vid ci, int up, int version) {
  if (version == 1) return dst;                      // louvain_gl.cu:578
  return up ? (dst < ci ? dst : ci) : (dst > ci ? dst : ci);  // louvain.cu:579-585
}


// Record position of v = perm[gi] (its own record and the id its self-loop
// carries in A.col): in degree-bin order the bin's vertices hold consecutive
// positions, so neither the rpos gather nor a scattered record read is needed.
__device__ __forceinline__ vid rec_pos(const MoveArgs& A, vid v, int64_t gi) {
  return A.pos0 >= 0 ? (vid)(A.pos0 + gi) : (A.rpos ? A.rpos[v] : v);
}

// Insert (key, w) into a table of 2^lg slots; the creator of a key also
// stores its community total. Keys are written once (EMPTY -> key) by a
// compare-and-swap that also reads the slot; collisions probe by double
// hashing.
template <class ST, class AT>
__device__ __forceinline__ void ins_tab(vid* K, ST* V, AT* Av, vid key, ST w, AT ak, uint32_t H,
                                        bool pow2, int lg, ProbeCount& pc) {
  uint32_t h = pow2 ? slot_hash(key, lg) : any_slot(key, H);
  const uint32_t st = probe_step(key, H, pow2, lg);
  for (uint32_t i = 0; i < H; ++i) {
    const vid cur = atomicCAS(&K[h], MGL_EMPTY, key);
    if (cur == MGL_EMPTY) Av[h] = ak;
    if (cur == MGL_EMPTY || cur == key) {
      atom_add(&V[h], w);
      pc.add(i + 1);
      return;
    }
    h = probe_next(h, st, H);
  }
  pc.over(H);
}
template <class ST, class AT>
__device__ __forceinline__ void ins_pow2(vid* K, ST* V, AT* Av, vid key, ST w, AT ak, int lg,
                                         ProbeCount& pc) {
  ins_tab<ST, AT>(K, V, Av, key, w, ak, 1u << lg, true, lg, pc);
}
// Same for a prime table size H (multiply-high slot).
template <class ST, class AT>
__device__ __forceinline__ void ins_any(vid* K, ST* V, AT* Av, vid key, ST w, AT ak, uint32_t H,
                                        ProbeCount& pc) {
  ins_tab<ST, AT>(K, V, Av, key, w, ak, H, false, 0, pc);
}

#define MOVE_UNROLL 4

// Walks the row of v with `stride` lanes starting at `lane`, MOVE_UNROLL
// entries per lane per step (their gathers are issued together), inserting
// candidates and summing the own community's weight (eici, self-loop).
template <class M, class AT, bool POW2>
__device__ __forceinline__ void scan_row(const MoveArgs& A, vid vp, vid ci, eid e0, eid e1,
                                         int lane, int stride, vid* K, typename M::S* V, AT* Av,
                                         uint32_t tab, typename M::S& eici,
                                         typename M::S& selfw, ProbeCount& pc) {
  typedef typename M::S ST;
  const Rec<AT>* R = (const Rec<AT>*)A.rec;
  const bool v1 = A.version == 1;
  for (eid e = e0 + lane; e < e1; e += (eid)stride * MOVE_UNROLL) {
    vid u[MOVE_UNROLL];
    Rec<AT> r[MOVE_UNROLL];
#pragma unroll
    for (int j = 0; j < MOVE_UNROLL; ++j) {
      const eid ej = e + (eid)j * stride;
      u[j] = ej < e1 ? colload(A.col, ej) : MGL_EMPTY;
    }
#pragma unroll
    for (int j = 0; j < MOVE_UNROLL; ++j)
      if (u[j] != MGL_EMPTY) r[j] = R[u[j]];
#pragma unroll
    for (int j = 0; j < MOVE_UNROLL; ++j) {
      if (u[j] == MGL_EMPTY) continue;
      const ST wv = wload<M>(A.w, e + (eid)j * stride);
      const vid cj = r[j].c;
      const bool cand = v1 ? (A.up ? cj < ci : cj > ci) : (cj != ci);
      if (cand) {
        if (POW2) ins_pow2<ST, AT>(K, V, Av, cj, wv, r[j].a, (int)tab, pc);
        else ins_any<ST, AT>(K, V, Av, cj, wv, r[j].a, tab, pc);
      } else if (cj == ci) {
        if (u[j] != vp) eici += wv; else selfw += wv;
      }
    }
  }
}

// SHAPE_GROUP: S lanes per vertex, 256/S vertices per workgroup, 2^LG LDS
// slots per vertex.
template <class M, class AT, int S, int LG, bool CNT>
__global__ __launch_bounds__(256) void k_move_group(MoveArgs A, const vid* __restrict__ perm,
                                                    int64_t nb) {
  typedef typename M::S ST;
  typedef typename M::K KT;
  constexpr int GPB = 256 / S;
  constexpr int H = 1 << LG;
  __shared__ vid sk[GPB * H];
  __shared__ ST sv[GPB * H];
  __shared__ AT sa[GPB * H];
  __shared__ unsigned long long sdist;
  const int grp = threadIdx.x / S, lane = threadIdx.x & (S - 1);
  const int64_t gi = (int64_t)blockIdx.x * GPB + grp;
  vid* K = sk + grp * H;
  ST* V = sv + grp * H;
  AT* Av = sa + grp * H;
#pragma unroll
  for (int i = lane; i < H; i += S) { K[i] = MGL_EMPTY; V[i] = (ST)0; }
  if (threadIdx.x == 0) sdist = 0;
  vid v = 0;
  int g = 0;
  bool ok = gi < nb;
  if (ok) { v = perm[gi]; ok = vertex_active<M>(A, v, g); }
  __syncthreads();
  ST eici = 0, selfw = 0;
  vid ci = 0, vp = 0;
  KT acik = 0;
  ProbeCount pc;
  if (ok) {
    vp = rec_pos(A, v, gi);
    const Rec<AT> rv = ((const Rec<AT>*)A.rec)[vp];
    ci = rv.c;
    acik = (KT)rv.a;
    scan_row<M, AT, true>(A, vp, ci, A.off[vp], A.off[vp + 1], lane, S, K, V, Av, LG, eici, selfw,
                          pc);
  }
  __syncthreads();
  if (ok) {
    eici = group_sum<S>(eici);
    selfw = group_sum<S>(selfw);
    const KT ki = ((const KT*)A.k)[vp];
    const KT aci_k = acik - ki;
    const GainIn G = gain_in<M, AT>(A.gs[g].mass, A.gs[g].mass_i, ki, aci_k, eici);
    const bool small = A.tie_small ? true : (A.up != 0);
    Argmax<M, AT> am(ci);
    int cnt = 0;
#pragma unroll
    for (int i = lane; i < H; i += S) {
      const vid cj = K[i];
      if (cj != MGL_EMPTY) {
        ++cnt;
        am.offer(G, V[i], Av[i], cj, small);
      }
    }
    double best = am.best;
    vid dst = am.dst;
    group_argmax<S>(best, dst, small);
    if (CNT && A.stats) {
      cnt = group_sum<S>(cnt);
      if (lane == 0) atomicAdd(&sdist, (unsigned long long)cnt);
    }
    if (lane == 0) {
      A.c_new[v] = clamp_dst(dst, ci, A.up, A.version);
      if (A.vin) ((KT*)A.vin)[v] = (KT)eici + (KT)selfw;
    }
  }
  if constexpr (CNT) flush_probes(A.probes, pc);
  flag_overflow(A.err, pc);
  if (CNT && A.stats) {
    __syncthreads();
    if (threadIdx.x == 0 && sdist) atomicAdd(A.stats, sdist);
  }
}

// ---------------------------------------------------------------------------
// LDS layout of a table of H slots: keys, values, totals (16-byte aligned parts)
__host__ __device__ inline size_t tab_bytes(uint32_t H, size_t sv, size_t sa) {
  return ((4 * (size_t)H + 15) / 16) * 16 + ((sv * H + 15) / 16) * 16 + sa * H;
}

// SHAPE_BLOCK: one workgroup per vertex, an LDS table of H slots (any H).
template <class M, class AT, int BS, bool CNT>
__global__ __launch_bounds__(BS) void k_move_block(MoveArgs A, const vid* __restrict__ perm,
                                                   int64_t nb, uint32_t H) {
  typedef typename M::S ST;
  typedef typename M::K KT;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double rg[BS / 64];
  __shared__ vid rd[BS / 64];
  __shared__ ST rs[BS / 64];
  __shared__ int rc[BS / 64];
  const int64_t gi = blockIdx.x;
  if (gi >= nb) return;
  const vid v = perm[gi];
  int g;
  if (!vertex_active<M>(A, v, g)) return;  // uniform over the workgroup
  vid* K = (vid*)smem;
  ST* V = (ST*)(smem + ((4 * (size_t)H + 15) / 16) * 16);
  AT* Av = (AT*)((unsigned char*)V + ((sizeof(ST) * (size_t)H + 15) / 16) * 16);
  for (uint32_t i = threadIdx.x; i < H; i += BS) { K[i] = MGL_EMPTY; V[i] = (ST)0; }
  const vid vp = rec_pos(A, v, gi);
  const Rec<AT> rv = ((const Rec<AT>*)A.rec)[vp];
  const vid ci = rv.c;
  __syncthreads();
  ST eici = 0, selfw = 0;
  ProbeCount pc;
  scan_row<M, AT, false>(A, vp, ci, A.off[vp], A.off[vp + 1], threadIdx.x, BS, K, V, Av, H, eici,
                         selfw, pc);
  if constexpr (CNT) flush_probes(A.probes, pc);
  flag_overflow(A.err, pc);
  __syncthreads();
  eici = block_sum<BS>(eici, rs);
  selfw = block_sum<BS>(selfw, rs);
  const KT ki = ((const KT*)A.k)[vp];
  const KT aci_k = (KT)rv.a - ki;
  const GainIn G = gain_in<M, AT>(A.gs[g].mass, A.gs[g].mass_i, ki, aci_k, eici);
  const bool small = A.tie_small ? true : (A.up != 0);
  Argmax<M, AT> am(ci);
  int cnt = 0;
  for (uint32_t i = threadIdx.x; i < H; i += BS) {
    const vid cj = K[i];
    if (cj != MGL_EMPTY) {
      ++cnt;
      am.offer(G, V[i], Av[i], cj, small);
    }
  }
  double best = am.best;
  vid dst = am.dst;
  block_argmax<BS, false>(best, dst, small, rg, rd);  // rg, rd unused before
  if (CNT && A.stats) cnt = block_sum<BS>(cnt, rc);
  if (threadIdx.x == 0) {
    A.c_new[v] = clamp_dst(dst, ci, A.up, A.version);
    if (A.vin) ((KT*)A.vin)[v] = (KT)eici + (KT)selfw;
    if (CNT && A.stats) atomicAdd(A.stats, (unsigned long long)cnt);
  }
}

// ---------------------------------------------------------------------------
// Issues the column loads and record gathers of the (at most IT per lane)
// entries of a row; own_insert consumes them. In between, the kernels wait at
// an LDS-only barrier for their table to be cleared, so the gathers are in
// flight while the workgroup synchronises.
template <class AT, int IT>
__device__ __forceinline__ void own_gather(const MoveArgs& A, eid e0, eid e1, int lane,
                                           int stride, vid (&u)[IT], Rec<AT> (&r)[IT]) {
  const Rec<AT>* R = (const Rec<AT>*)A.rec;
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    const eid e = e0 + lane + (eid)j * stride;
    u[j] = e < e1 ? colload(A.col, e) : MGL_EMPTY;
  }
#pragma unroll
  for (int j = 0; j < IT; ++j)
    if (u[j] != MGL_EMPTY) r[j] = R[u[j]];
}

// Inserts the gathered entries: candidates into the table (own[j] = the slot
// if this lane created the key), the own community's entries into eici and
// the self-loop.
template <class M, class AT, int IT, class TB>
__device__ __forceinline__ void own_insert(const MoveArgs& A, vid vp, vid ci, eid e0, int lane,
                                           int stride, TB& tb, uint32_t H, bool pow2, int lg,
                                           const vid (&u)[IT], const Rec<AT> (&r)[IT],
                                           int (&own)[IT], AT (&oa)[IT], typename M::S& eici,
                                           typename M::S& selfw, ProbeCount& pc) {
  typedef typename M::S ST;
  const bool v1 = A.version == 1;
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    own[j] = -1;
    if (u[j] == MGL_EMPTY) continue;
    const eid e = e0 + lane + (eid)j * stride;
    const ST wv = wload<M>(A.w, e);
    const vid cj = r[j].c;
    const bool cand = v1 ? (A.up ? cj < ci : cj > ci) : (cj != ci);
    if (cand) {
      own[j] = tb.ins(cj, wv, H, pow2, lg, pc);
      oa[j] = r[j].a;
    } else if (cj == ci) {
      if (u[j] != vp) eici += wv; else selfw += wv;
    }
  }
}

template <class M, class AT, int S, int LG, int IT, bool CNT>
__global__ __launch_bounds__(256) void k_move_group_own(MoveArgs A, const vid* __restrict__ perm,
                                                        int64_t nb) {
  typedef typename M::S ST;
  typedef typename M::K KT;
  constexpr bool PACK = sizeof(ST) == 4;
  typedef OwnTab<ST, PACK> TB;
  constexpr int GPB = 256 / S;
  constexpr int H = 1 << LG;
  __shared__ __align__(16) unsigned char tabs[GPB * H * TB::slot + 16 * GPB];
  __shared__ unsigned long long sdist;
  const int grp = threadIdx.x / S, lane = threadIdx.x & (S - 1);
  const int64_t gi = (int64_t)blockIdx.x * GPB + grp;
  TB tb;
  tb.init(tabs + (size_t)grp * (H * TB::slot + 16), H);
#pragma unroll
  for (int i = lane; i < H; i += S) tb.clear(i);
  if (threadIdx.x == 0) sdist = 0;
  vid v = 0;
  int g = 0;
  bool ok = gi < nb;
  if (ok) { v = perm[gi]; ok = vertex_active<M>(A, v, g); }
  ST eici = 0, selfw = 0;
  vid ci = 0, vp = 0;
  KT acik = 0;
  eid e0 = 0, e1 = 0;
  int own[IT];
  AT oa[IT];
  vid u[IT];
  Rec<AT> r[IT];
  ProbeCount pc;
#pragma unroll
  for (int j = 0; j < IT; ++j) own[j] = -1;
  if (ok) {
    vp = rec_pos(A, v, gi);
    const Rec<AT> rv = ((const Rec<AT>*)A.rec)[vp];
    ci = rv.c;
    acik = (KT)rv.a;
    e0 = A.off[vp];
    e1 = A.off[vp + 1];
  }
  own_gather<AT, IT>(A, e0, e1, lane, S, u, r);  // no entries unless ok
  lds_barrier();  // the tables are cleared
  if (ok)
    own_insert<M, AT, IT, TB>(A, vp, ci, e0, lane, S, tb, H, true, LG, u, r, own, oa, eici,
                              selfw, pc);
  __syncthreads();
  if (ok) {
    eici = group_sum<S>(eici);
    selfw = group_sum<S>(selfw);
    const KT ki = ((const KT*)A.k)[vp];
    const KT aci_k = acik - ki;
    const GainIn G = gain_in<M, AT>(A.gs[g].mass, A.gs[g].mass_i, ki, aci_k, eici);
    const bool small = A.tie_small ? true : (A.up != 0);
    double best = 0.0;
    vid dst = ci;
    int cnt = 0;
    own_gains<M, AT, IT, TB>(tb, own, oa, G, small, best, dst, cnt);
    group_argmax<S>(best, dst, small);
    if (CNT && A.stats) {
      cnt = group_sum<S>(cnt);
      if (lane == 0) atomicAdd(&sdist, (unsigned long long)cnt);
    }
    if (lane == 0) {
      A.c_new[v] = clamp_dst(dst, ci, A.up, A.version);
      if (A.vin) ((KT*)A.vin)[v] = (KT)eici + (KT)selfw;
    }
  }
  if constexpr (CNT) flush_probes(A.probes, pc);
  flag_overflow(A.err, pc);
  if (CNT && A.stats) {
    __syncthreads();
    if (threadIdx.x == 0 && sdist) atomicAdd(A.stats, sdist);
  }
}

template <class M, class AT, int BS, int IT, bool CNT>
__global__ __launch_bounds__(BS) void k_move_block_own(MoveArgs A, const vid* __restrict__ perm,
                                                       int64_t nb, uint32_t H) {
  typedef typename M::S ST;
  typedef typename M::K KT;
  constexpr bool PACK = sizeof(ST) == 4;
  typedef OwnTab<ST, PACK> TB;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double rg[BS / 64];
  __shared__ vid rd[BS / 64];
  __shared__ ST rs[BS / 64];
  __shared__ ST rs2[BS / 64];
  __shared__ int rc[BS / 64];
  const int64_t gi = blockIdx.x;
  if (gi >= nb) return;
  const vid v = perm[gi];
  int g;
  if (!vertex_active<M>(A, v, g)) return;  // uniform over the workgroup
  TB tb;
  tb.init(smem, H);
  for (uint32_t i = threadIdx.x; i < H; i += BS) tb.clear(i);
  const vid vp = rec_pos(A, v, gi);
  const Rec<AT> rv = ((const Rec<AT>*)A.rec)[vp];
  const vid ci = rv.c;
  const eid e0 = A.off[vp], e1 = A.off[vp + 1];
  ST eici = 0, selfw = 0;
  int own[IT];
  AT oa[IT];
  vid u[IT];
  Rec<AT> r[IT];
  ProbeCount pc;
  // the table must be cleared before any insert: workgroups of up to 512
  // threads wait at an LDS-only barrier after issuing their gathers; a
  // 1024-thread workgroup (one per CU) waits before them, which measured faster
  if constexpr (BS >= 1024) __syncthreads();
  own_gather<AT, IT>(A, e0, e1, threadIdx.x, BS, u, r);
  if constexpr (BS < 1024) lds_barrier();
  own_insert<M, AT, IT, TB>(A, vp, ci, e0, threadIdx.x, BS, tb, H, false, 0, u, r, own, oa, eici,
                            selfw, pc);
  if constexpr (CNT) flush_probes(A.probes, pc);
  flag_overflow(A.err, pc);
  __syncthreads();  // the table is complete
  block_sum2<BS>(eici, selfw, rs, rs2);  // eici and the self-loop, one barrier
  // (loaded here: a load issued before the barrier above would be waited for
  // there; on R-MAT 24 rows of 4,097-8,192 entries that cost 0.08 ms a pass)
  const KT ki = ((const KT*)A.k)[vp];
  const KT aci_k = (KT)rv.a - ki;
  const GainIn G = gain_in<M, AT>(A.gs[g].mass, A.gs[g].mass_i, ki, aci_k, eici);
  const bool small = A.tie_small ? true : (A.up != 0);
  double best = 0.0;
  vid dst = ci;
  int cnt = 0;
  own_gains<M, AT, IT, TB>(tb, own, oa, G, small, best, dst, cnt);
  block_argmax<BS, false>(best, dst, small, rg, rd);  // rg, rd unused before
  if (CNT && A.stats) cnt = block_sum<BS>(cnt, rc);
  if (threadIdx.x == 0) {
    A.c_new[v] = clamp_dst(dst, ci, A.up, A.version);
    if (A.vin) ((KT*)A.vin)[v] = (KT)eici + (KT)selfw;
    if (CNT && A.stats) atomicAdd(A.stats, (unsigned long long)cnt);
  }
}

// ---------------------------------------------------------------------------
// SHAPE_LANE: one lane per vertex for rows of at most D entries. The row is
// gathered into registers and equal communities are merged by pairwise
// comparison (the first occurrence accumulates), so no LDS, barrier or
// cross-lane operation is involved.
template <class M, class AT, int D, bool CNT>
__global__ __launch_bounds__(256) void k_move_lane(MoveArgs A, const vid* __restrict__ perm,
                                                   int64_t nb) {
  typedef typename M::S ST;
  typedef typename M::K KT;
  __shared__ unsigned long long sdist;
  if (threadIdx.x == 0) sdist = 0;
  if (CNT && A.stats) __syncthreads();
  const int64_t gi = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int g = 0;
  vid v = 0;
  bool ok = gi < nb;
  if (ok) { v = perm[gi]; ok = vertex_active<M>(A, v, g); }
  int cnt = 0;
  if (ok) {
    const Rec<AT>* R = (const Rec<AT>*)A.rec;
    const vid vp = rec_pos(A, v, gi);
    const Rec<AT> rv = R[vp];
    const vid ci = rv.c;
    const eid e0 = A.off[vp];
    const int d = (int)(A.off[vp + 1] - e0);
    vid u[D];
    Rec<AT> r[D];
    ST w[D];
#pragma unroll
    for (int j = 0; j < D; ++j) u[j] = j < d ? colload(A.col, e0 + j) : MGL_EMPTY;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      if (u[j] != MGL_EMPTY) r[j] = R[u[j]];
      w[j] = (u[j] != MGL_EMPTY) ? wload<M>(A.w, e0 + j) : (ST)0;
    }
    const bool v1 = A.version == 1;
    ST eici = 0, selfw = 0;
    bool cand[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      cand[j] = false;
      if (u[j] == MGL_EMPTY) continue;
      const vid cj = r[j].c;
      cand[j] = v1 ? (A.up ? cj < ci : cj > ci) : (cj != ci);
      if (!cand[j] && cj == ci) {
        if (u[j] != vp) eici += w[j]; else selfw += w[j];
      }
    }
    // merge duplicates into their first occurrence
#pragma unroll
    for (int j = 1; j < D; ++j) {
#pragma unroll
      for (int i = 0; i < j; ++i) {
        if (cand[i] && cand[j] && r[i].c == r[j].c) {
          w[i] += w[j];
          cand[j] = false;
        }
      }
    }
    const KT ki = ((const KT*)A.k)[vp];
    const KT aci_k = (KT)rv.a - ki;
    const GainIn G = gain_in<M, AT>(A.gs[g].mass, A.gs[g].mass_i, ki, aci_k, eici);
    const bool small = A.tie_small ? true : (A.up != 0);
    Argmax<M, AT> am(ci);
#pragma unroll
    for (int j = 0; j < D; ++j) {
      if (!cand[j]) continue;
      ++cnt;
      am.offer(G, w[j], r[j].a, r[j].c, small);
    }
    A.c_new[v] = clamp_dst(am.dst, ci, A.up, A.version);
    if (A.vin) ((KT*)A.vin)[v] = (KT)eici + (KT)selfw;
  }
  if (CNT && A.stats) {
    cnt = group_sum<64>(cnt);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(&sdist, (unsigned long long)cnt);
    __syncthreads();
    if (threadIdx.x == 0 && sdist) atomicAdd(A.stats, sdist);
  }
}


// Kernel names as profilers print them (template arguments spelled out).
template <class M> inline const char* mode_name();
template <> inline const char* mode_name<ModeUnit>() { return "ModeUnit"; }
template <> inline const char* mode_name<ModeInt>() { return "ModeInt"; }
template <> inline const char* mode_name<ModeI32>() { return "ModeI32"; }
template <> inline const char* mode_name<ModeF64>() { return "ModeF64"; }
template <> inline const char* mode_name<ModeF32>() { return "ModeF32"; }
template <class T> inline const char* type_name();
template <> inline const char* type_name<int32_t>() { return "int"; }
template <> inline const char* type_name<int64_t>() { return "long"; }
template <> inline const char* type_name<double>() { return "double"; }

// Launches the move kernel of a bin; with name != nullptr it only writes the
// name of the kernel it would launch.
template <class M, class AT, bool CNT>
static int move_bin_t(const BinSpec& bs, const MoveArgs& A, const vid* perm, int64_t nb,
                      hipStream_t st, char* name = nullptr, int len = 0) {
  typedef typename M::S ST;
  if (nb <= 0 || bs.shape == SHAPE_NONE) return 0;
#define NAME_(...)                                                  \
  if (name) {                                                       \
    snprintf(name, (size_t)len, __VA_ARGS__);                       \
    return 0;                                                       \
  }
  const char* mn = mode_name<M>();
  const char* an = type_name<AT>();
  const char* cn = CNT ? "true" : "false";
  if (bs.shape == SHAPE_LANE) {
    const unsigned grid = (unsigned)((nb + 255) / 256);
#define L_(D_)                                                                \
  if (bs.upper <= D_) {                                                       \
    NAME_("k_move_lane<%s, %s, %d, %s>", mn, an, D_, cn)                              \
    k_move_lane<M, AT, D_, CNT><<<grid, 256, 0, st>>>(A, perm, nb);                \
    return 0;                                                                 \
  }
    L_(4) L_(8) L_(16)
#undef L_
    return -1;
  }
  if (bs.shape == SHAPE_GROUP) {
    const int gpb = 256 / bs.s;
    const unsigned grid = (unsigned)((nb + gpb - 1) / gpb);
#define G_(S_, LG_)                                                                     \
  if (bs.s == S_ && bs.size == LG_) {                                                   \
    NAME_("k_move_group<%s, %s, %d, %d, %s>", mn, an, S_, LG_, cn)                              \
    k_move_group<M, AT, S_, LG_, CNT><<<grid, 256, 0, st>>>(A, perm, nb);                    \
    return 0;                                                                           \
  }
// owner-based instances (lanes S, table 2^LG, entries per lane IT); the
// smallest IT that covers the bin is taken, longer rows use the scan kernel
#define O_(S_, LG_, IT_)                                                                \
  if (bs.s == S_ && bs.size == LG_ && need <= IT_) {                                    \
    NAME_("k_move_group_own<%s, %s, %d, %d, %d, %s>", mn, an, S_, LG_, IT_, cn)                 \
    k_move_group_own<M, AT, S_, LG_, IT_, CNT><<<grid, 256, 0, st>>>(A, perm, nb);           \
    return 0;                                                                           \
  }
    const int64_t need = (bs.upper + bs.s - 1) / bs.s;
    O_(4, 3, 1) O_(4, 3, 2) O_(4, 4, 4) O_(4, 5, 4)
    O_(8, 4, 1) O_(8, 5, 2) O_(8, 5, 4) O_(8, 6, 4)
    O_(16, 5, 1) O_(16, 6, 2) O_(16, 6, 4) O_(16, 7, 4)
    O_(32, 6, 1) O_(32, 7, 2) O_(32, 7, 4) O_(32, 8, 4)
    O_(64, 7, 1) O_(64, 8, 2) O_(64, 8, 4) O_(64, 9, 4) O_(64, 9, 8) O_(64, 10, 8)
    G_(4, 3) G_(4, 4) G_(8, 4) G_(8, 5) G_(16, 5) G_(16, 6) G_(32, 6) G_(32, 7)
    G_(64, 7) G_(64, 8) G_(64, 9) G_(64, 10)
#undef O_
#undef G_
  } else if (bs.shape == SHAPE_BLOCK) {
    const uint32_t H = (uint32_t)bs.size;
    const size_t lds_own = (sizeof(ST) == 4) ? 8 * (size_t)H
                                             : ((4 * (size_t)H + 15) / 16) * 16 + sizeof(ST) * (size_t)H;
    const int64_t it = (bs.upper + bs.s - 1) / bs.s;
#define B_(BS_, IT_)                                                                        \
  if (bs.s == BS_ && it <= IT_) {                                                           \
    NAME_("k_move_block_own<%s, %s, %d, %d, %s>", mn, an, BS_, IT_, cn)                             \
    k_move_block_own<M, AT, BS_, IT_, CNT><<<(unsigned)nb, BS_, lds_own, st>>>(A, perm, nb, H);  \
    return 0;                                                                               \
  }
    B_(256, 1) B_(256, 2) B_(256, 4) B_(256, 8) B_(256, 16)
    B_(512, 1) B_(512, 2) B_(512, 4) B_(512, 8) B_(512, 16)
    B_(1024, 1) B_(1024, 2) B_(1024, 4) B_(1024, 8) B_(1024, 16)
#undef B_
    const size_t lds = tab_bytes(H, sizeof(ST), sizeof(AT));
    if (bs.s == 256 || bs.s == 512 || bs.s == 1024) NAME_("k_move_block<%s, %s, %d, %s>", mn, an, bs.s, cn)
    if (bs.s == 256) {
      k_move_block<M, AT, 256, CNT><<<(unsigned)nb, 256, lds, st>>>(A, perm, nb, H);
      return 0;
    }
    if (bs.s == 512) {
      k_move_block<M, AT, 512, CNT><<<(unsigned)nb, 512, lds, st>>>(A, perm, nb, H);
      return 0;
    }
    if (bs.s == 1024) {
      k_move_block<M, AT, 1024, CNT><<<(unsigned)nb, 1024, lds, st>>>(A, perm, nb, H);
      return 0;
    }
  }
#undef NAME_
  return -1;
}

// Opt the LDS-table kernels of one (mode, total type) into the device's
// largest workgroup LDS size.
template <class M, class AT, bool CNT>
static void move_allow_lds_cnt(int bytes) {
#define A_(BS_)                                                                      \
  (void)hipFuncSetAttribute((const void*)k_move_block<M, AT, BS_, CNT>,              \
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
#define O_(BS_, IT_)                                                                 \
  (void)hipFuncSetAttribute((const void*)k_move_block_own<M, AT, BS_, IT_, CNT>,     \
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
#define P_(BS_) A_(BS_) O_(BS_, 1) O_(BS_, 2) O_(BS_, 4) O_(BS_, 8) O_(BS_, 16)
  P_(256) P_(512) P_(1024)
#undef P_
#undef O_
#undef A_
}
template <class M, class AT>
static void move_allow_lds_t(int bytes) {
  move_allow_lds_cnt<M, AT, false>(bytes);
  move_allow_lds_cnt<M, AT, true>(bytes);
}

// Launch (or, with name != nullptr, name) the move kernel of a bin for one
// (mode, total type) pair; cnt selects the counting variant.
template <class M, class AT>
static int move_bin_mode(bool cnt, const BinSpec& bs, const MoveArgs& A, const vid* perm,
                         int64_t nb, hipStream_t st, char* name, int len) {
  return cnt ? move_bin_t<M, AT, true>(bs, A, perm, nb, st, name, len)
             : move_bin_t<M, AT, false>(bs, A, perm, nb, st, name, len);
}
