// gfx950 kernels of the Louvain pass: degree binning, binned modularity-gain
// moves, per-vertex row reductions, community-weight update, per-graph
// modularity reduction and the device-side inner-loop control.
//
// Reference behaviour (NikodemGapski/MG-alphaGCD):
//   move rule          src/louvain/louvain.cu:454-585, :633-805, :855-1026
//   v1 filter          src/louvain/louvain_gl.cu:504,578
//   vertex weights     src/louvain/louvain.cu:30-74
//   weight update      src/louvain/louvain.cu:1299-1396; louvain_baseline.cu:1319-1372
//   modularity         src/louvain/louvain.cu:105-238
//   binning            src/bin/BIN.cu:33-121, 298-342, 468-499
//   inner loop control src/louvain/louvain.cu:1548-1660
//
// The gain is evaluated in FP64 in the reference's exact expression order;
// this file is compiled with -ffp-contract=off and the pragma below so that no
// multiply-add is fused, which keeps the gains bit-identical to the oracle's.
#pragma clang fp contract(off)

#include "mgl_kernels.h"
#include "mgl_tables.h"
#include <climits>

static unsigned grid_for(int64_t n, int per_block, unsigned cap) {
  int64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

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

// ---------------------------------------------------------------------------
// Per-vertex record {c[u], a[c[u]]}, rebuilt before every move pass, so that
// one gather per CSR entry brings both the neighbour's community and that
// community's total. The slot creator stores the total in the table, so the
// gain loop reads no global memory. AT = int32 in the unit mode while
// 2m < 2^31 (an 8-byte record), else the mode's total type (16 bytes).
// Records of a large level are stored in degree-bin order (rpos): the
// frequently gathered records of high-degree vertices then share cache lines.
template <class KT, class AT>
__global__ __launch_bounds__(256) void k_build_rec(int64_t n, const vid* __restrict__ c,
                                                   const KT* __restrict__ a,
                                                   Rec<AT>* __restrict__ rec,
                                                   const vid* __restrict__ list,
                                                   const vid* __restrict__ rpos,
                                                   const vid* __restrict__ rinv) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    int64_t v, at;
    if (list) {
      v = (int64_t)list[i];
      at = rpos ? (int64_t)rpos[v] : v;
    } else if (rinv) {
      v = (int64_t)rinv[i];
      at = i;
    } else {
      v = i;
      at = i;
    }
    const vid x = c[v];
    Rec<AT> r;
    r.c = x;
    r.a = (AT)a[x];
    rec[at] = r;
  }
}

void launch_build_rec(int mode, int rec32, int64_t n, const vid* c, const void* a, void* rec,
                      hipStream_t st, const vid* list, const vid* rpos, const vid* rinv) {
  if (n <= 0) return;
  const unsigned grid = grid_for(n, 256, 8192);
  if ((mode == MODE_UNIT || mode == MODE_I32) && rec32)
    k_build_rec<int64_t, int32_t><<<grid, 256, 0, st>>>(n, c, (const int64_t*)a,
                                                        (Rec<int32_t>*)rec, list, rpos, rinv);
  else if (mode_is_int(mode))
    k_build_rec<int64_t, int64_t><<<grid, 256, 0, st>>>(n, c, (const int64_t*)a,
                                                        (Rec<int64_t>*)rec, list, rpos, rinv);
  else
    k_build_rec<double, double><<<grid, 256, 0, st>>>(n, c, (const double*)a, (Rec<double>*)rec,
                                                      list, rpos, rinv);
}

__global__ __launch_bounds__(256) void k_record_layout(const vid* __restrict__ perm, int64_t n,
                                                       vid* __restrict__ rpos,
                                                       vid* __restrict__ rinv) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const vid v = perm[i];
    rinv[i] = v;
    rpos[v] = (vid)i;
  }
}

__global__ __launch_bounds__(256) void k_remap_cols(const vid* __restrict__ col, int64_t nnz,
                                                    const vid* __restrict__ rpos,
                                                    vid* __restrict__ colr) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < nnz;
       e += (int64_t)gridDim.x * 256)
    colr[e] = rpos[colload(col, e)];
}

void launch_record_layout(const vid* perm, int64_t n, vid* rpos, vid* rinv, hipStream_t st) {
  if (n <= 0) return;
  k_record_layout<<<grid_for(n, 256, 16384), 256, 0, st>>>(perm, n, rpos, rinv);
}

void launch_remap_cols(const vid* col, int64_t nnz, const vid* rpos, vid* colr, hipStream_t st) {
  if (nnz <= 0) return;
  k_remap_cols<<<grid_for(nnz, 256, 65536), 256, 0, st>>>(col, nnz, rpos, colr);
}

// Record position of v (its own record and the id its self-loop carries in A.col).
__device__ __forceinline__ vid rec_pos(const MoveArgs& A, vid v) {
  return A.rpos ? A.rpos[v] : v;
}

// Insert (key, w) into a table of 2^lg slots; the creator of a key also
// stores its community total. Keys are written once (EMPTY -> key) by a
// compare-and-swap that also reads the slot; collisions probe by double
// hashing.
template <class ST, class AT>
__device__ __forceinline__ void ins_tab(vid* K, ST* V, AT* Av, vid key, ST w, AT ak, uint32_t H,
                                        bool pow2, int lg) {
  uint32_t h = pow2 ? slot_hash(key, lg) : any_slot(key, H);
  const uint32_t st = probe_step(key, H, pow2, lg);
  for (uint32_t i = 0; i < H; ++i) {
    const vid cur = atomicCAS(&K[h], MGL_EMPTY, key);
    if (cur == MGL_EMPTY) Av[h] = ak;
    if (cur == MGL_EMPTY || cur == key) { atom_add(&V[h], w); return; }
    h = probe_next(h, st, H);
  }
}
template <class ST, class AT>
__device__ __forceinline__ void ins_pow2(vid* K, ST* V, AT* Av, vid key, ST w, AT ak, int lg) {
  ins_tab<ST, AT>(K, V, Av, key, w, ak, 1u << lg, true, lg);
}
// Same for a prime table size H (multiply-high slot).
template <class ST, class AT>
__device__ __forceinline__ void ins_any(vid* K, ST* V, AT* Av, vid key, ST w, AT ak, uint32_t H) {
  ins_tab<ST, AT>(K, V, Av, key, w, ak, H, false, 0);
}

#define MOVE_UNROLL 4

// Walks the row of v with `stride` lanes starting at `lane`, MOVE_UNROLL
// entries per lane per step (their gathers are issued together), inserting
// candidates and summing the own community's weight (eici, self-loop).
template <class M, class AT, bool POW2>
__device__ __forceinline__ void scan_row(const MoveArgs& A, vid vp, vid ci, eid e0, eid e1,
                                         int lane, int stride, vid* K, typename M::S* V, AT* Av,
                                         uint32_t tab, typename M::S& eici,
                                         typename M::S& selfw) {
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
        if (POW2) ins_pow2<ST, AT>(K, V, Av, cj, wv, r[j].a, (int)tab);
        else ins_any<ST, AT>(K, V, Av, cj, wv, r[j].a, tab);
      } else if (cj == ci) {
        if (u[j] != vp) eici += wv; else selfw += wv;
      }
    }
  }
}

// SHAPE_GROUP: S lanes per vertex, 256/S vertices per workgroup, 2^LG LDS
// slots per vertex.
template <class M, class AT, int S, int LG>
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
  vid ci = 0;
  KT acik = 0;
  if (ok) {
    const vid vp = rec_pos(A, v);
    const Rec<AT> rv = ((const Rec<AT>*)A.rec)[vp];
    ci = rv.c;
    acik = (KT)rv.a;
    scan_row<M, AT, true>(A, vp, ci, A.off[v], A.off[v + 1], lane, S, K, V, Av, LG, eici, selfw);
  }
  __syncthreads();
  if (ok) {
    eici = group_sum<S>(eici);
    selfw = group_sum<S>(selfw);
    const double Mh = A.gs[g].mass / 2.0;
    const KT ki = ((const KT*)A.k)[v];
    const KT aci_k = acik - ki;
    const double deici = (double)eici, dki = (double)ki, daci = (double)aci_k;
    const bool small = A.tie_small ? true : (A.up != 0);
    double best = 0.0;
    vid dst = ci;
    int cnt = 0;
#pragma unroll
    for (int i = lane; i < H; i += S) {
      const vid cj = K[i];
      if (cj != MGL_EMPTY) {
        ++cnt;
        const double wij = gain((double)V[i], deici, dki, daci, (double)Av[i], Mh);
        if (better(wij, cj, best, dst, small)) { best = wij; dst = cj; }
      }
    }
    group_argmax<S>(best, dst, small);
    if (A.stats) {
      cnt = group_sum<S>(cnt);
      if (lane == 0) atomicAdd(&sdist, (unsigned long long)cnt);
    }
    if (lane == 0) {
      A.c_new[v] = clamp_dst(dst, ci, A.up, A.version);
      if (A.vin) ((KT*)A.vin)[v] = (KT)eici + (KT)selfw;
    }
  }
  if (A.stats) {
    __syncthreads();
    if (threadIdx.x == 0 && sdist) atomicAdd(A.stats, sdist);
  }
}

// ---------------------------------------------------------------------------
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

template <int BS>
__device__ __forceinline__ void block_argmax(double& bg, vid& bd, bool small, double* rg,
                                             vid* rd) {
  group_argmax<64>(bg, bd, small);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) { rg[w] = bg; rd[w] = bd; }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < BS / 64; ++i)
    if (better(rg[i], rd[i], bg, bd, small)) { bg = rg[i]; bd = rd[i]; }
}

// LDS layout of a table of H slots: keys, values, totals (16-byte aligned parts)
__host__ __device__ inline size_t tab_bytes(uint32_t H, size_t sv, size_t sa) {
  return ((4 * (size_t)H + 15) / 16) * 16 + ((sv * H + 15) / 16) * 16 + sa * H;
}

// SHAPE_BLOCK: one workgroup per vertex, an LDS table of H slots (any H).
template <class M, class AT, int BS>
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
  const vid vp = rec_pos(A, v);
  const Rec<AT> rv = ((const Rec<AT>*)A.rec)[vp];
  const vid ci = rv.c;
  __syncthreads();
  ST eici = 0, selfw = 0;
  scan_row<M, AT, false>(A, vp, ci, A.off[v], A.off[v + 1], threadIdx.x, BS, K, V, Av, H, eici,
                         selfw);
  __syncthreads();
  eici = block_sum<BS>(eici, rs);
  selfw = block_sum<BS>(selfw, rs);
  const double Mh = A.gs[g].mass / 2.0;
  const KT ki = ((const KT*)A.k)[v];
  const KT aci_k = (KT)rv.a - ki;
  const double deici = (double)eici, dki = (double)ki, daci = (double)aci_k;
  const bool small = A.tie_small ? true : (A.up != 0);
  double best = 0.0;
  vid dst = ci;
  int cnt = 0;
  for (uint32_t i = threadIdx.x; i < H; i += BS) {
    const vid cj = K[i];
    if (cj != MGL_EMPTY) {
      ++cnt;
      const double wij = gain((double)V[i], deici, dki, daci, (double)Av[i], Mh);
      if (better(wij, cj, best, dst, small)) { best = wij; dst = cj; }
    }
  }
  block_argmax<BS>(best, dst, small, rg, rd);
  if (A.stats) cnt = block_sum<BS>(cnt, rc);
  if (threadIdx.x == 0) {
    A.c_new[v] = clamp_dst(dst, ci, A.up, A.version);
    if (A.vin) ((KT*)A.vin)[v] = (KT)eici + (KT)selfw;
    if (A.stats) atomicAdd(A.stats, (unsigned long long)cnt);
  }
}

// ---------------------------------------------------------------------------
// Gathers and inserts the (at most IT per lane) entries of v's row.
template <class M, class AT, int IT, class TB>
__device__ __forceinline__ void own_row(const MoveArgs& A, vid vp, vid ci, eid e0, eid e1,
                                        int lane, int stride, TB& tb, uint32_t H, bool pow2,
                                        int lg, int (&own)[IT], AT (&oa)[IT],
                                        typename M::S& eici, typename M::S& selfw) {
  typedef typename M::S ST;
  const Rec<AT>* R = (const Rec<AT>*)A.rec;
  const bool v1 = A.version == 1;
  vid u[IT];
  Rec<AT> r[IT];
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    const eid e = e0 + lane + (eid)j * stride;
    u[j] = e < e1 ? colload(A.col, e) : MGL_EMPTY;
  }
#pragma unroll
  for (int j = 0; j < IT; ++j)
    if (u[j] != MGL_EMPTY) r[j] = R[u[j]];
#pragma unroll
  for (int j = 0; j < IT; ++j) {
    own[j] = -1;
    if (u[j] == MGL_EMPTY) continue;
    const eid e = e0 + lane + (eid)j * stride;
    const ST wv = wload<M>(A.w, e);
    const vid cj = r[j].c;
    const bool cand = v1 ? (A.up ? cj < ci : cj > ci) : (cj != ci);
    if (cand) {
      own[j] = tb.ins(cj, wv, H, pow2, lg);
      oa[j] = r[j].a;
    } else if (cj == ci) {
      if (u[j] != vp) eici += wv; else selfw += wv;
    }
  }
}

template <class M, class AT, int S, int LG, int IT>
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
  __syncthreads();
  ST eici = 0, selfw = 0;
  vid ci = 0;
  KT acik = 0;
  int own[IT];
  AT oa[IT];
#pragma unroll
  for (int j = 0; j < IT; ++j) own[j] = -1;
  if (ok) {
    const vid vp = rec_pos(A, v);
    const Rec<AT> rv = ((const Rec<AT>*)A.rec)[vp];
    ci = rv.c;
    acik = (KT)rv.a;
    own_row<M, AT, IT, TB>(A, vp, ci, A.off[v], A.off[v + 1], lane, S, tb, H, true, LG, own, oa,
                           eici, selfw);
  }
  __syncthreads();
  if (ok) {
    eici = group_sum<S>(eici);
    selfw = group_sum<S>(selfw);
    const double Mh = A.gs[g].mass / 2.0;
    const KT ki = ((const KT*)A.k)[v];
    const KT aci_k = acik - ki;
    const bool small = A.tie_small ? true : (A.up != 0);
    double best = 0.0;
    vid dst = ci;
    int cnt = 0;
    own_gains<M, AT, IT, TB>(tb, own, oa, (double)eici, (double)ki, (double)aci_k, Mh, small,
                             best, dst, cnt);
    group_argmax<S>(best, dst, small);
    if (A.stats) {
      cnt = group_sum<S>(cnt);
      if (lane == 0) atomicAdd(&sdist, (unsigned long long)cnt);
    }
    if (lane == 0) {
      A.c_new[v] = clamp_dst(dst, ci, A.up, A.version);
      if (A.vin) ((KT*)A.vin)[v] = (KT)eici + (KT)selfw;
    }
  }
  if (A.stats) {
    __syncthreads();
    if (threadIdx.x == 0 && sdist) atomicAdd(A.stats, sdist);
  }
}

template <class M, class AT, int BS, int IT>
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
  __shared__ int rc[BS / 64];
  const int64_t gi = blockIdx.x;
  if (gi >= nb) return;
  const vid v = perm[gi];
  int g;
  if (!vertex_active<M>(A, v, g)) return;  // uniform over the workgroup
  TB tb;
  tb.init(smem, H);
  for (uint32_t i = threadIdx.x; i < H; i += BS) tb.clear(i);
  const vid vp = rec_pos(A, v);
  const Rec<AT> rv = ((const Rec<AT>*)A.rec)[vp];
  const vid ci = rv.c;
  __syncthreads();
  ST eici = 0, selfw = 0;
  int own[IT];
  AT oa[IT];
  own_row<M, AT, IT, TB>(A, vp, ci, A.off[v], A.off[v + 1], threadIdx.x, BS, tb, H, false, 0, own,
                         oa, eici, selfw);
  __syncthreads();
  eici = block_sum<BS>(eici, rs);
  selfw = block_sum<BS>(selfw, rs);
  const double Mh = A.gs[g].mass / 2.0;
  const KT ki = ((const KT*)A.k)[v];
  const KT aci_k = (KT)rv.a - ki;
  const bool small = A.tie_small ? true : (A.up != 0);
  double best = 0.0;
  vid dst = ci;
  int cnt = 0;
  own_gains<M, AT, IT, TB>(tb, own, oa, (double)eici, (double)ki, (double)aci_k, Mh, small, best,
                           dst, cnt);
  block_argmax<BS>(best, dst, small, rg, rd);
  if (A.stats) cnt = block_sum<BS>(cnt, rc);
  if (threadIdx.x == 0) {
    A.c_new[v] = clamp_dst(dst, ci, A.up, A.version);
    if (A.vin) ((KT*)A.vin)[v] = (KT)eici + (KT)selfw;
    if (A.stats) atomicAdd(A.stats, (unsigned long long)cnt);
  }
}

// ---------------------------------------------------------------------------
// SHAPE_LANE: one lane per vertex for rows of at most D entries. The row is
// gathered into registers and equal communities are merged by pairwise
// comparison (the first occurrence accumulates), so no LDS, barrier or
// cross-lane operation is involved.
template <class M, class AT, int D>
__global__ __launch_bounds__(256) void k_move_lane(MoveArgs A, const vid* __restrict__ perm,
                                                   int64_t nb) {
  typedef typename M::S ST;
  typedef typename M::K KT;
  __shared__ unsigned long long sdist;
  if (threadIdx.x == 0) sdist = 0;
  if (A.stats) __syncthreads();
  const int64_t gi = (int64_t)blockIdx.x * 256 + threadIdx.x;
  int g = 0;
  vid v = 0;
  bool ok = gi < nb;
  if (ok) { v = perm[gi]; ok = vertex_active<M>(A, v, g); }
  int cnt = 0;
  if (ok) {
    const Rec<AT>* R = (const Rec<AT>*)A.rec;
    const vid vp = rec_pos(A, v);
    const Rec<AT> rv = R[vp];
    const vid ci = rv.c;
    const eid e0 = A.off[v];
    const int d = (int)(A.off[v + 1] - e0);
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
    const double Mh = A.gs[g].mass / 2.0;
    const KT ki = ((const KT*)A.k)[v];
    const KT aci_k = (KT)rv.a - ki;
    const double deici = (double)eici, dki = (double)ki, daci = (double)aci_k;
    const bool small = A.tie_small ? true : (A.up != 0);
    double best = 0.0;
    vid dst = ci;
#pragma unroll
    for (int j = 0; j < D; ++j) {
      if (!cand[j]) continue;
      ++cnt;
      const double wij = gain((double)w[j], deici, dki, daci, (double)r[j].a, Mh);
      if (better(wij, r[j].c, best, dst, small)) { best = wij; dst = r[j].c; }
    }
    A.c_new[v] = clamp_dst(dst, ci, A.up, A.version);
    if (A.vin) ((KT*)A.vin)[v] = (KT)eici + (KT)selfw;
  }
  if (A.stats) {
    cnt = group_sum<64>(cnt);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(&sdist, (unsigned long long)cnt);
    __syncthreads();
    if (threadIdx.x == 0 && sdist) atomicAdd(A.stats, sdist);
  }
}

// ---------------------------------------------------------------------------
BinLayout default_bins(int lds_bytes, int slot_bytes) {
  BinLayout L;
  int i = 0;
  L.b[i++] = {0, SHAPE_NONE, 0, 0};
  L.b[i++] = {4, SHAPE_LANE, 1, 0};
  L.b[i++] = {8, SHAPE_LANE, 1, 0};
  L.b[i++] = {16, SHAPE_GROUP, 4, 5};
  L.b[i++] = {32, SHAPE_GROUP, 8, 6};
  L.b[i++] = {64, SHAPE_GROUP, 16, 7};
  L.b[i++] = {128, SHAPE_GROUP, 32, 8};
  L.b[i++] = {256, SHAPE_GROUP, 64, 9};
  L.b[i++] = {512, SHAPE_GROUP, 64, 10};
  L.b[i++] = {2048, SHAPE_BLOCK, 512, 4096};
  L.b[i++] = {4096, SHAPE_BLOCK, 1024, 8192};
  // the largest table the workgroup LDS holds, filled to at most 3/4, and at
  // most 16 entries per thread (owner-based kernel)
  const int64_t hmax = (int64_t)(lds_bytes - 2048) / slot_bytes / 16 * 16;
  int64_t up = hmax * 3 / 4;
  if (up > 16 * 1024) up = 16 * 1024;
  // below it, rows up to 8192 entries in tables of half the LDS (two
  // workgroups of 512 threads per CU) when such a table holds them at <= 7/8
  const int64_t half = (int64_t)(lds_bytes / 2 - 1024) / slot_bytes;
  if (up > 8192 && half * 7 >= 8192 * 8) L.b[i++] = {8192, SHAPE_BLOCK, 512, (int)half};
  if (up > 4096) L.b[i++] = {up, SHAPE_BLOCK, 1024, (int)hmax};
  L.b[i++] = {INT64_MAX, SHAPE_GLOBAL, 0, 0};
  L.nb = i;
  return L;
}

template <class M, class AT>
static int move_bin_t(const BinSpec& bs, const MoveArgs& A, const vid* perm, int64_t nb,
                      hipStream_t st) {
  typedef typename M::S ST;
  if (nb <= 0 || bs.shape == SHAPE_NONE) return 0;
  if (bs.shape == SHAPE_LANE) {
    const unsigned grid = (unsigned)((nb + 255) / 256);
    if (bs.upper <= 4) { k_move_lane<M, AT, 4><<<grid, 256, 0, st>>>(A, perm, nb); return 0; }
    if (bs.upper <= 8) { k_move_lane<M, AT, 8><<<grid, 256, 0, st>>>(A, perm, nb); return 0; }
    if (bs.upper <= 16) { k_move_lane<M, AT, 16><<<grid, 256, 0, st>>>(A, perm, nb); return 0; }
    return -1;
  }
  if (bs.shape == SHAPE_GROUP) {
    const int gpb = 256 / bs.s;
    const unsigned grid = (unsigned)((nb + gpb - 1) / gpb);
#define G_(S_, LG_)                                                                     \
  if (bs.s == S_ && bs.size == LG_) {                                                   \
    k_move_group<M, AT, S_, LG_><<<grid, 256, 0, st>>>(A, perm, nb);                    \
    return 0;                                                                           \
  }
// owner-based instances (lanes S, table 2^LG, entries per lane IT); the
// smallest IT that covers the bin is taken, longer rows use the scan kernel
#define O_(S_, LG_, IT_)                                                                \
  if (bs.s == S_ && bs.size == LG_ && need <= IT_) {                                    \
    k_move_group_own<M, AT, S_, LG_, IT_><<<grid, 256, 0, st>>>(A, perm, nb);           \
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
    k_move_block_own<M, AT, BS_, IT_><<<(unsigned)nb, BS_, lds_own, st>>>(A, perm, nb, H);  \
    return 0;                                                                               \
  }
    B_(256, 1) B_(256, 2) B_(256, 4) B_(256, 8) B_(256, 16)
    B_(512, 1) B_(512, 2) B_(512, 4) B_(512, 8) B_(512, 16)
    B_(1024, 1) B_(1024, 2) B_(1024, 4) B_(1024, 8) B_(1024, 16)
#undef B_
    const size_t lds = tab_bytes(H, sizeof(ST), sizeof(AT));
    if (bs.s == 256) {
      k_move_block<M, AT, 256><<<(unsigned)nb, 256, lds, st>>>(A, perm, nb, H);
      return 0;
    }
    if (bs.s == 512) {
      k_move_block<M, AT, 512><<<(unsigned)nb, 512, lds, st>>>(A, perm, nb, H);
      return 0;
    }
    if (bs.s == 1024) {
      k_move_block<M, AT, 1024><<<(unsigned)nb, 1024, lds, st>>>(A, perm, nb, H);
      return 0;
    }
  }
  return -1;
}

int launch_move_bin(int mode, int rec32, const BinSpec& bs, const MoveArgs& A, const vid* perm,
                    int64_t nb, hipStream_t st) {
  switch (mode) {
    case MODE_UNIT:
      return rec32 ? move_bin_t<ModeUnit, int32_t>(bs, A, perm, nb, st)
                   : move_bin_t<ModeUnit, int64_t>(bs, A, perm, nb, st);
    case MODE_INT: return move_bin_t<ModeInt, int64_t>(bs, A, perm, nb, st);
    case MODE_I32:
      return rec32 ? move_bin_t<ModeI32, int32_t>(bs, A, perm, nb, st)
                   : move_bin_t<ModeI32, int64_t>(bs, A, perm, nb, st);
    case MODE_F64: return move_bin_t<ModeF64, double>(bs, A, perm, nb, st);
    case MODE_F32: return move_bin_t<ModeF32, double>(bs, A, perm, nb, st);
  }
  return -1;
}

// Opt the LDS-table kernels into the device's largest workgroup LDS size.
void move_kernels_allow_lds(int bytes) {
#define A_(M_, AT_, BS_) \
  (void)hipFuncSetAttribute((const void*)k_move_block<M_, AT_, BS_>, \
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
#define O_(M_, AT_, BS_, IT_) \
  (void)hipFuncSetAttribute((const void*)k_move_block_own<M_, AT_, BS_, IT_>, \
                            hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
#define P_(M_, AT_, BS_) O_(M_, AT_, BS_, 1) O_(M_, AT_, BS_, 2) O_(M_, AT_, BS_, 4) \
  O_(M_, AT_, BS_, 8) O_(M_, AT_, BS_, 16)
#define B_(M_, AT_) A_(M_, AT_, 256) A_(M_, AT_, 512) A_(M_, AT_, 1024) \
  P_(M_, AT_, 256) P_(M_, AT_, 512) P_(M_, AT_, 1024)
  B_(ModeUnit, int32_t) B_(ModeUnit, int64_t) B_(ModeInt, int64_t) B_(ModeF64, double)
  B_(ModeF32, double) B_(ModeI32, int32_t) B_(ModeI32, int64_t)
#undef B_
#undef P_
#undef O_
#undef A_
}

// ---------------------------------------------------------------------------
// Row reductions: S lanes per vertex (S <= 64) or a workgroup (S = 256).
//   OP_WEIGHT: k[v] = sum_row w (self-loops included, louvain.cu:30-74),
//              vin[v] = self-loop weight
//   OP_INTRA : vin[v] = sum of w over entries with c[u] == c[v] (self-loop
//              included, louvain.cu:187-206)
template <class M, int S, int OP>
__global__ __launch_bounds__(256) void k_rows(RowArgs A, const vid* __restrict__ perm, int64_t nb) {
  typedef typename M::K KT;
  constexpr int GPB = (S >= 256) ? 1 : 256 / S;
  __shared__ KT red[4];
  const int grp = (S >= 256) ? 0 : threadIdx.x / S;
  const int lane = (S >= 256) ? threadIdx.x : (threadIdx.x & (S - 1));
  const int64_t gi = (int64_t)blockIdx.x * GPB + grp;
  bool ok = gi < nb;
  vid v = 0;
  if (ok) {
    v = perm[gi];
    if (OP == OP_INTRA && A.gs) {
      const int g = A.gid ? A.gid[v] : 0;
      ok = A.gs[g].active != 0;
    }
  }
  if (S >= 256 && !ok) return;  // uniform per workgroup
  KT s0 = 0, s1 = 0;
  if (ok) {
    const eid e0 = A.off[v], e1 = A.off[v + 1];
    const vid cv = (OP == OP_INTRA) ? A.c[v] : 0;
    for (eid e = e0 + lane; e < e1; e += (S >= 256 ? 256 : S)) {
      const vid u = colload(A.col, e);
      const KT wv = (KT)wload<M>(A.w, e);
      if (OP == OP_WEIGHT) {
        s0 += wv;
        if (u == v) s1 += wv;
      } else {
        if (A.c[u] == cv) s1 += wv;
      }
    }
  }
  if (S >= 256) {
    s0 = block_sum<256>(s0, red);
    s1 = block_sum<256>(s1, red);
  } else {
    if (!ok) return;
    s0 = group_sum<S>(s0);
    s1 = group_sum<S>(s1);
  }
  if (lane == 0) {
    if (OP == OP_WEIGHT) ((KT*)A.k)[v] = s0;
    ((KT*)A.vin)[v] = s1;
  }
}

template <class M, int OP>
static void rows_bin_t(const BinSpec& bs, const RowArgs& A, const vid* perm, int64_t nb,
                       hipStream_t st) {
  if (nb <= 0 || bs.shape == SHAPE_NONE) return;
  if (bs.upper <= 4) {
    k_rows<M, 4, OP><<<(unsigned)((nb + 63) / 64), 256, 0, st>>>(A, perm, nb);
  } else if (bs.upper <= 16) {
    k_rows<M, 16, OP><<<(unsigned)((nb + 15) / 16), 256, 0, st>>>(A, perm, nb);
  } else if (bs.upper <= 512) {
    k_rows<M, 64, OP><<<(unsigned)((nb + 3) / 4), 256, 0, st>>>(A, perm, nb);
  } else {
    k_rows<M, 256, OP><<<(unsigned)nb, 256, 0, st>>>(A, perm, nb);
  }
}

void launch_rows_bin(int mode, int op, const BinSpec& bs, const RowArgs& A, const vid* perm,
                     int64_t nb, hipStream_t st) {
#define R_(M_)                                                  \
  if (op == OP_WEIGHT) rows_bin_t<M_, OP_WEIGHT>(bs, A, perm, nb, st); \
  else rows_bin_t<M_, OP_INTRA>(bs, A, perm, nb, st);
  switch (mode) {
    case MODE_UNIT: R_(ModeUnit) break;
    case MODE_INT: R_(ModeInt) break;
    case MODE_I32: R_(ModeI32) break;
    case MODE_F64: R_(ModeF64) break;
    case MODE_F32: R_(ModeF32) break;
  }
#undef R_
}

// ---------------------------------------------------------------------------
// Degree binning (BIN.cu:33-121 histogram, :298-342 permutation). Within a
// bin the order is whatever the atomics give; no result depends on it.
struct BinBounds {
  int64_t up[MGL_MAX_BINS];
  int nb;
};
__device__ __forceinline__ int bin_of(int64_t d, const BinBounds& B) {
  int b = 0;
  while (b < B.nb - 1 && d > B.up[b]) ++b;
  return b;
}
static BinBounds bounds_of(const BinLayout& L) {
  BinBounds B;
  B.nb = L.nb;
  for (int i = 0; i < L.nb; ++i) B.up[i] = L.b[i].upper;
  return B;
}

__global__ __launch_bounds__(256) void k_bin_count(const eid* __restrict__ off, int64_t n,
                                                   BinBounds B,
                                                   unsigned long long* __restrict__ counts) {
  __shared__ unsigned int h[MGL_MAX_BINS];
  if (threadIdx.x < MGL_MAX_BINS) h[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < n; v += (int64_t)gridDim.x * 256)
    atomicAdd(&h[bin_of(off[v + 1] - off[v], B)], 1u);
  __syncthreads();
  if (threadIdx.x < B.nb && h[threadIdx.x])
    atomicAdd(&counts[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_bin_scatter(const eid* __restrict__ off, int64_t n,
                                                     BinBounds B,
                                                     unsigned long long* __restrict__ cursor,
                                                     vid* __restrict__ perm) {
  __shared__ unsigned int h[MGL_MAX_BINS];
  __shared__ unsigned long long base[MGL_MAX_BINS];
  for (int64_t v0 = (int64_t)blockIdx.x * 256; v0 < n; v0 += (int64_t)gridDim.x * 256) {
    if (threadIdx.x < MGL_MAX_BINS) h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t v = v0 + threadIdx.x;
    int b = -1;
    unsigned int r = 0;
    if (v < n) {
      b = bin_of(off[v + 1] - off[v], B);
      r = atomicAdd(&h[b], 1u);
    }
    __syncthreads();
    if (threadIdx.x < B.nb && h[threadIdx.x])
      base[threadIdx.x] = atomicAdd(&cursor[threadIdx.x], (unsigned long long)h[threadIdx.x]);
    __syncthreads();
    if (b >= 0) perm[base[b] + r] = (vid)v;
    __syncthreads();
  }
}


void launch_bin_count(const eid* off, int64_t n, const BinLayout& L, unsigned long long* counts,
                      hipStream_t st) {
  if (n <= 0) return;
  k_bin_count<<<grid_for(n, 256, 4096), 256, 0, st>>>(off, n, bounds_of(L), counts);
}
void launch_bin_scatter(const eid* off, int64_t n, const BinLayout& L, unsigned long long* cursor,
                        vid* perm, hipStream_t st) {
  if (n <= 0) return;
  k_bin_scatter<<<grid_for(n, 256, 4096), 256, 0, st>>>(off, n, bounds_of(L), cursor, perm);
}

// (first entry, end) of the rows of perm[0..nb)
__global__ void k_row_ranges(const eid* off, const vid* perm, int64_t nb, int64_t* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const vid v = perm[i];
  out[2 * i] = off[v];
  out[2 * i + 1] = off[v + 1];
}
void launch_row_ranges(const eid* off, const vid* perm, int64_t nb, int64_t* out,
                       hipStream_t st) {
  if (nb <= 0) return;
  k_row_ranges<<<grid_for(nb, 256, 1u << 30), 256, 0, st>>>(off, perm, nb, out);
}

// ---------------------------------------------------------------------------
// Community-weight update: a[c] -= k, a[c_new] += k for every moved vertex of
// an active graph (louvain_baseline.cu:1339-1361 form; louvain.cu:1328-1336
// with 1377-1381 routes the same integers through a zeroed delta buffer).
template <class KT>
__global__ __launch_bounds__(256) void k_update(int64_t n, const vid* __restrict__ c,
                                                const vid* __restrict__ cn,
                                                const KT* __restrict__ k, KT* a,
                                                const int32_t* __restrict__ gid, GState* gs,
                                                const vid* __restrict__ list) {
  __shared__ unsigned int moved;
  if (threadIdx.x == 0) moved = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const int64_t v = list ? (int64_t)list[i] : i;
    const vid x = c[v], y = cn[v];
    if (x != y) {
      const int g = gid ? gid[v] : 0;
      if (gs[g].active) {
        const KT kv = k[v];
        atom_add(&a[x], -kv);
        atom_add(&a[y], kv);
        atomicAdd(&moved, 1u);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && moved) atomicAdd((unsigned long long*)&gs[0].moved, (unsigned long long)moved);
}

// Integer form: the moves of a workgroup are first summed per community in an
// LDS table (many vertices of a pass join the same few communities, and their
// global atomics would serialise on one address); each touched community then
// takes one global atomic per workgroup. Integer sums make the order
// irrelevant. A community whose probe sequence meets UPD_PROBES occupied slots
// goes straight to the global atomic.
#define UPD_SLOTS 1021
#define UPD_PROBES 8
__device__ __forceinline__ void upd_add(vid* tk, int64_t* tv, int64_t* a, vid key, int64_t x) {
  uint32_t h = any_slot(key, UPD_SLOTS);
  const uint32_t st = probe_step(key, UPD_SLOTS, false, 0);
  for (int t = 0; t < UPD_PROBES; ++t) {
    const vid cur = atomicCAS(&tk[h], MGL_EMPTY, key);
    if (cur == MGL_EMPTY || cur == key) {
      atom_add(&tv[h], x);
      return;
    }
    h = probe_next(h, st, UPD_SLOTS);
  }
  atom_add(&a[key], x);
}

__global__ __launch_bounds__(256) void k_update_int(int64_t n, const vid* __restrict__ c,
                                                    const vid* __restrict__ cn,
                                                    const int64_t* __restrict__ k, int64_t* a,
                                                    const int32_t* __restrict__ gid, GState* gs,
                                                    const vid* __restrict__ list) {
  __shared__ vid tk[UPD_SLOTS];
  __shared__ int64_t tv[UPD_SLOTS];
  __shared__ unsigned int moved;
  for (int s = threadIdx.x; s < UPD_SLOTS; s += 256) {
    tk[s] = MGL_EMPTY;
    tv[s] = 0;
  }
  if (threadIdx.x == 0) moved = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const int64_t v = list ? (int64_t)list[i] : i;
    const vid x = c[v], y = cn[v];
    if (x != y) {
      const int g = gid ? gid[v] : 0;
      if (gs[g].active) {
        const int64_t kv = k[v];
        upd_add(tk, tv, a, x, -kv);
        upd_add(tk, tv, a, y, kv);
        atomicAdd(&moved, 1u);
      }
    }
  }
  __syncthreads();
  for (int s = threadIdx.x; s < UPD_SLOTS; s += 256)
    if (tk[s] != MGL_EMPTY && tv[s] != 0) atom_add(&a[tk[s]], tv[s]);
  if (threadIdx.x == 0 && moved) atomicAdd((unsigned long long*)&gs[0].moved, (unsigned long long)moved);
}

void launch_update(int mode, int64_t n, const vid* c, const vid* cn, const void* k, void* a,
                   const int32_t* gid, GState* gs, hipStream_t st, const vid* list) {
  if (n <= 0) return;
  const unsigned grid = grid_for(n, 256, 8192);
  if (mode_is_int(mode))
    k_update_int<<<grid, 256, 0, st>>>(n, c, cn, (const int64_t*)k, (int64_t*)a, gid, gs, list);
  else
    k_update<double><<<grid, 256, 0, st>>>(n, c, cn, (const double*)k, (double*)a, gid, gs,
                                           list);
}

// ---------------------------------------------------------------------------
// Per-graph reduction of sum_in (from vin) and sum_c a_c^2 (louvain.cu:146-165,
// 187-206). Graphs occupy contiguous vertex ranges, so each wavefront walks a
// contiguous range, keeps lane-private partials for the current graph and
// flushes them (exact 64/128-bit atomics, or FP64 atomics) when the graph id
// changes and at the end.
template <bool EXACT>
struct RedAcc;
template <>
struct RedAcc<true> {
  int64_t sin = 0;
  u128 sa2 = 0;
  __device__ void add(int64_t vin, int64_t a) {
    sin += vin;
    const u128 x = (u128)(uint64_t)(a < 0 ? -a : a);
    sa2 += x * x;
  }
  __device__ void flush(GState* s) {
    int64_t t = sin;
    uint64_t lo = (uint64_t)sa2, hi = (uint64_t)(sa2 >> 64);
    t = group_sum<64>(t);
    // 128-bit wave sum: sum the halves separately with carries
    uint64_t lo_sum = lo, hi_sum = hi;
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
      uint64_t olo = (uint64_t)shfl_xor_t((int64_t)lo_sum, m, 64);
      uint64_t ohi = (uint64_t)shfl_xor_t((int64_t)hi_sum, m, 64);
      uint64_t nlo = lo_sum + olo;
      hi_sum = hi_sum + ohi + (nlo < lo_sum ? 1ull : 0ull);
      lo_sum = nlo;
    }
    if ((threadIdx.x & 63) == 0) {
      if (t) atomicAdd((unsigned long long*)&s->sin_i, (unsigned long long)t);
      atomic_add_u128(&s->sa2_lo, &s->sa2_hi, ((u128)hi_sum << 64) | lo_sum);
    }
    sin = 0;
    sa2 = 0;
  }
};
template <>
struct RedAcc<false> {
  double sin = 0, sa2 = 0;
  __device__ void add(double vin, double a) { sin += vin; sa2 += a * a; }
  __device__ void flush(GState* s) {
    double t = group_sum<64>(sin), u = group_sum<64>(sa2);
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(&s->sin_d, t);
      atomicAdd(&s->sa2_d, u);
    }
    sin = 0;
    sa2 = 0;
  }
};

template <class KT, bool EXACT>
__global__ __launch_bounds__(256) void k_reduce(int64_t n, const KT* __restrict__ vin,
                                                const KT* __restrict__ a,
                                                const int32_t* __restrict__ gid, GState* gs,
                                                int64_t per_wave, const vid* __restrict__ list) {
  const int64_t wave = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t beg = wave * per_wave;
  if (beg >= n) return;
  const int64_t end = beg + per_wave < n ? beg + per_wave : n;
  RedAcc<EXACT> acc;
  auto vx = [&](int64_t i) -> int64_t { return list ? (int64_t)list[i] : i; };
  int gcur = gid ? gid[vx(beg)] : 0;
  for (int64_t v0 = beg; v0 < end; v0 += 64) {
    const int64_t i = v0 + lane;
    const bool ok = i < end;
    const int64_t v = ok ? vx(i) : 0;
    int g = gcur;
    if (gid) {
      g = ok ? gid[v] : gid[vx(end - 1)];
      const int gf = __shfl(g, 0);
      const int gl = __shfl(g, 63);
      if (gf != gcur || gl != gcur) {
        // the wave straddles a graph boundary: flush, then handle lane by lane
        acc.flush(&gs[gcur]);
        if (gf == gl) {
          gcur = gf;
        } else {
          for (int j = 0; j < 64; ++j) {
            const int gj = __shfl(g, j);
            RedAcc<EXACT> one;
            if (lane == j && ok && gs[gj].active) one.add(vin ? vin[v] : (KT)0, a[v]);
            one.flush(&gs[gj]);
          }
          gcur = gl;
          continue;
        }
      }
    }
    if (ok && gs[gcur].active) acc.add(vin ? vin[v] : (KT)0, a[v]);
  }
  acc.flush(&gs[gcur]);
}

// One graph: grid-stride partials, one flush per wavefront.
template <class KT, bool EXACT>
__global__ __launch_bounds__(256) void k_reduce1(int64_t n, const KT* __restrict__ vin,
                                                 const KT* __restrict__ a, GState* gs) {
  if (!gs[0].active) return;
  RedAcc<EXACT> acc;
  for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < n; v += (int64_t)gridDim.x * 256)
    acc.add(vin ? vin[v] : (KT)0, a[v]);
  acc.flush(&gs[0]);
}

void launch_reduce(int mode, int64_t n, const void* vin, const void* a, const int32_t* gid,
                   GState* gs, hipStream_t st, const vid* list) {
  if (n <= 0) return;
  if (!gid && !list) {
    const unsigned grid = grid_for(n, 256 * 8, 1024);
    if (mode_is_int(mode))
      k_reduce1<int64_t, true><<<grid, 256, 0, st>>>(n, (const int64_t*)vin, (const int64_t*)a, gs);
    else
      k_reduce1<double, false><<<grid, 256, 0, st>>>(n, (const double*)vin, (const double*)a, gs);
    return;
  }
  const int64_t per_wave = 256;
  const int64_t waves = (n + per_wave - 1) / per_wave;
  const unsigned grid = (unsigned)((waves + 3) / 4);
  if (mode_is_int(mode))
    k_reduce<int64_t, true><<<grid, 256, 0, st>>>(n, (const int64_t*)vin, (const int64_t*)a,
                                                  gid, gs, per_wave, list);
  else
    k_reduce<double, false><<<grid, 256, 0, st>>>(n, (const double*)vin, (const double*)a, gid,
                                                  gs, per_wave, list);
}

// ---------------------------------------------------------------------------
// Inner-loop control, one thread per graph (louvain.cu:1548-1660).
__device__ __forceinline__ double q_of(const GState& s, bool exact, bool prev) {
  const double M = s.mass;
  if (M == 0.0) return 0.0;
  double sin, sa2;
  if (exact) {
    sin = (double)s.sin_i;
    sa2 = prev ? u128_to_double(s.sa2p_hi, s.sa2p_lo) : u128_to_double(s.sa2_hi, s.sa2_lo);
  } else {
    sin = s.sin_d;
    sa2 = prev ? s.sa2p_d : s.sa2_d;
  }
  return sin / M - sa2 / (M * M);  // louvain.cu:234
}

__device__ __forceinline__ void reset_acc(GState& s) {
  s.sin_i = 0; s.sin_d = 0; s.sa2_lo = 0; s.sa2_hi = 0; s.sa2_d = 0;
}

// Level start, step 1: graphs listed in run[] become active, accumulators cleared.
__global__ void k_level_begin(int ng, GState* gs, const int32_t* run) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  GState& s = gs[g];
  s.active = run[g] ? 1 : 0;
  s.iters = 0;
  s.fix = 0;
  reset_acc(s);
  s.sa2p_lo = 0; s.sa2p_hi = 0; s.sa2p_d = 0;
}
void launch_level_begin(int ng, GState* gs, const int32_t* run, hipStream_t st) {
  k_level_begin<<<(ng + 255) / 256, 256, 0, st>>>(ng, gs, run);
}

// Level start, step 2: q_cur = Q(singletons) from the accumulators
// (louvain.cu:1509-1533).
__global__ void k_level_init(int ng, GState* gs, int exact) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  GState& s = gs[g];
  if (s.active) s.q_cur = q_of(s, exact != 0, false);
  reset_acc(s);
}
void launch_level_init(int ng, GState* gs, int exact, hipStream_t st) {
  k_level_init<<<(ng + 255) / 256, 256, 0, st>>>(ng, gs, exact);
}

// pre = 1: after the first move of a level (fused mode) only stores sum a^2 of
// the first proposal; no decision is taken.
__device__ void decide_one(GState& s, int g, double thr, int max_iter, int fused, int exact,
                           double* traj, int traj_max, int pre);
__global__ void k_decide(int ng, GState* gs, double thr, int max_iter, int fused, int exact,
                         double* traj, int traj_max, int pre, int* nactive) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  GState& s = gs[g];
  if (!s.active) { s.fix = 0; reset_acc(s); return; }
  decide_one(s, g, thr, max_iter, fused, exact, traj, traj_max, pre);
  if (nactive && s.active) atomicAdd(nactive, 1);
}
__device__ void decide_one(GState& s, int g, double thr, int max_iter, int fused, int exact,
                           double* traj, int traj_max, int pre) {
  if (pre) {
    s.sa2p_lo = s.sa2_lo; s.sa2p_hi = s.sa2_hi; s.sa2p_d = s.sa2_d;
    s.fix = 0;
    reset_acc(s);
    return;
  }
  const double Q = q_of(s, exact != 0, fused != 0);
  const int it = s.iters++;
  if ((Q - s.q_cur) > thr) {
    s.q_cur = Q;
    if (traj && it < traj_max) traj[(int64_t)g * traj_max + it] = Q;
    const bool stop = (max_iter > 0 && s.iters >= max_iter);
    if (fused) s.fix = stop ? 2 : 0;
    else s.fix = 3;
    if (stop) s.active = 0;
  } else {
    if (traj && it < traj_max) traj[(int64_t)g * traj_max + it] = s.q_cur;
    s.active = 0;
    s.fix = fused ? 1 : 0;
  }
  if (fused) { s.sa2p_lo = s.sa2_lo; s.sa2p_hi = s.sa2_hi; s.sa2p_d = s.sa2_d; }
  reset_acc(s);
}

void launch_decide(int ng, GState* gs, double thr, int max_iter, int fused, int exact,
                   double* traj, int traj_max, int pre, hipStream_t st, int* nactive) {
  k_decide<<<(ng + 255) / 256, 256, 0, st>>>(ng, gs, thr, max_iter, fused, exact, traj, traj_max,
                                             pre, nactive);
}

// Buffer fix-up after a decision:
//   fix 1 (fused, rejected)        : c, p, n := c
//   fix 2 (fused, accepted, stop)  : c, p, n := p
//   fix 3 (unfused, accepted)      : c := p
__global__ __launch_bounds__(256) void k_fixup(int64_t n, vid* c, vid* p, vid* nx,
                                               const int32_t* __restrict__ gid,
                                               const GState* __restrict__ gs,
                                               const vid* __restrict__ list) {
  if (!gid && gs[0].fix == 0) return;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const int64_t v = list ? (int64_t)list[i] : i;
    const int g = gid ? gid[v] : 0;
    const int f = gs[g].fix;
    if (f == 0) continue;
    if (f == 3) { c[v] = p[v]; continue; }
    const vid x = (f == 1) ? c[v] : p[v];
    c[v] = x;
    p[v] = x;
    if (nx) nx[v] = x;
  }
}
void launch_fixup(int64_t n, vid* c, vid* p, vid* nx, const int32_t* gid, const GState* gs,
                  hipStream_t st, const vid* list) {
  if (n <= 0) return;
  k_fixup<<<grid_for(n, 256, 8192), 256, 0, st>>>(n, c, p, nx, gid, gs, list);
}

// Keeps, per bin, the vertices of graphs still active: perm_out[off[b] + j].
__global__ __launch_bounds__(256) void k_filter_bin(const vid* __restrict__ perm, int64_t nb,
                                                    const int32_t* __restrict__ gid,
                                                    const GState* __restrict__ gs,
                                                    vid* __restrict__ out,
                                                    unsigned long long* __restrict__ cnt) {
  __shared__ unsigned int lc;
  __shared__ unsigned long long base;
  for (int64_t i0 = (int64_t)blockIdx.x * 256; i0 < nb; i0 += (int64_t)gridDim.x * 256) {
    if (threadIdx.x == 0) lc = 0;
    __syncthreads();
    const int64_t i = i0 + threadIdx.x;
    vid v = 0;
    bool keep = false;
    unsigned int r = 0;
    if (i < nb) {
      v = perm[i];
      keep = gs[gid ? gid[v] : 0].active != 0;
      if (keep) r = atomicAdd(&lc, 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0 && lc) base = atomicAdd(cnt, (unsigned long long)lc);
    __syncthreads();
    if (keep) out[base + r] = v;
    __syncthreads();
  }
}
void launch_filter_bin(const vid* perm, int64_t nb, const int32_t* gid, const GState* gs,
                       vid* out, unsigned long long* cnt, hipStream_t st) {
  if (nb <= 0) return;
  k_filter_bin<<<grid_for(nb, 256, 4096), 256, 0, st>>>(perm, nb, gid, gs, out, cnt);
}

// flag[v] = 1 for non-isolated vertices of active graphs
__global__ void k_flag_active(int64_t n, const eid* __restrict__ off,
                              const int32_t* __restrict__ gid, const GState* __restrict__ gs,
                              int32_t* __restrict__ flag) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    flag[v] = (off[v + 1] > off[v] && gs[gid ? gid[v] : 0].active) ? 1 : 0;
}
__global__ void k_compact_ids(int64_t n, const int32_t* __restrict__ flag,
                              const int64_t* __restrict__ pos, vid* __restrict__ out) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * blockDim.x)
    if (flag[v]) out[pos[v]] = (vid)v;
}
void launch_active_list(int64_t n, const eid* off, const int32_t* gid, const GState* gs,
                        int32_t* flag, int64_t* pos, vid* out, void* scan_tmp, hipStream_t st) {
  if (n <= 0) return;
  k_flag_active<<<grid_for(n, 256, 8192), 256, 0, st>>>(n, off, gid, gs, flag);
  scan_exclusive_i32(flag, pos, n, scan_tmp, st);
  k_compact_ids<<<grid_for(n, 256, 8192), 256, 0, st>>>(n, flag, pos, out);
}

// ---------------------------------------------------------------------------
__global__ void k_iota(vid* x, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = (vid)i;
}
void launch_iota(vid* x, int64_t n, hipStream_t st) {
  if (n > 0) k_iota<<<grid_for(n, 256, 8192), 256, 0, st>>>(x, n);
}
template <class T>
__global__ void k_copy(const T* s, T* d, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[i];
}
void launch_copy_k_to_a(int mode, int64_t n, const void* k, void* a, hipStream_t st) {
  if (n <= 0) return;
  if (mode_is_int(mode))
    k_copy<int64_t><<<grid_for(n, 256, 8192), 256, 0, st>>>((const int64_t*)k, (int64_t*)a, n);
  else
    k_copy<double><<<grid_for(n, 256, 8192), 256, 0, st>>>((const double*)k, (double*)a, n);
}
void launch_copy_vid(const vid* s, vid* d, int64_t n, hipStream_t st) {
  if (n > 0) k_copy<vid><<<grid_for(n, 256, 8192), 256, 0, st>>>(s, d, n);
}
__global__ void k_zero(uint32_t* p, int64_t nw) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nw;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}
void launch_fill_bytes0(void* p, size_t bytes, hipStream_t st) {
  // a fill kernel (not hipMemsetAsync) so the sequence can be graph-captured
  const int64_t nw = (int64_t)((bytes + 3) / 4);
  if (nw > 0) k_zero<<<grid_for(nw, 256, 8192), 256, 0, st>>>((uint32_t*)p, nw);
}

// float input weights -> level-0 storage of the chosen mode. MODE_INT scales
// by 2^scale (exact for dyadic weights, checked on the host).
__global__ void k_cast(int to_mode, const float* w, void* out, int64_t nnz, double mul) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double x = (double)w[i];
    if (to_mode == MODE_INT) ((int64_t*)out)[i] = (int64_t)(x * mul);
    else if (to_mode == MODE_F64) ((double*)out)[i] = x;
    else ((float*)out)[i] = w[i];
  }
}
void launch_cast_weights(int to_mode, const float* w, void* out, int64_t nnz, int scale,
                         hipStream_t st) {
  if (nnz <= 0) return;
  double mul = 1.0;
  for (int i = 0; i < scale; ++i) mul *= 2.0;
  k_cast<<<grid_for(nnz, 256, 8192), 256, 0, st>>>(to_mode, w, out, nnz, mul);
}

__global__ void k_gid_fill(const int64_t* voff, int ng, int32_t* gid) {
  const int g = blockIdx.x;
  if (g >= ng) return;
  for (int64_t v = voff[g] + threadIdx.x; v < voff[g + 1]; v += blockDim.x) gid[v] = g;
}
void launch_gid_fill(const int64_t* voff, int ng, int32_t* gid, hipStream_t st) {
  if (ng > 0) k_gid_fill<<<ng, 256, 0, st>>>(voff, ng, gid);
}
