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
#include <cstdio>

static unsigned grid_for(int64_t n, int per_block, unsigned cap) {
  int64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Per-vertex record {c[u], a[c[u]]}, rebuilt before every move pass, so that
// one gather per CSR entry brings both the neighbour's community and that
// community's total. The slot creator stores the total in the table, so the
// gain loop reads no global memory. AT = int32 in the unit mode while
// 2m < 2^31 (an 8-byte record), else the mode's total type (16 bytes).
// Records of a large level are stored in degree-bin order (rpos): the
// frequently gathered records of high-degree vertices then share cache lines.
template <class KT, class AT>
__global__ __launch_bounds__(256) void k_build_rec(int64_t first, int64_t n,
                                                   const vid* __restrict__ c,
                                                   const KT* __restrict__ a,
                                                   Rec<AT>* __restrict__ rec,
                                                   const vid* __restrict__ list,
                                                   const vid* __restrict__ rpos,
                                                   const vid* __restrict__ rinv) {
  for (int64_t i = first + (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
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

// Degree-bin record layout built in vertex order: c[v] and a[c[v]] are read
// in order (a community id is mostly its member's own id or near it) and each
// record goes to rpos[v]. Consecutive vertices of one bin have consecutive
// positions, so the stores come in short runs, while the position-order loop
// above reads c and a once per bin (about seven sweeps of both arrays on
// R-MAT 24). Records at positions below `first` (isolated vertices) are not
// built. All loads of a thread are issued before its stores.
#define REC_IPT 8
template <class KT, class AT>
__global__ __launch_bounds__(256) void k_build_rec_vo(int64_t first, int64_t n,
                                                      const vid* __restrict__ c,
                                                      const KT* __restrict__ a,
                                                      Rec<AT>* __restrict__ rec,
                                                      const vid* __restrict__ rpos) {
  const int64_t v0 = (int64_t)blockIdx.x * 256 * REC_IPT + threadIdx.x;
  vid x[REC_IPT];
  int64_t at[REC_IPT];
  KT av[REC_IPT];
#pragma unroll
  for (int j = 0; j < REC_IPT; ++j) {
    const int64_t v = v0 + (int64_t)j * 256;
    at[j] = v < n ? (int64_t)rpos[v] : -1;
    x[j] = v < n ? c[v] : 0;
  }
#pragma unroll
  for (int j = 0; j < REC_IPT; ++j) av[j] = at[j] >= first ? a[x[j]] : (KT)0;
#pragma unroll
  for (int j = 0; j < REC_IPT; ++j) {
    if (at[j] < first) continue;  // beyond n (at = -1) or an isolated vertex
    Rec<AT> r;
    r.c = x[j];
    r.a = (AT)av[j];
    rec[at[j]] = r;
  }
}

template <class KT, class AT>
static void build_rec_t(int64_t n, const vid* c, const void* a, void* rec, hipStream_t st,
                        const vid* list, const vid* rpos, const vid* rinv, int64_t first) {
  if (!list && rpos && rinv) {
    const unsigned gv = (unsigned)((n + 256 * REC_IPT - 1) / (256 * REC_IPT));
    k_build_rec_vo<KT, AT><<<gv, 256, 0, st>>>(first, n, c, (const KT*)a, (Rec<AT>*)rec, rpos);
    return;
  }
  const unsigned grid = grid_for(n - first, 256, 8192);
  k_build_rec<KT, AT><<<grid, 256, 0, st>>>(first, n, c, (const KT*)a, (Rec<AT>*)rec, list, rpos,
                                            rinv);
}

void launch_build_rec(int mode, int rec32, int64_t n, const vid* c, const void* a, void* rec,
                      hipStream_t st, const vid* list, const vid* rpos, const vid* rinv,
                      int64_t first) {
  if (n - first <= 0) return;
  if ((mode == MODE_UNIT || mode == MODE_I32) && rec32)
    build_rec_t<int64_t, int32_t>(n, c, a, rec, st, list, rpos, rinv, first);
  else if (mode_is_int(mode))
    build_rec_t<int64_t, int64_t>(n, c, a, rec, st, list, rpos, rinv, first);
  else
    build_rec_t<double, double>(n, c, a, rec, st, list, rpos, rinv, first);
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

// Rows in record-position order: deg[p] = degree of the vertex at position p
// (rinv[p]), deg[n] = 0; its exclusive scan is the offsets of the placed rows.
__global__ __launch_bounds__(256) void k_pos_degrees(const eid* __restrict__ off,
                                                     const vid* __restrict__ rinv, int64_t n,
                                                     int64_t* __restrict__ deg) {
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p <= n;
       p += (int64_t)gridDim.x * 256) {
    if (p == n) {
      deg[p] = 0;
    } else {
      const vid v = rinv[p];
      deg[p] = off[v + 1] - off[v];
    }
  }
}

// Places the rows of positions p0 .. p0+np-1: the row of v = rinv[p] goes to
// offr[p] .. offr[p+1], its columns as record positions (rpos), its weights
// (wb bytes each, 0 for unit weights) unchanged. S lanes per row, a workgroup
// per row when S = 256.
template <int S>
__global__ __launch_bounds__(256) void k_place_rows(const eid* __restrict__ off,
                                                    const vid* __restrict__ col,
                                                    const void* __restrict__ w, int wb,
                                                    const vid* __restrict__ rinv,
                                                    const vid* __restrict__ rpos,
                                                    const eid* __restrict__ offr, int64_t p0,
                                                    int64_t np, vid* __restrict__ colr,
                                                    void* __restrict__ wr) {
  constexpr int GPB = S >= 256 ? 1 : 256 / S;
  constexpr int STEP = S >= 256 ? 256 : S;
  const int grp = S >= 256 ? 0 : (int)threadIdx.x / S;
  const int lane = S >= 256 ? (int)threadIdx.x : (int)(threadIdx.x & (S - 1));
  for (int64_t i = (int64_t)blockIdx.x * GPB + grp; i < np; i += (int64_t)gridDim.x * GPB) {
    const int64_t p = p0 + i;
    const vid v = rinv[p];
    const eid s0 = off[v], s1 = off[v + 1], d0 = offr[p];
    for (eid e = s0 + lane; e < s1; e += STEP) {
      const eid q = d0 + (e - s0);
      colr[q] = rpos[colload(col, e)];
      if (wb == 4) ((uint32_t*)wr)[q] = ((const uint32_t*)w)[e];
      else if (wb == 8) ((uint64_t*)wr)[q] = ((const uint64_t*)w)[e];
    }
  }
}

void launch_record_layout(const vid* perm, int64_t n, vid* rpos, vid* rinv, hipStream_t st) {
  if (n <= 0) return;
  k_record_layout<<<grid_for(n, 256, 16384), 256, 0, st>>>(perm, n, rpos, rinv);
}

void launch_pos_degrees(const eid* off, const vid* rinv, int64_t n, int64_t* deg,
                        hipStream_t st) {
  k_pos_degrees<<<grid_for(n + 1, 256, 16384), 256, 0, st>>>(off, rinv, n, deg);
}

void launch_place_rows(int64_t upper, const eid* off, const vid* col, const void* w, int wb,
                       const vid* rinv, const vid* rpos, const eid* offr, int64_t p0, int64_t np,
                       vid* colr, void* wr, hipStream_t st) {
  if (np <= 0) return;
#define P_(S_)                                                                                 \
  {                                                                                            \
    constexpr int GPB = S_ >= 256 ? 1 : 256 / S_;                                              \
    k_place_rows<S_><<<grid_for(np, GPB, 65536), 256, 0, st>>>(off, col, w, wb, rinv, rpos, \
                                                               offr, p0, np, colr, wr);     \
    return;                                                                                    \
  }
  if (upper <= 8) P_(4)
  if (upper <= 64) P_(16)
  if (upper <= 2048) P_(64)
  P_(256)
#undef P_
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
  // above it one 1024-thread workgroup per CU with the whole LDS: on R-MAT 24
  // a phased kernel (two 512-thread workgroups per CU, half tables, two
  // phases) took 2.36 ms for these rows against 1.86 ms
  if (up > 4096) L.b[i++] = {up, SHAPE_BLOCK, 1024, (int)hmax};
  L.b[i++] = {INT64_MAX, SHAPE_GLOBAL, 0, 0};
  L.nb = i;
  return L;
}

// ---------------------------------------------------------------------------
// Move-kernel dispatch over the per-mode instantiation units (mgl_move_*.hip).
static int move_bin_any(int mode, int rec32, bool cnt, const BinSpec& bs, const MoveArgs& A,
                        const vid* perm, int64_t nb, hipStream_t st, char* name, int len) {
  switch (mode) {
    case MODE_UNIT: return move_bin_unit(mode, rec32, cnt, bs, A, perm, nb, st, name, len);
    case MODE_I32: return move_bin_i32(mode, rec32, cnt, bs, A, perm, nb, st, name, len);
    case MODE_INT:
    case MODE_F64: return move_bin_int(mode, rec32, cnt, bs, A, perm, nb, st, name, len);
    case MODE_F32: return move_bin_f32(mode, rec32, cnt, bs, A, perm, nb, st, name, len);
  }
  return -1;
}

int move_kernel_name(int mode, int rec32, const BinSpec& bs, char* buf, int len, bool cnt) {
  if (!buf || len <= 0) return -1;
  buf[0] = 0;
  MoveArgs A{};
  return move_bin_any(mode, rec32, cnt, bs, A, nullptr, 1, 0, buf, len);
}

int launch_move_bin(int mode, int rec32, const BinSpec& bs, const MoveArgs& A, const vid* perm,
                    int64_t nb, hipStream_t st) {
  const bool cnt = A.stats != nullptr || A.probes != nullptr;
  return move_bin_any(mode, rec32, cnt, bs, A, perm, nb, st, nullptr, 0);
}

void move_kernels_allow_lds(int bytes) {
  move_allow_lds_unit(bytes);
  move_allow_lds_i32(bytes);
  move_allow_lds_int(bytes);
  move_allow_lds_f32(bytes);
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

__global__ void k_gather_i32(const int32_t* src, const vid* idx, int64_t n, int32_t* out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = src[idx[i]];
}
void launch_gather_i32(const int32_t* src, const vid* idx, int64_t n, int32_t* out,
                       hipStream_t st) {
  if (n <= 0) return;
  k_gather_i32<<<grid_for(n, 256, 4096), 256, 0, st>>>(src, idx, n, out);
}

__global__ void k_gather_u64(const uint64_t* src, const vid* idx, int64_t n, uint64_t* out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = src[idx[i]];
}
void launch_gather_u64(const void* src, const vid* idx, int64_t n, void* out, hipStream_t st) {
  if (n <= 0) return;
  k_gather_u64<<<grid_for(n, 256, 16384), 256, 0, st>>>((const uint64_t*)src, idx, n,
                                                         (uint64_t*)out);
}

// ---------------------------------------------------------------------------
// Community-weight update: a[c] -= k, a[c_new] += k for every moved vertex of
// an active graph (louvain_baseline.cu:1339-1361 form; louvain.cu:1328-1336
// with 1377-1381 routes the same integers through a zeroed delta buffer).
// A workgroup of UPD_BS threads takes UPD_IPT consecutive strides of UPD_BS
// vertices and issues all their c / c_new loads before any update. The moved
// count is summed per workgroup: one global atomic per workgroup, since
// atomics on one address serialise (about 10 ns each).
#define UPD_BS 1024
#define UPD_IPT 16

template <class KT>
__global__ __launch_bounds__(UPD_BS) void k_update(int64_t n, const vid* __restrict__ c,
                                                   const vid* __restrict__ cn,
                                                   const KT* __restrict__ k, KT* a,
                                                   const int32_t* __restrict__ gid, GState* gs,
                                                   const vid* __restrict__ list) {
  __shared__ unsigned int ws[UPD_BS / 64];
  const int64_t i0 = (int64_t)blockIdx.x * UPD_BS * UPD_IPT + threadIdx.x;
  int64_t v[UPD_IPT];
  vid x[UPD_IPT], y[UPD_IPT];
#pragma unroll
  for (int j = 0; j < UPD_IPT; ++j) {
    const int64_t i = i0 + (int64_t)j * UPD_BS;
    v[j] = i < n ? (list ? (int64_t)list[i] : i) : -1;
  }
#pragma unroll
  for (int j = 0; j < UPD_IPT; ++j) {
    x[j] = v[j] >= 0 ? c[v[j]] : 0;
    y[j] = v[j] >= 0 ? cn[v[j]] : 0;
  }
  const bool act0 = gid ? true : gs[0].active != 0;
  unsigned int cnt = 0;
#pragma unroll
  for (int j = 0; j < UPD_IPT; ++j) {
    if (x[j] == y[j]) continue;
    if (!(gid ? gs[gid[v[j]]].active != 0 : act0)) continue;
    const KT kv = k[v[j]];
    atom_add(&a[x[j]], -kv);
    atom_add(&a[y[j]], kv);
    ++cnt;
  }
  const unsigned int tot = block_sum<UPD_BS, unsigned int>(cnt, ws);
  if (threadIdx.x == 0 && tot) atomicAdd((unsigned long long*)&gs[0].moved, (unsigned long long)tot);
}

// Integer form: the additions to the joined communities of a workgroup's
// moves are first summed per community in an LDS table (many vertices of a
// pass join the same few communities, and their global atomics would
// serialise on one address); each touched community then takes one global
// atomic per workgroup. The subtractions from the left communities, which are
// spread over many communities, go straight to global atomics (timed faster
// than through the table, also when a third of the moved vertices leave the
// same thousand communities). Integer sums make the order irrelevant. A community
// whose probe sequence meets UPD_PROBES occupied slots goes straight to the
// global atomic.
#define UPD_SLOTS 4093
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

__global__ __launch_bounds__(UPD_BS) void k_update_int(int64_t n, const vid* __restrict__ c,
                                                       const vid* __restrict__ cn,
                                                       const int64_t* __restrict__ k, int64_t* a,
                                                       const int32_t* __restrict__ gid, GState* gs,
                                                       const vid* __restrict__ list) {
  __shared__ vid tk[UPD_SLOTS];
  __shared__ int64_t tv[UPD_SLOTS];
  __shared__ unsigned int ws[UPD_BS / 64];
  for (int s = threadIdx.x; s < UPD_SLOTS; s += UPD_BS) {
    tk[s] = MGL_EMPTY;
    tv[s] = 0;
  }
  const int64_t i0 = (int64_t)blockIdx.x * UPD_BS * UPD_IPT + threadIdx.x;
  int64_t v[UPD_IPT];
  vid x[UPD_IPT], y[UPD_IPT];
#pragma unroll
  for (int j = 0; j < UPD_IPT; ++j) {
    const int64_t i = i0 + (int64_t)j * UPD_BS;
    v[j] = i < n ? (list ? (int64_t)list[i] : i) : -1;
  }
#pragma unroll
  for (int j = 0; j < UPD_IPT; ++j) {
    x[j] = v[j] >= 0 ? c[v[j]] : 0;
    y[j] = v[j] >= 0 ? cn[v[j]] : 0;
  }
  const bool act0 = gid ? true : gs[0].active != 0;  // one graph: its flag, read once
  __syncthreads();  // table cleared
  unsigned int cnt = 0;
#pragma unroll
  for (int j = 0; j < UPD_IPT; ++j) {
    if (x[j] == y[j]) continue;
    if (!(gid ? gs[gid[v[j]]].active != 0 : act0)) continue;
    const int64_t kv = k[v[j]];
    atom_add(&a[x[j]], -kv);
    upd_add(tk, tv, a, y[j], kv);
    ++cnt;
  }
  __syncthreads();
  for (int s = threadIdx.x; s < UPD_SLOTS; s += UPD_BS)
    if (tk[s] != MGL_EMPTY && tv[s] != 0) atom_add(&a[tk[s]], tv[s]);
  const unsigned int tot = block_sum<UPD_BS, unsigned int>(cnt, ws);
  if (threadIdx.x == 0 && tot) atomicAdd((unsigned long long*)&gs[0].moved, (unsigned long long)tot);
}

void launch_update(int mode, int64_t n, const vid* c, const vid* cn, const void* k, void* a,
                   const int32_t* gid, GState* gs, hipStream_t st, const vid* list) {
  if (n <= 0) return;
  const unsigned grid = (unsigned)((n + (int64_t)UPD_BS * UPD_IPT - 1) / ((int64_t)UPD_BS * UPD_IPT));
  if (mode_is_int(mode))
    k_update_int<<<grid, UPD_BS, 0, st>>>(n, c, cn, (const int64_t*)k, (int64_t*)a, gid, gs,
                                          list);
  else
    k_update<double><<<grid, UPD_BS, 0, st>>>(n, c, cn, (const double*)k, (double*)a, gid, gs,
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
  // wavefront sums (valid in every lane): sum_in, and sum a^2 as two 64-bit
  // halves added with carries
  __device__ void wave_sum(int64_t& t, uint64_t& lo_sum, uint64_t& hi_sum) const {
    t = group_sum<64>(sin);
    lo_sum = (uint64_t)sa2;
    hi_sum = (uint64_t)(sa2 >> 64);
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
      uint64_t olo = (uint64_t)shfl_xor_t((int64_t)lo_sum, m, 64);
      uint64_t ohi = (uint64_t)shfl_xor_t((int64_t)hi_sum, m, 64);
      uint64_t nlo = lo_sum + olo;
      hi_sum = hi_sum + ohi + (nlo < lo_sum ? 1ull : 0ull);
      lo_sum = nlo;
    }
  }
  __device__ void flush(GState* s) {
    int64_t t;
    uint64_t lo_sum, hi_sum;
    wave_sum(t, lo_sum, hi_sum);
    if ((threadIdx.x & 63) == 0) {
      if (t) atomicAdd((unsigned long long*)&s->sin_i, (unsigned long long)t);
      atomic_add_u128(&s->sa2_lo, &s->sa2_hi, ((u128)hi_sum << 64) | lo_sum);
    }
    sin = 0;
    sa2 = 0;
  }
  // Workgroup flush (every thread of the BS-thread workgroup calls it): the
  // wavefront sums are combined in LDS and thread 0 issues the atomics.
  template <int BS>
  __device__ void flush_block(GState* s) {
    __shared__ int64_t wt[BS / 64];
    __shared__ uint64_t wl[BS / 64], wh[BS / 64];
    int64_t t;
    uint64_t lo_sum, hi_sum;
    wave_sum(t, lo_sum, hi_sum);
    if ((threadIdx.x & 63) == 0) {
      wt[threadIdx.x >> 6] = t;
      wl[threadIdx.x >> 6] = lo_sum;
      wh[threadIdx.x >> 6] = hi_sum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t T = 0;
      u128 S = 0;
      for (int i = 0; i < BS / 64; ++i) {
        T += wt[i];
        S += ((u128)wh[i] << 64) | wl[i];
      }
      if (T) atomicAdd((unsigned long long*)&s->sin_i, (unsigned long long)T);
      atomic_add_u128(&s->sa2_lo, &s->sa2_hi, S);
    }
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
  template <int BS>
  __device__ void flush_block(GState* s) {
    __shared__ double wt[BS / 64], wu[BS / 64];
    const double t = group_sum<64>(sin), u = group_sum<64>(sa2);
    if ((threadIdx.x & 63) == 0) {
      wt[threadIdx.x >> 6] = t;
      wu[threadIdx.x >> 6] = u;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double T = 0, U = 0;
      for (int i = 0; i < BS / 64; ++i) {
        T += wt[i];
        U += wu[i];
      }
      atomicAdd(&s->sin_d, T);
      atomicAdd(&s->sa2_d, U);
    }
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

// One graph: grid-stride partials with RED_UN independent loads per thread and
// step, one flush per workgroup (atomics on one address serialise).
#define RED_BS 1024
#define RED_UN 4
template <class KT, bool EXACT>
__global__ __launch_bounds__(RED_BS) void k_reduce1(int64_t n, const KT* __restrict__ vin,
                                                    const KT* __restrict__ a, GState* gs) {
  if (!gs[0].active) return;  // uniform over the workgroup
  RedAcc<EXACT> acc;
  const int64_t stride = (int64_t)gridDim.x * RED_BS;
  for (int64_t v0 = (int64_t)blockIdx.x * RED_BS + threadIdx.x; v0 < n; v0 += stride * RED_UN) {
    KT xv[RED_UN], xa[RED_UN];
#pragma unroll
    for (int u = 0; u < RED_UN; ++u) {
      const int64_t v = v0 + u * stride;
      xv[u] = (v < n && vin) ? vin[v] : (KT)0;
      xa[u] = v < n ? a[v] : (KT)0;
    }
#pragma unroll
    for (int u = 0; u < RED_UN; ++u) acc.add(xv[u], xa[u]);
  }
  acc.template flush_block<RED_BS>(&gs[0]);
}

void launch_reduce(int mode, int64_t n, const void* vin, const void* a, const int32_t* gid,
                   GState* gs, hipStream_t st, const vid* list) {
  if (n <= 0) return;
  if (!gid && !list) {
    const unsigned grid = grid_for(n, RED_BS * RED_UN, 256);
    if (mode_is_int(mode))
      k_reduce1<int64_t, true><<<grid, RED_BS, 0, st>>>(n, (const int64_t*)vin, (const int64_t*)a,
                                                        gs);
    else
      k_reduce1<double, false><<<grid, RED_BS, 0, st>>>(n, (const double*)vin, (const double*)a,
                                                        gs);
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
  if (s.active) {
    s.q_cur = q_of(s, exact != 0, false);
    s.q0 = s.q_cur;
    s.tlast = wall_clock64();
  }
  reset_acc(s);
}
void launch_level_init(int ng, GState* gs, int exact, hipStream_t st) {
  k_level_init<<<(ng + 255) / 256, 256, 0, st>>>(ng, gs, exact);
}

// pre = 1: after the first move of a level (fused mode) only stores sum a^2 of
// the first proposal; no decision is taken.
__device__ void decide_one(GState& s, int g, double thr, int max_iter, int fused, int exact,
                           double* traj, int64_t* ttraj, int traj_max, int pre);
__global__ void k_decide(int ng, GState* gs, double thr, int max_iter, int fused, int exact,
                         double* traj, int64_t* ttraj, int traj_max, int pre, int* nactive) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  GState& s = gs[g];
  if (!s.active) { s.fix = 0; reset_acc(s); return; }
  decide_one(s, g, thr, max_iter, fused, exact, traj, ttraj, traj_max, pre);
  if (nactive && s.active) atomicAdd(nactive, 1);
}
// Besides the decision, the device clock ticks since the level start or the
// previous decision go to ttraj (the iteration's time).
__device__ void decide_one(GState& s, int g, double thr, int max_iter, int fused, int exact,
                           double* traj, int64_t* ttraj, int traj_max, int pre) {
  if (pre) {
    s.sa2p_lo = s.sa2_lo; s.sa2p_hi = s.sa2_hi; s.sa2p_d = s.sa2_d;
    s.fix = 0;
    reset_acc(s);
    return;
  }
  const double Q = q_of(s, exact != 0, fused != 0);
  const int it = s.iters++;
  const int64_t now = wall_clock64();
  if (ttraj && it < traj_max) ttraj[(int64_t)g * traj_max + it] = now - s.tlast;
  s.tlast = now;
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
                   double* traj, int traj_max, int pre, hipStream_t st, int* nactive,
                   int64_t* ttraj) {
  k_decide<<<(ng + 255) / 256, 256, 0, st>>>(ng, gs, thr, max_iter, fused, exact, traj, ttraj,
                                             traj_max, pre, nactive);
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

// Input weights (float or double) -> level-0 storage of the chosen mode.
// MODE_INT scales by 2^scale (exact for dyadic weights, checked on the host);
// MODE_F32 rounds a double to the nearest float.
template <class T>
__global__ void k_cast(int to_mode, const T* w, void* out, int64_t nnz, double mul) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double x = (double)w[i];
    if (to_mode == MODE_INT) ((int64_t*)out)[i] = (int64_t)(x * mul);
    else if (to_mode == MODE_F64) ((double*)out)[i] = x;
    else ((float*)out)[i] = (float)w[i];
  }
}
static double pow2(int scale) {
  double mul = 1.0;
  for (int i = 0; i < scale; ++i) mul *= 2.0;
  return mul;
}
void launch_cast_weights(int to_mode, const float* w, void* out, int64_t nnz, int scale,
                         hipStream_t st) {
  if (nnz <= 0) return;
  k_cast<float><<<grid_for(nnz, 256, 8192), 256, 0, st>>>(to_mode, w, out, nnz, pow2(scale));
}
void launch_cast_weights_f64(int to_mode, const double* w, void* out, int64_t nnz, int scale,
                             hipStream_t st) {
  if (nnz <= 0) return;
  k_cast<double><<<grid_for(nnz, 256, 8192), 256, 0, st>>>(to_mode, w, out, nnz, pow2(scale));
}

__global__ void k_gid_fill(const int64_t* voff, int ng, int32_t* gid) {
  const int g = blockIdx.x;
  if (g >= ng) return;
  for (int64_t v = voff[g] + threadIdx.x; v < voff[g + 1]; v += blockDim.x) gid[v] = g;
}
void launch_gid_fill(const int64_t* voff, int ng, int32_t* gid, hipStream_t st) {
  if (ng > 0) k_gid_fill<<<ng, 256, 0, st>>>(voff, ng, gid);
}
